"""Pins of the FP64 oracle against closed forms, invariants and brute force.

Each check uses an independent textbook statement (Hertz with the effective
modulus, Mindlin tangential stiffness from the shear modulus, exact constant-
acceleration kinematics, Rodrigues rotation, O(N^2) neighbour search, numpy
stable sort) rather than the oracle's own formula.
"""
import numpy as np
import pytest

import nbc
import orc
from util import cube_wall, cylinder_wall, make_params, quat_to_matrix, rodrigues, single_sphere_clumps

E, NU = 1e8, 0.3


def hertz(Rstar, delta, E=E, nu=NU):
    Estar = E / (2.0 * (1.0 - nu * nu))  # two identical elastic bodies
    return 4.0 / 3.0 * Estar * np.sqrt(Rstar) * delta ** 1.5


def mindlin_ks(Rstar, delta, E=E, nu=NU):
    G = E / (2.0 * (1.0 + nu))
    Gstar = G / (2.0 * (2.0 - nu))
    return 8.0 * Gstar * np.sqrt(Rstar * delta)


def run(bed, steps=1, wall=None, pene=None, opt=None):
    p = orc.finalize_params(bed)
    o = orc.Oracle(p, opt)
    o.upload(bed, wall, pene)
    o.step(steps)
    return o


def by_file(o, a):
    """Reorder a sorted-order per-sphere array to file order."""
    fid = o.spheres()["fid"]
    out = np.empty_like(a)
    out[fid] = a
    return out


def test_hertz_normal_force_two_spheres():
    r, d = 1e-3, 1e-5
    bed = single_sphere_clumps([[0, 0, 0], [2 * r - d, 0, 0]], r)
    o = run(bed)
    f, t = o.sphere_forces(1)
    f = by_file(o, f)
    Fn = hertz(r * r / (2 * r), d)
    assert f[0] == pytest.approx([-Fn, 0, 0], rel=1e-12, abs=1e-18)
    assert f[1] == pytest.approx([Fn, 0, 0], rel=1e-12, abs=1e-18)
    assert np.abs(t).max() == 0.0


def test_hertz_polydisperse_and_diagonal():
    r1, r2, d = 0.8e-3, 1.3e-3, 2e-5
    n = np.array([1.0, 2.0, -0.5]) / np.linalg.norm([1.0, 2.0, -0.5])
    bed = single_sphere_clumps([[0, 0, 0], (r1 + r2 - d) * n], [r1, r2])
    o = run(bed)
    f = by_file(o, o.sphere_forces(1)[0])
    Fn = hertz(r1 * r2 / (r1 + r2), d)
    assert np.allclose(f[1], Fn * n, rtol=1e-11, atol=0)
    assert np.allclose(f[0] + f[1], 0.0, atol=1e-18)


def test_coulomb_cap_and_drag_direction():
    r, d = 1e-3, 1e-5
    bed = single_sphere_clumps([[0, 0, 0], [2 * r - d, 0, 0]], r, vel=[[0, 0, 0], [0, 1.0, 0]])
    o = run(bed)
    f = by_file(o, o.sphere_forces(1)[0])
    pos = by_file(o, o.spheres()["pos"])
    n = (pos[1] - pos[0]) / np.linalg.norm(pos[1] - pos[0])
    fn = np.dot(f[0], n) * n
    ft = f[0] - fn
    mu = 0.5
    assert ft[1] > 0, "friction drags the lower sphere along the partner's motion"
    assert np.linalg.norm(ft) == pytest.approx(mu * np.linalg.norm(fn), rel=1e-9)
    assert np.linalg.norm(fn) == pytest.approx(hertz(0.5 * r, d), rel=1e-2)  # plus normal damping


def test_sticking_history_equals_minus_ks_dt_vt():
    r, d, dt, v = 1e-3, 1e-5, 1e-6, 1e-4
    bed = single_sphere_clumps([[0, 0, 0], [2 * r - d, 0, 0]], r, vel=[[0, 0, 0], [0, v, 0]], StepSize=dt)
    o = run(bed)
    h = o.pair_history()
    assert h.shape[0] == 1
    Rstar = 0.5 * r
    pos = by_file(o, o.spheres()["pos"])
    n = (pos[1] - pos[0]) / np.linalg.norm(pos[1] - pos[0])
    delta = 2 * r - np.linalg.norm(pos[1] - pos[0])
    ks = mindlin_ks(Rstar, delta)
    vrel = np.array([0.0, v, 0.0])
    vt = vrel - np.dot(vrel, n) * n
    assert np.allclose(h[0, 0:3], -ks * dt * vt, rtol=1e-9, atol=1e-22)


@pytest.mark.parametrize("which", ["rolling", "twisting"])
def test_rolling_and_twisting_caps(which):
    r, d = 1e-3, 1e-5
    w = [[0, 0, 2e6], [0, 0, 0]] if which == "rolling" else [[2e6, 0, 0], [0, 0, 0]]
    bed = single_sphere_clumps([[0, 0, 0], [2 * r - d, 0, 0]], r, angv=w)
    o = run(bed)
    h = o.pair_history()[0]
    Rstar, beta, mu = 0.5 * r, 0.5, 0.5
    nf = hertz(Rstar, d)
    if which == "rolling":
        assert np.linalg.norm(h[6:9]) == pytest.approx(0.525 * beta * Rstar * nf, rel=1e-9)
    else:
        assert np.linalg.norm(h[3:6]) == pytest.approx(0.65 * mu * beta * Rstar * nf, rel=1e-9)


def test_free_fall_is_exact():
    g, dt, n = -9.81, 1e-5, 400
    bed = single_sphere_clumps([[0, 0, 0.01]], 1e-3, vel=[[0.1, 0, 0.2]], Gravity=g, StepSize=dt, Thresh=1.0)
    o = run(bed, steps=n)
    c = o.clumps()
    t = n * dt
    assert c["pos"][0] == pytest.approx([0.1 * t, 0.0, 0.01 + 0.2 * t + 0.5 * g * t * t], rel=1e-12, abs=1e-15)
    assert c["vel"][0] == pytest.approx([0.1, 0.0, 0.2 + g * t], rel=1e-12, abs=1e-15)


def test_torque_free_isotropic_rotation():
    # constant world-frame spin: R(t) = exp([w] t) R0; the quaternion Taylor
    # step is second order, so halving dt cuts the error about four times
    w = np.array([300.0, -200.0, 500.0])
    q0 = np.array([0.2, -0.3, 0.1, 0.9])
    q0 /= np.linalg.norm(q0)
    errs = []
    for dt, n in ((1e-5, 1000), (5e-6, 2000)):
        bed = single_sphere_clumps([[0, 0, 0]], 1e-3, angv=[w], quat=[q0], StepSize=dt, Thresh=1.0)
        o = run(bed, steps=n)
        q = o.clumps()["quat"][0]
        assert np.linalg.norm(q) == pytest.approx(1.0, abs=1e-14)
        R = quat_to_matrix(q) @ quat_to_matrix(q0).T
        expect = rodrigues(w, np.linalg.norm(w) * n * dt)
        errs.append(np.abs(R - expect).max())
    assert errs[0] < 5e-5
    assert 3.0 < errs[0] / errs[1] < 5.0


def test_quaternion_convention_body_to_world():
    # a clump of two spheres with local offset along body x, rotated 90 deg about z
    q = np.array([0, 0, np.sin(np.pi / 4), np.cos(np.pi / 4)])
    p = make_params(2, 1, Thresh=1.0)
    local = np.array([[1e-3, 0, 0], [-1e-3, 0, 0]])
    bed = nbc.Bed(p, np.zeros(2, np.int32), np.zeros((2, 3)), local, np.ones(2) * 1e-5, np.ones(2) * 1e-3,
                  np.zeros((1, 3)), np.zeros((1, 3)), q[None], np.zeros((1, 3)), np.array([[1e-10] * 3 + [2e-5]]), None)
    o = run(bed)
    pos = by_file(o, o.spheres()["pos"])
    assert pos[0] == pytest.approx([0, 1e-3, 0], abs=1e-18)
    assert pos[1] == pytest.approx([0, -1e-3, 0], abs=1e-18)


def test_bruteforce_neighbour_set_and_stable_sort():
    bed = nbc.generate_bed(1, 150, seed=4)
    p = orc.finalize_params(bed)
    o = orc.Oracle(p)
    o.upload(bed, None)
    o.rebuild()
    g = o.grid()
    cell = g["cell"]
    assert np.all(np.diff(cell) >= 0), "cells sorted ascending"
    perm = g["perm"]
    assert np.array_equal(np.sort(perm), np.arange(bed.n))
    same = cell[1:] == cell[:-1]
    assert np.all(perm[1:][same] > perm[:-1][same]), "stable within a cell"
    counts = np.bincount(cell, minlength=p.CellNum)
    occupied = counts > 0
    assert np.array_equal(g["end"][occupied] - g["beg"][occupied], counts[occupied])
    assert np.all(g["beg"][~occupied] == -1)
    # neighbour set = brute force O(N^2) on the same positions
    pr = o.pairs()
    own = np.repeat(np.arange(bed.n), np.diff(pr["foff"]))
    grid_pairs = set(zip(own.tolist(), pr["fj"].tolist()))
    bi, bj = o.bruteforce_pairs()
    assert grid_pairs == set(zip(bi.tolist(), bj.tolist()))
    # back lists are the transpose of the front lists
    bown = np.repeat(np.arange(bed.n), np.diff(pr["boff"]))
    assert set(zip(pr["bi"].tolist(), bown.tolist())) == grid_pairs
    slots = pr["bslot"]
    assert np.all(slots >= 0)
    assert np.array_equal(pr["fj"][slots], bown)


def test_cell_hash_known_points():
    bed = single_sphere_clumps([[0.0012, 0.0034, 0.0051], [0.02, -0.01, 0.03]], 1e-3)
    p = orc.finalize_params(bed)
    o = orc.Oracle(p)
    o.upload(bed, None)
    o.rebuild()
    cs = p.CellSize[0]
    assert cs == 3e-3
    cells = []
    for x in bed.pos:
        c = np.floor((x - np.array(p.WorldOrigin[:])) / cs).astype(int)
        cells.append(c[2] * p.GridDim[1] * p.GridDim[0] + c[1] * p.GridDim[0] + c[0])
    assert sorted(cells) == o.grid()["cell"].tolist()


def test_newton_third_law_in_a_bed():
    bed = nbc.generate_bed(1, 300, seed=8)
    bed.params.StepSize = 1e-5  # let the deposited bed settle onto its floor
    o = run(bed, steps=300, wall=bed.wall, opt=nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER))
    f, _ = o.sphere_forces(1)
    assert np.abs(f).sum() > 0
    assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).sum()


def test_cube_floor_contact_is_hertz_with_sphere_radius():
    r, d = 1e-3, 2e-5
    wall = cube_wall((0.01, 0.01, 0.01), (0.0, 0.0, 0.01))
    bed = single_sphere_clumps([[0.0, 0.0, r - d]], r, isWallForce=1, lo=(-0.01, -0.01, 0.0), size=(0.02, 0.02, 0.02))
    o = run(bed, wall=wall)
    f = o.sphere_forces(2)[0][0]
    assert f == pytest.approx([0, 0, hertz(r, d)], rel=1e-12, abs=1e-18)


def test_cylinder_side_contact_is_hertz():
    r, d, R = 1e-3, 1e-5, 0.02
    wall = cylinder_wall(R, 0.05, (0.0, 0.0, 0.05))
    bed = single_sphere_clumps([[R - r + d, 0.0, 0.05]], r, isWallForce=1, lo=(-0.02, -0.02, 0.0), size=(0.04, 0.04, 0.1))
    o = run(bed, wall=wall, opt=nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER))
    f = o.sphere_forces(2)[0][0]
    assert f == pytest.approx([-hertz(r, d), 0, 0], rel=1e-9, abs=1e-18)


def test_sphere_penetrator_contact_is_hertz():
    r, Rp, d = 1e-3, 5e-3, 1e-5
    bed = single_sphere_clumps([[0.0, 0.0, 0.0]], r, isIntruder=1)
    pene = nbc.sphere_penetrator(Rp, (0.0, 0.0, r + Rp - d), (0.0, 0.0, 0.0))
    o = run(bed, pene=pene)
    f = o.sphere_forces(3)[0][0]
    assert f == pytest.approx([0, 0, -hertz(r * Rp / (r + Rp), d)], rel=1e-9, abs=1e-18)
    pen = o.penetrator()
    assert pen["cache_f"][0] == pytest.approx([0, 0, hertz(r * Rp / (r + Rp), d)], rel=1e-9, abs=1e-18)


def test_clump_torque_from_offset_sphere():
    # two-sphere clump, only the +x sphere touches the floor
    r, d = 1e-3, 2e-5
    p = make_params(2, 1, isWallForce=1, lo=(-0.01, -0.01, 0.0), size=(0.02, 0.02, 0.02), Thresh=1.0)
    local = np.array([[2e-3, 0, 0], [-2e-3, 0, 0]])
    cpos = np.array([[0.0, 0.0, 0.005]])
    q = np.array([[0.0, np.sin(0.3 / 2), 0.0, np.cos(0.3 / 2)]])  # tilt about y: +x sphere lowest
    R = quat_to_matrix(q[0])
    pos = cpos + local @ R.T
    shift = pos[0, 2] - (r - d)
    cpos[0, 2] -= shift
    pos = cpos + local @ R.T
    assert pos[1, 2] > r
    bed = nbc.Bed(p, np.zeros(2, np.int32), pos, local, np.ones(2) * 4e-6, np.ones(2) * r, cpos, np.zeros((1, 3)), q,
                  np.zeros((1, 3)), np.array([[1e-11, 1e-11, 1e-11, 8e-6]]), None)
    wall = cube_wall((0.01, 0.01, 0.01), (0.0, 0.0, 0.01))
    o = run(bed, wall=wall)
    c = o.clumps()
    f = by_file(o, o.sphere_forces(2)[0])
    sp = by_file(o, o.spheres()["pos"])
    arm = sp[0] - c["pos"][0]
    assert c["F"][0] == pytest.approx(f.sum(axis=0), rel=1e-12)
    assert c["T"][0] == pytest.approx(np.cross(arm, f[0]), rel=1e-9, abs=1e-20)


def test_restitution_band_monotone():
    def cor(e):
        r, v = 1e-3, 0.1
        bed = single_sphere_clumps([[0, 0, 0], [2 * r + 1e-6, 0, 0]], r, vel=[[v / 2, 0, 0], [-v / 2, 0, 0]],
                                   NormRes=e, StepSize=1e-7, Thresh=1.0)
        o = run(bed, steps=2000)
        cv = o.clumps()["vel"]
        return (cv[1, 0] - cv[0, 0]) / v

    e1, e2 = cor(0.3), cor(0.7)
    assert 0.0 < e1 < e2 < 1.0

"""Pins of the DEMmain single-sphere loops in the FP64 oracle
(src/dem_simulation_{cube,cylind,split,hgrid}.cu run as one-sphere clumps).

Independent statements used: constant-acceleration kinematics (the drift sum
and the step at which |cum| first passes the rebuild threshold), the linear map
of the glue type-1 drive iterated with numpy.linalg.matrix_power, the clump
loop on the same one-sphere clumps (its contact and cylinder-wall laws differ
from DEMmain's only in how the spin term is associated), the Hertz-Mindlin
closed forms of the DEMmain cube wall (normal force, twice-damped tangential
force, no twisting torque), and a two-pair bed whose list slots swap at a
rebuild, where the cylinder loop's history exchange and the cube loop's
inheritance must give different tangential forces.
"""
import numpy as np
import pytest

import nbc
import orc
from util import cube_wall, cylinder_wall, single_sphere_clumps


def part_of(bed, types=None):
    p = nbc.Particles.from_bed(bed)
    if types is not None:
        p.type[:] = types
    return p


def oracle(part, driver, wall, levels=0):
    opt = nbc.make_options(driver=driver, hgrid_levels=levels)
    o = orc.Oracle(orc.finalize_params(part, 0.0, levels), opt)
    o.upload_particles(part, wall)
    return o


def big_cube():
    return cube_wall((0.05, 0.05, 0.05), (0.0, 0.0, 0.0))


def big_cylinder():
    return cylinder_wall(0.05, 0.05, (0.0, 0.0, 0.0))


@pytest.mark.parametrize("driver", [nbc.NBC_DRIVER_CUBE, nbc.NBC_DRIVER_CYLINDER, nbc.NBC_DRIVER_SPLIT])
def test_free_fall_displacement_and_rebuild_step(driver):
    # one particle in free fall: z_n = z0 + g (n dt)^2 / 2; the first step's
    # rebuild clears cum after its drift, so cum = z_n - z_1 and the next
    # rebuild comes at the first n with |z_n - z_1| past the threshold
    g, dt, thresh, r = -9.81, 1e-4, 2e-4, 1e-3
    bed = single_sphere_clumps([[0.0, 0.0, 0.0]], r, Gravity=g, StepSize=dt, Thresh=thresh)
    wall = big_cube() if driver == nbc.NBC_DRIVER_CUBE else big_cylinder()
    o = oracle(part_of(bed), driver, wall)
    limit = 0.1 * r if driver == nbc.NBC_DRIVER_SPLIT else thresh
    def drop(n):
        return 0.5 * abs(g) * ((n * dt) ** 2 - dt ** 2)
    n_star = next(n for n in range(1, 1000) if drop(n) > limit)
    assert drop(n_star - 1) < limit * (1 - 1e-9) and drop(n_star) > limit * (1 + 1e-9)  # far from rounding
    o.step(n_star - 1)
    assert o.stats().rebuilds == 1
    z = o.clumps()["pos"][0, 2]
    assert abs(z - 0.5 * g * ((n_star - 1) * dt) ** 2) <= 1e-15
    cum = o.spheres()["cum"][0]
    assert abs(cum[2] - (z - 0.5 * g * dt ** 2)) <= 1e-17 and cum[0] == 0.0 and cum[1] == 0.0
    o.step(1)
    assert o.stats().rebuilds == 2
    assert np.all(o.spheres()["cum"][0] == 0.0)


def test_glue_types_follow_their_kinematics():
    # type 1 at (x, y) driven by v = (-0.2 y, -0.2 x, 0): (x, y)_{n+1} = M (x, y)_n,
    # M = [[1, -0.2 dt], [-0.2 dt, 1]]; type 2 never moves, even when pressed
    dt, r = 1e-6, 1e-3
    c = [[0.010, 0.004, 0.0], [-0.02, 0.0, 0.0], [-0.02 + 1.9e-3, 0.0, 0.0]]
    v = [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [-0.3, 0.0, 0.0]]
    w = [[0.0, 0.0, 5.0], [1.0, 2.0, 3.0], [0.0, 0.0, 0.0]]
    bed = single_sphere_clumps(c, r, vel=v, angv=w, StepSize=dt, Thresh=1.0, Gravity=-9.81)
    part = part_of(bed, [1, 2, 0])
    o = oracle(part, nbc.NBC_DRIVER_SPLIT, big_cylinder())
    n = 400
    o.step(1)
    # the fixed particle is pressed by the free one in the first step
    assert o.stats().touching_ss == 1
    f, _ = o.sphere_forces(1)
    assert np.abs(f[o.spheres()["fid"] == 1]).max() > 0
    o.step(n - 1)
    cl = o.clumps()
    M = np.array([[1.0, -0.2 * dt], [-0.2 * dt, 1.0]])
    xy = np.linalg.matrix_power(M, n) @ np.array(c[0][:2])
    assert np.abs(cl["pos"][0, :2] - xy).max() <= 1e-14 * np.abs(xy).max()
    assert cl["pos"][0, 2] == 0.0
    prev = np.linalg.matrix_power(M, n - 1) @ np.array(c[0][:2])
    assert np.abs(cl["vel"][0] - [-0.2 * prev[1], -0.2 * prev[0], 0.0]).max() <= 1e-14 * 0.2 * np.abs(prev).max()
    assert np.array_equal(cl["angv"][0], w[0]) and np.array_equal(cl["quat"][0], [0, 0, 0, 1])
    # the fixed particle did not move
    assert np.array_equal(cl["pos"][1], c[1]) and np.array_equal(cl["vel"][1], v[1])
    assert np.array_equal(cl["angv"][1], w[1]) and np.array_equal(cl["quat"][1], [0, 0, 0, 1])
    # the free particle bounced off
    assert cl["vel"][2, 0] > 0.0


def test_contact_law_matches_the_clump_loop():
    # oblique contact with spins: DEMmain's An (u x wJ) + Ap (u x wI) and the
    # clump loop's wJ x (-An u) - wI x (Ap u) are the same vector
    r = np.array([1.0e-3, 1.3e-3])
    c = [[0.0, 0.0, 0.0], [1.6e-3, 1.0e-3, 0.4e-3]]
    v = [[0.1, -0.2, 0.05], [-0.3, 0.1, 0.2]]
    w = [[30.0, -10.0, 50.0], [-20.0, 40.0, 10.0]]
    res = []
    for driver in (nbc.NBC_DRIVER_CLUMP, nbc.NBC_DRIVER_CYLINDER):
        bed = single_sphere_clumps(c, r, vel=v, angv=w, isWallForce=1, Gravity=0.0)
        wall = big_cylinder()
        if driver == nbc.NBC_DRIVER_CLUMP:
            opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER)
            o = orc.Oracle(orc.finalize_params(bed), opt)
            o.upload(bed, wall)
        else:
            o = oracle(part_of(bed), driver, wall)
        o.step(3)
        f, t = o.sphere_forces(1)
        s = o.spheres()
        res.append((f[np.argsort(s["fid"])], t[np.argsort(s["fid"])]))
    (fa, ta), (fb, tb) = res
    assert np.abs(fa).max() > 0 and np.abs(ta).max() > 0
    assert np.abs(fa - fb).max() <= 1e-13 * np.abs(fa).max()
    assert np.abs(ta - tb).max() <= 1e-13 * np.abs(ta).max()
    # Newton's third law on the forces
    assert np.abs(fb[0] + fb[1]).max() <= 1e-15 * np.abs(fb).max()


def hertz(P, r, m, dn):
    lncor = np.log(P.NormRes)
    damping = 2.0 * np.sqrt(5.0 / 6.0) * lncor / np.sqrt(lncor ** 2 + np.pi ** 2)
    kn = 2.0 * P.Ymod * np.sqrt(r * dn) / (3.0 * (1.0 - P.Prat ** 2))
    ks = 2.0 * P.Ymod / (1.0 + P.Prat) / (2.0 - P.Prat) * np.sqrt(r * dn)
    return kn, damping * np.sqrt(kn * m), ks, damping * np.sqrt(ks * m)


def test_cube_wall_law():
    # a sphere on the floor of the cube, overlap dn, sliding at vt along x and
    # spinning at wz about the normal: normal Kn dn, tangential Cs vt with the
    # factor 2 in Cs (below the Coulomb cap), and no wall torque about z
    # (the DEMmain cube wall does not apply its twisting moment)
    r, dn, vt, wz = 1e-3, 2e-6, 1e-3, 50.0
    bed = single_sphere_clumps([[0.0, 0.0, -0.02 + r - dn]], r, vel=[[vt, 0.0, 0.0]], angv=[[0.0, 0.0, wz]],
                               Gravity=0.0)
    m = bed.mass[0]
    wall = cube_wall((0.02, 0.02, 0.02), (0.0, 0.0, 0.0))
    o = oracle(part_of(bed), nbc.NBC_DRIVER_CUBE, wall)
    o.step(1)
    f, t = o.sphere_forces(2)
    P = bed.params
    # position after the first drift (no force yet): the overlap is unchanged
    kn, cn, ks, cs = hertz(P, r, m, dn)
    assert f[0, 2] == pytest.approx(kn * dn, rel=1e-12)
    # Ds = dt vt, tangential -Ks Ds + 2 Cs' vt, both along -x / +x
    tf = -ks * P.StepSize * vt + 2.0 * cs * vt
    assert abs(tf) < P.Frcoe * kn * dn
    assert f[0, 0] == pytest.approx(tf, rel=1e-12)
    assert t[0, 2] == 0.0
    # torque -Ap u x tf with u = (0, 0, -1) pointing from the contact to the centre? here
    # u = RVn / |RVn| = (0, 0, +1): T = -(r - dn) (u x tf e_x) = -(r - dn) tf e_y
    assert t[0, 1] == pytest.approx(-(r - dn) * tf, rel=1e-12)
    # the cylinder loop on its bottom cap applies the twisting moment
    wall2 = cylinder_wall(0.02, 0.02, (0.0, 0.0, 0.0))
    o2 = oracle(part_of(bed), nbc.NBC_DRIVER_CYLINDER, wall2)
    o2.step(1)
    _, t2 = o2.sphere_forces(2)
    assert abs(t2[0, 2]) > 0


def slot_swap_bed():
    # pair A (a1, a2) rises through a grid-layer boundary (z) while a1 spins,
    # so the contact slips and carries a tangential history; pair B (b1, b2)
    # rests. Both overlap by 1 nm, little enough to stay in contact for the
    # few hundred steps. A sorts before B at the first list and after it at
    # the next one (same layer, larger y), so their list slots swap.
    r = 1e-3
    gap = 2 * r - 1e-9
    c = [[0.010, 0.030, 0.0209], [0.010, 0.030 + gap, 0.0209],
         [0.010, 0.010, 0.0225], [0.010, 0.010 + gap, 0.0225]]
    v = [[0.0, 0.0, 1.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
    w = [[0.0, 0.0, 200.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
    return single_sphere_clumps(c, r, vel=v, angv=w, Gravity=0.0, StepSize=1e-6, Thresh=2e-4, lo=(0.0, 0.0, 0.0),
                                size=(0.05, 0.05, 0.05))


def test_cylinder_loop_swaps_histories_at_a_rebuild():
    out = {}
    for driver, wall in ((nbc.NBC_DRIVER_CUBE, cube_wall((0.025, 0.025, 0.025), (0.025, 0.025, 0.025))),
                         (nbc.NBC_DRIVER_CYLINDER, cylinder_wall(0.024, 0.025, (0.025, 0.025, 0.025)))):
        o = oracle(part_of(slot_swap_bed()), driver, wall)
        o.step(1)
        fid0 = o.spheres()["fid"].copy()
        while o.stats().rebuilds < 2:
            o.step(1)
        s = o.spheres()
        assert not np.array_equal(s["fid"], fid0)  # the order changed
        st = o.stats()
        assert st.touching_ss == 2 and st.touching_wall == 0
        f, _ = o.sphere_forces(1)
        out[driver] = f[np.argsort(s["fid"])]
    fc, fy = out[nbc.NBC_DRIVER_CUBE], out[nbc.NBC_DRIVER_CYLINDER]
    assert np.abs(fc[2:, 1]).max() > 0 and np.abs(fc[:2, 0]).max() > 0
    # B rests: with inherited histories its force is normal (y)
    assert np.abs(fc[2:, [0, 2]]).max() <= 1e-12 * np.abs(fc[2:, 1]).max()
    # with the exchange, B read A's slipping history from the slot it now holds
    assert np.abs(fy[2:, 0]).max() > 0.1 * np.abs(fy[2:, 1]).max()


def pene_case(driver, cohesion=0.0):
    # one particle under a sphere penetrator pressing down at constant velocity,
    # the particle spinning and sliding: the penetrator law and the drift
    r = 1e-3
    bed = single_sphere_clumps([[0.0, 0.0, 0.0]], r, vel=[[0.02, 0.0, 0.0]], angv=[[5.0, -3.0, 20.0]],
                               Gravity=0.0, Cohesion=cohesion, isIntruder=1, isWallForce=1)
    pene = nbc.sphere_penetrator(2e-3, (0.0, 0.0, r + 2e-3 - 2e-6), (0.0, 0.0, -0.01))
    if driver == nbc.NBC_DRIVER_CLUMP:
        opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER)
        o = orc.Oracle(orc.finalize_params(bed), opt)
        o.upload(bed, big_cylinder(), pene)
    else:
        opt = nbc.make_options(driver=driver)
        o = orc.Oracle(orc.finalize_params(part_of(bed)), opt)
        o.upload_particles(part_of(bed), big_cylinder(), pene)
    return o, bed


@pytest.mark.parametrize("cohesion", [0.0, 2e4])
def test_penetrator_law_matches_the_clump_loop(cohesion):
    # the DEMmain penetrator differs from the clump one by the association of
    # the spin term and by the cohesion term -Coh (Beta Rij)^2 n along the
    # contact normal (src/dem_force_penetrator.cu:91, 119)
    out = {}
    for driver in (nbc.NBC_DRIVER_CLUMP, nbc.NBC_DRIVER_CYLINDER):
        o, bed = pene_case(driver, cohesion)
        o.step(1)
        f, t = o.sphere_forces(3)
        out[driver] = (f[0], t[0], o.clumps()["pos"][0] - o.penetrator()["pos"])
        assert o.stats().touching_wall == 1
    (fc, tc, pc), (fd, td, pd) = out[nbc.NBC_DRIVER_CLUMP], out[nbc.NBC_DRIVER_CYLINDER]
    assert np.abs(tc - td).max() <= 1e-13 * np.abs(tc).max()
    r, R, P = 1e-3, 2e-3, bed.params
    rij = r * R / (r + R)
    n = pd / np.linalg.norm(pd)  # from the penetrator's centre to the particle's, at the force evaluation
    want = -cohesion * (P.Beta * rij) ** 2 * n
    assert np.abs((fd - fc) - want).max() <= 1e-13 * np.abs(fc).max() + 1e-15
    if cohesion:
        assert np.abs(fd - fc).max() > 1e-6 * np.abs(fc).max()


def test_demmain_drift_leaves_the_penetrator_out():
    # the clump force the next drift uses is Fc + Fw in the DEMmain loops
    # (src/dem_integrate.cu:23-25) and Fc + Fw + Fi in the clump loop
    for driver in (nbc.NBC_DRIVER_CLUMP, nbc.NBC_DRIVER_CYLINDER):
        o, _ = pene_case(driver)
        o.step(2)
        F = o.clumps()["F"][0]
        f1, t1 = o.sphere_forces(1)
        f2, t2 = o.sphere_forces(2)
        f3, t3 = o.sphere_forces(3)
        assert np.abs(f3).max() > 0
        want = f1[0] + f2[0] if driver else f1[0] + f2[0] + f3[0]
        assert np.array_equal(F, want), driver


def write_grid_bt(path, p):
    with open(path, "w") as f:
        for t in range(p.Levels):
            d = p.LevelDim[t]
            f.write(f"{t} {p.LevelCell[t]!r} {d[0]} {d[1]} {d[2]} {p.LevelOff[t]} {d[0] * d[1] * d[2]}\n")


def test_grid_bt_levels_by_type(tmp_path):
    # a grid.bt holding the generated 3-level table, and every particle's type
    # set to its radius level: the same lists and forces as the radius rule;
    # every particle on level 0 instead: the KD-tree pair set (complete search)
    import ctypes as C
    from scipy.spatial import cKDTree
    bed = nbc.generate_bed(3, 1500, seed=9, container=(0.012, 0, 0, 0))
    # widen the radii so that three levels are populated
    rng = np.random.default_rng(9)
    part = nbc.Particles.from_bed(bed)
    part.radius[:] = part.radius * rng.choice([0.4, 0.7, 1.0], part.n)
    opt_r = nbc.make_options(driver=nbc.NBC_DRIVER_HGRID, hgrid_levels=3)
    pr = orc.finalize_params(part, 0.0, 3)
    write_grid_bt(tmp_path / "grid.bt", pr)
    pg = nbc.Params()
    C.memmove(C.byref(pg), C.byref(pr), C.sizeof(nbc.Params))
    for t in range(3):
        pg.LevelCell[t] = pg.LevelDim[t][0] = pg.LevelOff[t] = 0
    pg.GridLevels = 3
    assert nbc.lib().nbc_load_grid_levels(str(tmp_path / "grid.bt").encode(), C.byref(pg)) == 0
    for t in range(3):
        assert pg.LevelCell[t] == pr.LevelCell[t] and list(pg.LevelDim[t]) == list(pr.LevelDim[t])
        assert pg.LevelOff[t] == pr.LevelOff[t]
    assert pg.CellNum == pr.CellNum and pg.Levels == 3
    lvl = np.array([next(t for t in range(3) if t == 2 or r > pr.LevelRmin[t]) for r in part.radius])
    assert set(lvl) == {0, 1, 2}
    res = []
    for by_type in (0, 1):
        q = nbc.Particles(part.params, lvl if by_type else part.type, part.pos, part.vel, part.quat, part.angv,
                          part.prop, part.radius, part.wall)
        opt = nbc.make_options(driver=nbc.NBC_DRIVER_HGRID, hgrid_levels=0 if by_type else 3,
                               level_from_type=by_type)
        o = orc.Oracle(pg if by_type else pr, opt)
        o.upload_particles(q, part.wall)
        o.step(3)
        res.append((o.pairs(), o.sphere_forces(1)[0], o.spheres()["fid"]))
    (pa, fa, ia), (pb, fb, ib) = res
    assert np.array_equal(ia, ib)
    for k in ("foff", "fj", "boff", "bi", "bslot"):
        assert np.array_equal(pa[k], pb[k]), k
    assert np.array_equal(fa, fb)
    # all on level 0: a uniform search at the coarsest cell
    q = nbc.Particles(part.params, np.zeros(part.n, np.int32), part.pos, part.vel, part.quat, part.angv, part.prop,
                      part.radius, part.wall)
    o = orc.Oracle(pg, nbc.make_options(driver=nbc.NBC_DRIVER_HGRID, level_from_type=1))
    o.upload_particles(q, part.wall)
    o.rebuild()
    s, p = o.spheres(), o.pairs()
    own = np.repeat(np.arange(part.n), np.diff(p["foff"]))
    rows = {tuple(sorted(x)) for x in zip(s["fid"][own].tolist(), s["fid"][p["fj"]].tolist())}
    tree = cKDTree(part.pos)
    cand = tree.query_pairs(3.0 * part.radius.max() * (1 + 1e-9), output_type="ndarray")
    d2 = ((part.pos[cand[:, 0]] - part.pos[cand[:, 1]]) ** 2).sum(1)
    lim = 1.5 * (part.radius[cand[:, 0]] + part.radius[cand[:, 1]])
    want = {tuple(sorted(x)) for x in cand[d2 < lim * lim].tolist()}
    assert rows == want


def test_grid_bt_rejects_bad_tables(tmp_path):
    import ctypes as C
    p = nbc.Params()
    p.GridLevels = 2
    p.NormRes = 0.5
    f = tmp_path / "grid.bt"
    f.write_text("0 0.003 10 10 10 0 1000\n1 0.001 30 30 30 999 27000\n")  # offset not contiguous
    assert nbc.lib().nbc_load_grid_levels(str(f).encode(), C.byref(p)) == nbc.NBC_E_BADINPUT
    f.write_text("0 0.001 30 30 30 0 27000\n1 0.003 10 10 10 27000 1000\n")  # finer level first
    assert nbc.lib().nbc_load_grid_levels(str(f).encode(), C.byref(p)) == nbc.NBC_E_BADINPUT
    f.write_text("0 0.003 10 10 10 0 1000\n")  # a row short
    assert nbc.lib().nbc_load_grid_levels(str(f).encode(), C.byref(p)) == nbc.NBC_E_IO
    f.write_text("0 0.003 10 10 10 0 1000\n1 0.001 30 30 30 1000 27000\n")
    assert nbc.lib().nbc_load_grid_levels(str(f).encode(), C.byref(p)) == 0
    assert p.Levels == 2 and p.CellNum == 28001 and p.LevelOff[1] == 1000 and p.finalized == 1

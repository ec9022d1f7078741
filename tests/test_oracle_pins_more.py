"""More pins of the FP64 oracle: penetrator shapes, torque-driven rotation,
the gravity-mode penetrator, the kinetic-energy sum, force-chain exports, the
invalid-input guard and the compact slot mode.

Each check rests on something the oracle's own code does not restate: an
independent point-to-mesh / point-to-cylinder distance, conservation laws of
the discrete integrator (derived in the docstrings), a hand-derived closed
form of the attitude update under a constant torque, the equivalence of the
gravity-mode sphere penetrator with a two-sphere contact, the radical plane of
two spheres, or a second mode of the oracle that must agree bit for bit.
"""
import numpy as np
import pytest

import nbc
import orc
from util import cube_wall, make_params, quat_to_matrix, rodrigues, single_sphere_clumps

E, NU = 1e8, 0.3
KH = 2.0 * E / (3.0 * (1.0 - NU * NU))  # F = KH sqrt(R*) delta^1.5 (Hertz, two equal elastic bodies)


def hertz(Rstar, delta):
    return KH * np.sqrt(Rstar) * delta ** 1.5


def overlap_from_force(F, Rstar):
    return (F / (KH * np.sqrt(Rstar))) ** (2.0 / 3.0)


def by_file(o, a):
    fid = o.spheres()["fid"]
    out = np.empty_like(a)
    out[fid] = a
    return out


def run(bed, steps=1, wall=None, pene=None, opt=None, **kw):
    o = orc.Oracle(orc.finalize_params(bed), opt, **kw)
    o.upload(bed, wall, pene)
    o.step(steps)
    return o


# ---------------------------------------------------------------------------
# independent geometry: closest point on a triangle mesh (plane projection with
# a barycentric inside test, else the nearest of the three edge segments)
def closest_on_mesh(p, tri):
    best_d, best_q = np.inf, None
    for a, b, c in tri:
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        q = p - np.dot(p - a, n) * n
        # barycentric coordinates of q in abc
        v0, v1, v2 = b - a, c - a, q - a
        d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
        d20, d21 = v2 @ v0, v2 @ v1
        den = d00 * d11 - d01 * d01
        v = (d11 * d20 - d01 * d21) / den
        w = (d00 * d21 - d01 * d20) / den
        cands = []
        if v >= 0 and w >= 0 and v + w <= 1:
            cands.append(q)
        else:
            for s0, s1 in ((a, b), (b, c), (c, a)):
                e = s1 - s0
                t = np.clip(np.dot(p - s0, e) / (e @ e), 0.0, 1.0)
                cands.append(s0 + t * e)
        for x in cands:
            d = np.linalg.norm(p - x)
            if d < best_d:
                best_d, best_q = d, x
    return best_d, best_q


def mesh_outside(p, tri, centre):
    """p outside the convex mesh: in front of some face (normals oriented away from centre)."""
    for a, b, c in tri:
        n = np.cross(b - a, c - a)
        if np.dot(n, a - centre) < 0:
            n = -n
        if np.dot(p - a, n) > 0:
            return True
    return False


def pene_bed(centres, radii, **kw):
    bed = single_sphere_clumps(centres, radii, isIntruder=1, BranchLmtLen=kw.pop("BranchLmtLen", 256), **kw)
    return bed


# ---------------------------------------------------------------------------
# cone (triangulated): no reference counterpart; the mesh contact is pinned
# against the independent distance above and Hertz with R* = sphere radius
# (src/dem_force_penetrator_clump.cu:62: Rij = Radius for a non-sphere shape)

def cone(R=4e-3, nseg=16):
    H = R / np.tan(np.radians(30.0))
    pene = nbc.cone_penetrator(R, H, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), nseg=nseg)
    return pene, pene._tri.reshape(-1, 3, 3), H


def test_cone_facet_contact_is_hertz_along_the_facet_normal():
    pene, tri, H = cone()
    apex = np.array([0.0, 0.0, -H])
    lateral = [t for t in tri if np.any(np.all(np.isclose(t, apex, atol=1e-15), axis=1))]
    t = lateral[3]
    c = t.mean(axis=0)
    n = np.cross(t[1] - t[0], t[2] - t[0])
    n /= np.linalg.norm(n)
    if np.dot(n, [c[0], c[1], 0.0]) < 0:
        n = -n
    r, d = 1e-3, 2e-5
    for gap, touching in ((r - d, True), (r + d, False)):
        bed = pene_bed([c + gap * n], r)
        o = run(bed, pene=pene)
        f = o.sphere_forces(3)[0][0]
        if touching:
            assert f == pytest.approx(hertz(r, d) * n, rel=1e-9, abs=1e-18)
            assert o.stats().touching_wall == 1
            assert o.penetrator()["cache_f"][0] == pytest.approx(-f, rel=1e-12)
        else:
            assert np.all(f == 0.0) and o.stats().touching_wall == 0


def test_cone_apex_and_rim_vertex_contacts():
    pene, tri, H = cone()
    r, d = 1e-3, 1.5e-5
    R = 4e-3
    verts = tri.reshape(-1, 3)
    rim = verts[np.argmin(np.linalg.norm(verts - [R, 0, 0], axis=1))]
    cases = [(np.array([0.0, 0.0, -H]), np.array([0.0, 0.0, -1.0])),            # below the apex
             (rim, np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0))]                   # out and up from a rim vertex
    for vtx, u in cases:
        p = vtx + (r - d) * u
        dist, q = closest_on_mesh(p, tri)
        assert np.allclose(q, vtx, atol=1e-15), "the vertex is the closest feature"
        o = run(pene_bed([p], r), pene=pene)
        f = o.sphere_forces(3)[0][0]
        assert f == pytest.approx(hertz(r, d) * u, rel=1e-8, abs=1e-18)


def test_cone_random_contacts_against_independent_distance():
    pene, tri, H = cone(nseg=12)
    rng = np.random.default_rng(7)
    centre = np.array([0.0, 0.0, -H / 3.0])
    pts, rad, exp_d, exp_n = [], [], [], []
    while len(pts) < 30:
        p = rng.uniform([-6e-3, -6e-3, -H - 2e-3], [6e-3, 6e-3, 2e-3])
        if not mesh_outside(p, tri, centre):
            continue
        dist, q = closest_on_mesh(p, tri)
        if not (2e-4 < dist < 9e-4):
            continue
        if pts and np.min(np.linalg.norm(np.array(pts) - p, axis=1)) < 3e-4:
            continue  # keep spheres apart (sphere-sphere forces are not in kind 3 anyway)
        delta = rng.uniform(1e-6, 4e-5)
        pts.append(p)
        rad.append(dist + delta)
        exp_d.append(delta)
        exp_n.append((p - q) / dist)
    o = run(pene_bed(pts, np.array(rad)), pene=pene)
    f = by_file(o, o.sphere_forces(3)[0])
    F = np.linalg.norm(f, axis=1)
    assert o.stats().touching_wall == len(pts)
    got_d = overlap_from_force(F, np.array(rad))
    assert np.allclose(got_d, exp_d, rtol=1e-8, atol=0)
    assert np.allclose(f / F[:, None], exp_n, atol=1e-9)


# ---------------------------------------------------------------------------
# cylinder penetrator (src/dem_contact_penetrator.cu:42-83): side, cap and
# rim contacts against the analytic distance to a solid cylinder

def cyl_distance(p, R, H):
    rho = np.hypot(p[0], p[1])
    er = np.array([p[0], p[1], 0.0]) / rho if rho > 0 else np.array([1.0, 0.0, 0.0])
    ez = np.array([0.0, 0.0, np.sign(p[2])])
    if abs(p[2]) <= H:
        return rho - R, er
    if rho <= R:
        return abs(p[2]) - H, ez
    d = np.hypot(rho - R, abs(p[2]) - H)
    return d, ((rho - R) * er + (abs(p[2]) - H) * ez) / d


@pytest.mark.parametrize("where", ["side", "top", "bottom", "rim", "random"])
def test_cylinder_penetrator_contacts(where):
    R, H = 3e-3, 5e-3
    pene = nbc.cylinder_penetrator(R, H, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    r, d = 1e-3, 2e-5
    if where == "random":
        rng = np.random.default_rng(3)
        pts, rad, exp_d, exp_n = [], [], [], []
        while len(pts) < 24:
            p = rng.uniform([-R - 1e-3, -R - 1e-3, -H - 1e-3], [R + 1e-3, R + 1e-3, H + 1e-3])
            dist, n = cyl_distance(p, R, H)
            if not (2e-4 < dist < 9e-4):
                continue
            if pts and np.min(np.linalg.norm(np.array(pts) - p, axis=1)) < 3e-4:
                continue
            delta = rng.uniform(1e-6, 4e-5)
            pts.append(p); rad.append(dist + delta); exp_d.append(delta); exp_n.append(n)
    else:
        u = {"side": np.array([np.cos(0.7), np.sin(0.7), 0.0]), "top": np.array([0.0, 0.0, 1.0]),
             "bottom": np.array([0.0, 0.0, -1.0]), "rim": np.array([1.0, 0.0, 1.0]) / np.sqrt(2)}[where]
        foot = {"side": R * u + [0, 0, 0.3 * H], "top": np.array([0.4 * R, -0.2 * R, H]),
                "bottom": np.array([-0.3 * R, 0.1 * R, -H]), "rim": np.array([R, 0.0, H])}[where]
        pts, rad, exp_d, exp_n = [foot + (r - d) * u], [r], [d], [u]
        assert cyl_distance(pts[0], R, H)[0] == pytest.approx(r - d, rel=1e-12)
    o = run(pene_bed(pts, np.array(rad)), pene=pene)
    f = by_file(o, o.sphere_forces(3)[0])
    F = np.linalg.norm(f, axis=1)
    assert o.stats().touching_wall == len(pts)
    assert np.allclose(overlap_from_force(F, np.array(rad)), exp_d, rtol=1e-8, atol=0)
    assert np.allclose(f / F[:, None], exp_n, atol=1e-9)
    # the reaction history records -sum F on the penetrator
    assert o.penetrator()["cache_f"][0] == pytest.approx(-f.sum(axis=0), rel=1e-12)


# ---------------------------------------------------------------------------
# gravity-mode penetrator (src/dem_integrate_penetrator.cu:3-41, 79-115)

def test_gravity_penetrator_free_fall_is_exact():
    g, dt, n = -9.81, 1e-5, 500
    bed = pene_bed([[0.0, 0.0, 0.0]], 1e-3, Gravity=g, StepSize=dt, Thresh=1.0, OutputInterval=n)
    z0, v0 = 0.04, -0.3
    pene = nbc.sphere_penetrator(3e-3, (0.0, 0.0, z0), (0.01, 0.0, v0), mass=2e-3, motion=nbc.NBC_PENE_GRAVITY)
    o = run(bed, steps=n, pene=pene)
    t = n * dt
    pen = o.penetrator()
    assert pen["pos"] == pytest.approx([0.01 * t, 0.0, z0 + v0 * t + 0.5 * g * t * t], rel=1e-12, abs=1e-15)
    assert pen["vel"] == pytest.approx([0.01, 0.0, v0 + g * t], rel=1e-12, abs=1e-15)
    assert np.all(pen["cache_f"] == 0.0)


def test_gravity_penetrator_collision_equals_two_sphere_contact():
    """A gravity-mode sphere penetrator (mass M, radius Rp) uses R* = r Rp/(r+Rp),
    M* = m M/(m+M), Cs with factor 1 and takes the reaction -sum F
    (src/dem_force_penetrator_clump.cu:60-70, src/dem_integrate_penetrator.cu):
    exactly the sphere-sphere law between two free clumps. A frictionless
    oblique collision run both ways must agree, and momentum is conserved."""
    r, m = 1e-3, 1.2e-5
    Rp, M = 1.5e-3, 3.0e-5
    dt, steps = 1e-7, 2500
    x1, v1 = np.array([0.0, 0.0, 0.0]), np.array([0.1, 0.0, 0.0])
    u = np.array([np.cos(0.45), np.sin(0.45), 0.0])  # oblique line of centres, 20 um apart
    xp, vp = x1 + (r + Rp + 2e-5) * u, np.array([-0.4, 0.0, 0.0])
    kw = dict(Gravity=0.0, StepSize=dt, Thresh=1.0, Frcoe=0.0, Beta=0.0, OutputInterval=steps)
    bed_a = single_sphere_clumps([x1], r, masses=[m], vel=[v1], isIntruder=1, **kw)
    pene = nbc.sphere_penetrator(Rp, tuple(xp), tuple(vp), mass=M, motion=nbc.NBC_PENE_GRAVITY)
    a = run(bed_a, steps=steps, pene=pene)
    bed_b = single_sphere_clumps([x1, xp], [r, Rp], masses=[m, M], vel=[v1, vp], **kw)
    b = run(bed_b, steps=steps)
    ca, cb, pa = a.clumps(), b.clumps(), a.penetrator()
    assert np.linalg.norm(cb["vel"][0] - v1) > 0.05, "the collision happened"
    assert np.linalg.norm(ca["vel"][0] - v1) > 0.05
    assert ca["vel"][0] == pytest.approx(cb["vel"][0], rel=1e-9, abs=1e-12)
    assert pa["vel"] == pytest.approx(cb["vel"][1], rel=1e-9, abs=1e-12)
    assert np.abs(pa["pos"] - cb["pos"][1]).max() < 1e-13
    p0 = m * v1 + M * vp
    p1 = m * ca["vel"][0] + M * pa["vel"]
    assert np.abs(p1 - p0).max() <= 1e-12 * np.abs(p0).max()


# ---------------------------------------------------------------------------
# torque-driven rotation (src/dem_integrate_clumpsphere.cu:3-43, 45-90, 110-141)

def two_clumps(offset_y, spin):
    """Two 3-sphere clumps with declared isotropic inertia, approaching
    off-centre; clump 0 spins."""
    r = 1e-3
    local = np.array([[-1.4e-3, 0, 0], [0.3e-3, 0.2e-3, 0], [1.5e-3, -0.1e-3, 0.2e-3]])
    q0 = np.array([0.1, 0.2, -0.3, 0.9]); q0 /= np.linalg.norm(q0)
    q1 = np.array([-0.4, 0.1, 0.2, 0.8]); q1 /= np.linalg.norm(q1)
    R0, R1 = quat_to_matrix(q0), quat_to_matrix(q1)
    c0 = np.array([0.0, 0.0, 0.0])
    # clump 1 placed just out of reach along x
    c1 = np.array([4.2e-3, offset_y, 0.0])
    # unequal radii, so that the lever arms Ap and An of a pair differ
    rad = np.array([r, 0.8 * r, 1.1 * r, 0.7 * r, 0.9 * r, 0.75 * r])
    s0, s1 = c0 + local @ R0.T, c1 + local @ R1.T
    gap = np.min(np.linalg.norm(s0[:, None, :] - s1[None, :, :], axis=2) - rad[:3, None] - rad[None, 3:])
    c1 = c1 - np.array([gap - 2e-5, 0, 0])
    s1 = c1 + local @ R1.T
    m, I = 3.3e-5, 4.0e-11
    p = make_params(6, 2, StepSize=1e-7, Thresh=1.0, Gravity=0.0)
    bed = nbc.Bed(p, np.array([0, 0, 0, 1, 1, 1], np.int32), np.vstack([s0, s1]), np.vstack([local, local]),
                  np.full(6, m / 3), rad, np.array([c0, c1]), np.array([[0.3, 0, 0], [-0.3, 0.05, 0]]),
                  np.array([q0, q1]), np.array([spin, [0, 0, 0]]), np.array([[I, I, I, m]] * 2), None)
    return bed, m, I


def angular_momentum(c, m, I):
    return (m * np.cross(c["pos"], c["vel"])).sum(axis=0) + I * c["angv"].sum(axis=0)


def test_angular_momentum_conserved_through_a_frictional_spinning_collision():
    """For clumps with isotropic inertia, velocity Verlet with the clump torque
    T = sum(T_s + (x_s - x_c) x F_s) gives L_{n+1} = L_n + h/2 (tau_n + tau_{n+1}),
    where tau = sum(x_s x F_s + T_s) over all spheres. For a sphere pair the
    lever arms Ap + An = |x_J - x_I| and the equal and opposite rolling and
    twisting moments make tau vanish, so L stays constant to rounding. A wrong
    alpha = T/I factor, a swapped Ap/An or a sign error in a moment breaks it."""
    bed, m, I = two_clumps(offset_y=0.9e-3, spin=[0, 0, 800.0])
    o = orc.Oracle(orc.finalize_params(bed))
    o.upload(bed, None)
    c = o.clumps()
    L0, w0 = angular_momentum(c, m, I), c["angv"].copy()
    scale = (m * np.abs(np.cross(c["pos"], c["vel"]))).sum() + I * np.abs(c["angv"]).sum()
    touched = 0
    for _ in range(30):
        o.step(100)
        touched += o.stats().touching_ss
        c = o.clumps()
        assert np.abs(angular_momentum(c, m, I) - L0).max() <= 1e-12 * scale
    assert touched > 5
    assert np.abs(c["angv"] - w0).max() > 5.0, "torques changed the spins"
    p = (m * c["vel"]).sum(axis=0)
    assert np.abs(p - m * np.array([0.0, 0.05, 0.0])).max() <= 1e-13 * m


def test_constant_torque_about_a_fixed_axis_closed_form():
    """Constant torque T = I A k along the spin axis k. With w_n = w0 + n h A
    the attitude step (half-kicked Omega = w_n + h A / 2 in qdot, alpha in
    qddot) is q_{n+1} = normalize((1 - h^2 Omega^2 / 8) q_n + (h Omega / 2 +
    h^2 A / 4) k (x) q_n): a rotation about k by 2 atan((h Omega/2 + h^2 A/4) /
    (1 - h^2 Omega^2 / 8)). Derived by hand from qdot = q (x) w_b / 2 and
    qddot = q (x) alpha_b / 2 + qdot (x) w_b / 2; the scheme is first order in
    A, with global angle error A t h / 2."""
    k = np.array([1.0, -2.0, 0.5]); k /= np.linalg.norm(k)
    q0 = np.array([0.3, -0.1, 0.2, 0.93]); q0 /= np.linalg.norm(q0)
    I, w0, A = 2e-11, 150.0, 4.0e5
    errs = []
    for h, n in ((1e-5, 400), (5e-6, 800)):
        bed = single_sphere_clumps([[0, 0, 0]], 1e-3, angv=[w0 * k], quat=[q0], inertia=I, StepSize=h, Thresh=1.0)
        o = orc.Oracle(orc.finalize_params(bed))
        o.upload(bed, None)
        o.integrate_only(n, np.zeros((1, 3)), (I * A * k)[None])
        c = o.clumps()
        assert c["angv"][0] == pytest.approx((w0 + n * h * A) * k, rel=1e-13)
        theta = 0.0
        for j in range(n):
            Om = w0 + (j + 0.5) * h * A
            theta += 2.0 * np.arctan2(h * Om / 2 + h * h * A / 4, 1.0 - h * h * Om * Om / 8)
        Rq = quat_to_matrix(c["quat"][0])
        assert np.abs(Rq - rodrigues(k, theta) @ quat_to_matrix(q0)).max() < 2e-12
        t = n * h
        errs.append(theta - (w0 * t + 0.5 * A * t * t))
        assert errs[-1] == pytest.approx(0.5 * A * t * h, rel=0.05)
    assert 1.9 < errs[0] / errs[1] < 2.1


def test_angular_acceleration_is_world_torque_over_principal_moments():
    """The reference divides the world-frame torque by the principal moments
    component by component (src/dem_integrate_clumpsphere.cu:26-28, 121-123),
    whatever the orientation."""
    prop_I = np.array([1e-11, 3e-11, 7e-11])
    T = np.array([2e-6, -1e-6, 5e-7])
    q0 = np.array([0.5, 0.1, -0.3, 0.8]); q0 /= np.linalg.norm(q0)
    h, n = 1e-6, 250
    bed = single_sphere_clumps([[0, 0, 0]], 1e-3, quat=[q0], StepSize=h, Thresh=1.0)
    bed.cprop[0, :3] = prop_I
    o = orc.Oracle(orc.finalize_params(bed))
    o.upload(bed, None)
    o.integrate_only(n, np.zeros((1, 3)), T[None])
    assert o.clumps()["angv"][0] == pytest.approx(n * h * T / prop_I, rel=1e-12)


# ---------------------------------------------------------------------------
# kinetic energy (the FP64 global sum reported in nbc_stats; no reference counterpart)

def test_kinetic_energy_closed_form_and_free_flight():
    bed = nbc.generate_bed(1, 60, seed=12)
    rng = np.random.default_rng(1)
    bed.cvel[:] = rng.normal(0, 0.05, bed.cvel.shape)
    bed.cangv[:] = rng.normal(0, 30.0, bed.cangv.shape)
    bed.params.Gravity = 0.0
    o = orc.Oracle(orc.finalize_params(bed))
    o.upload(bed, None)
    c = o.clumps()
    ke = 0.5 * (bed.cprop[:, 3] * (c["vel"] ** 2).sum(1)).sum() + 0.5 * (bed.cprop[:, :3] * c["angv"] ** 2).sum()
    assert o.stats().kinetic_energy == pytest.approx(ke, rel=1e-13)
    # force-free, torque-free flight: spread the clumps apart so nothing touches
    far = single_sphere_clumps(np.arange(8)[:, None] * [[0.01, 0, 0]], 1e-3, vel=rng.normal(0, 0.1, (8, 3)),
                               angv=rng.normal(0, 50, (8, 3)), Gravity=0.0, StepSize=1e-5, Thresh=1.0)
    o2 = orc.Oracle(orc.finalize_params(far))
    o2.upload(far, None)
    e0 = o2.stats().kinetic_energy
    o2.step(200)
    assert o2.stats().touching_ss == 0
    assert o2.stats().kinetic_energy == pytest.approx(e0, rel=1e-14)


# ---------------------------------------------------------------------------
# force-chain exports (src/dem_force_particle.cu:238-268: CacheFcn = -nf,
# CacheCtp = x_I + Ap n; front rows +Fcn, back rows -Fcn)

def test_force_chain_export_of_a_two_sphere_contact():
    r1, r2, d = 0.8e-3, 1.3e-3, 2.5e-5
    n = np.array([0.3, -0.5, 0.8]); n /= np.linalg.norm(n)
    bed = single_sphere_clumps([[0, 0, 0], (r1 + r2 - d) * n], [r1, r2])
    o = run(bed)
    pr = o.pairs()
    fcn, cpt = o.pair_exports()
    assert len(pr["fj"]) == 1
    s = o.spheres()
    I = int(np.repeat(np.arange(2), np.diff(pr["foff"]))[0])
    J = int(pr["fj"][0])
    xI, xJ, rI, rJ = s["pos"][I], s["pos"][J], s["radius"][I], s["radius"][J]
    u = (xJ - xI) / np.linalg.norm(xJ - xI)
    # normal force on I: -Hertz along I->J (static pair: no damping term)
    assert fcn[0] == pytest.approx(-hertz(r1 * r2 / (r1 + r2), d) * u, rel=1e-9)
    # the contact point lies on the line of centres and on the radical plane
    c = cpt[0]
    assert np.linalg.norm(np.cross(c - xI, u)) < 1e-15
    assert np.sum((c - xI) ** 2) - rI ** 2 == pytest.approx(np.sum((c - xJ) ** 2) - rJ ** 2, abs=1e-18)
    # and agrees with the force on I from the sphere-sphere sum
    fI = o.sphere_forces(1)[0][I]
    assert fI == pytest.approx(fcn[0], rel=1e-12)


# ---------------------------------------------------------------------------
# invalid input guard (src/dem_force_cube_wall_clump.cu:191-203, 295-297)

@pytest.mark.parametrize("bad", ["nan_pos", "zero_radius", "neg_mass", "clump_id"])
def test_invalid_sphere_input_is_refused(bad):
    centres = [[0.002 * k, 0.0, 0.005] for k in range(5)]
    wall = cube_wall((0.01, 0.01, 0.01), (0.0, 0.0, 0.01))
    bed = single_sphere_clumps(centres, 5e-4, isWallForce=1, lo=(-0.01, -0.01, 0.0), size=(0.02, 0.02, 0.02))
    k = 3
    if bad == "nan_pos":
        bed.pos[k, 1] = np.nan
        bed.cpos[k, 1] = np.nan
    elif bad == "zero_radius":
        bed.radius[k] = 0.0
    elif bad == "neg_mass":
        bed.mass[k] = -1e-6
    else:
        bed.clump_id[k] = 7
    o = orc.Oracle(orc.finalize_params(bed))
    if bad == "clump_id":
        with pytest.raises(nbc.NbcError) as e:
            o.upload(bed, wall)
        assert e.value.code == nbc.NBC_E_NONFINITE and "particle index 3" in str(e.value)
        return
    o.upload(bed, wall)
    with pytest.raises(nbc.NbcError) as e:
        o.step(1)
    assert e.value.code == nbc.NBC_E_NONFINITE
    # the reported index is the bad sphere's sorted position
    idx = int(str(e.value).split("sorted particle index ")[1].split(".")[0])
    assert o.spheres()["fid"][idx] == k


# ---------------------------------------------------------------------------
# compact slot mode: same arithmetic over the used slot range only

@pytest.mark.parametrize("driver", [nbc.NBC_DRIVER_CLUMP, nbc.NBC_DRIVER_CYLINDER])
def test_compact_mode_is_bit_identical(driver):
    if driver == nbc.NBC_DRIVER_CLUMP:
        bed = nbc.generate_bed(1, 200, seed=5)
        bed.params.Thresh = 2e-6
        bed.params.StepSize = 1e-5
        opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER)
        outs = []
        for compact in (False, True):
            o = orc.Oracle(orc.finalize_params(bed), opt, compact=compact)
            o.upload(bed, bed.wall)
            o.step(150)
            outs.append((o.clumps(), o.pair_history(), o.sphere_forces(0), o.stats().rebuilds))
    else:
        part = nbc.Particles.from_bed(nbc.generate_bed(3, 300, seed=2))
        part.vel[:] = np.random.default_rng(4).normal(0.0, 0.05, part.vel.shape)
        part.params.Thresh = 2e-6
        part.params.StepSize = 1e-5
        opt = nbc.make_options(driver=driver)
        outs = []
        for compact in (False, True):
            o = orc.Oracle(orc.finalize_params(part), opt, compact=compact)
            o.upload_particles(part, part.wall)
            o.step(150)
            outs.append((o.clumps(), o.pair_history(), o.sphere_forces(0), o.stats().rebuilds))
    (ca, ha, fa, ra), (cb, hb, fb, rb) = outs
    assert ra == rb and ra > 2
    for k in ca:
        assert np.array_equal(ca[k], cb[k])
    assert np.array_equal(ha, hb)
    assert np.array_equal(fa[0], fb[0]) and np.array_equal(fa[1], fb[1])

"""Penetrator paths on the GPU against the FP64 oracle: the triangulated
60-degree cone driven down into a deposited clumped bed (resistance force vs
depth), and the gravity-after-contact sphere."""
import numpy as np
import pytest

import nbc
import orc

pytestmark = pytest.mark.gpu


def cone_apex_clearance(bed, Rc, alpha):
    """Lowest apex height at which no sphere of the bed touches the lateral
    surface of a cone of half-angle alpha on the z axis (conservative near the
    apex): a sphere at axial distance rho, height z and radius r touches when
    rho*cos(alpha) - (z - apex)*sin(alpha) = r."""
    rho = np.hypot(bed.pos[:, 0], bed.pos[:, 1])
    m = rho < Rc + bed.radius
    return (bed.pos[m, 2] - (rho[m] * np.cos(alpha) - bed.radius[m]) / np.sin(alpha)).max()


def sphere_clearance(bed, Rp):
    """Lowest centre height of a sphere of radius Rp on the z axis that clears
    every sphere of the bed."""
    rho = np.hypot(bed.pos[:, 0], bed.pos[:, 1])
    m = rho < Rp + bed.radius
    return (bed.pos[m, 2] + np.sqrt((Rp + bed.radius[m]) ** 2 - rho[m] ** 2)).max()


def cone_case(nclumps=1500, R=0.012, steps=600):
    bed = nbc.generate_bed(1, nclumps, seed=31, container=(R, 0, 0, 0))
    bed.params.isIntruder = 1
    bed.params.OutputInterval = steps
    Rc = 0.004
    Hc = Rc / np.tan(np.radians(30.0))  # 60-degree apex
    # apex 20 um above first touch; 0.1 m/s would need 8e4 steps per 4 mm, so
    # the test drives at 2 m/s to cut into the bed within a few hundred steps
    apex = cone_apex_clearance(bed, Rc, np.radians(30.0)) + 20e-6
    pene = nbc.cone_penetrator(Rc, Hc, (0.0, 0.0, apex + Hc), (0.0, 0.0, -2.0), nseg=48)
    return bed, pene


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_cone_penetration_force_curve(prec):
    print(f"[case] cone penetrator, precision {prec}", flush=True)
    bed, pene = cone_case()
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER, pene_frcoe=0.4)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall, pene)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, bed.wall, pene)
    steps = bed.params.OutputInterval
    ctx.set_record(True)
    ctx.step(steps)
    ref.step(steps)
    fa, fb = ctx.penetrator()["cache_f"], ref.penetrator()["cache_f"]
    assert np.count_nonzero(fb[:, 2]) > 100, "cone never reached the bed"
    # reaction on the penetrator resists the downward motion
    assert fb[:, 2].sum() > 0
    scale = np.abs(fb).max()
    tol = 1e-9 if prec == nbc.NBC_FP64 else 2e-3
    assert np.abs(fa - fb).max() <= tol * scale, np.abs(fa - fb).max() / scale
    assert ctx.stats().touching_wall == ref.stats().touching_wall or prec == nbc.NBC_MIXED


def test_gravity_after_contact_sphere_penetrator():
    print("[case] gravity_after_contact sphere penetrator", flush=True)
    bed = nbc.generate_bed(1, 600, seed=2, container=(0.01, 0, 0, 0))
    bed.params.isIntruder = 1
    bed.params.OutputInterval = 400
    bed.params.StepSize = 2e-6
    Rp = 0.003
    pene = nbc.sphere_penetrator(Rp, (0.0, 0.0, sphere_clearance(bed, Rp) + 20e-6), (0.0, 0.0, -0.5), mass=2e-3,
                                 motion=nbc.NBC_PENE_GRAVITY)
    pene.wait_for_contact = 1
    opt = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall, pene)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, bed.wall, pene)
    ctx.step(400)
    ref.step(400)
    pa, pb = ctx.penetrator(), ref.penetrator()
    assert pb["state"] == 2 and pa["state"] == 2
    assert np.abs(pa["pos"] - pb["pos"]).max() <= 1e-12
    assert np.abs(pa["vel"] - pb["vel"]).max() <= 1e-9 * max(np.abs(pb["vel"]).max(), 1e-30)


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_fast_penetrator_wall_list_refresh(prec):
    """A sphere penetrator driven through the bed at 4 m/s with a small Thresh:
    it leaves its wall-list allowance (8 * Thresh) every few steps, so the
    wall list is rebuilt many times between and at neighbour-list rebuilds.
    The reaction history and the bed must follow the oracle."""
    print(f"[case] fast sphere penetrator, wall list refreshes, precision {prec}", flush=True)
    bed = nbc.generate_bed(1, 600, seed=4, container=(0.01, 0, 0, 0))
    bed.params.isIntruder = 1
    bed.params.OutputInterval = 500
    bed.params.StepSize = 1e-6
    bed.params.Thresh = 5e-6  # allowance 40 um: a new wall list every 10 steps
    Rp = 0.002
    # 1 mm above the bed: ~25 wall-list rebuilds over a resting bed, then contact
    zc = sphere_clearance(bed, Rp) + 1e-3
    pene = nbc.sphere_penetrator(Rp, (0.0, 0.0, zc), (0.0, 0.0, -4.0))
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall, pene)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, bed.wall, pene)
    ctx.step(500)
    ref.step(500)
    fa, fb = ctx.penetrator()["cache_f"], ref.penetrator()["cache_f"]
    assert np.count_nonzero(fb[:, 2]) > 150, "penetrator out of contact"
    scale = np.abs(fb).max()
    if prec == nbc.NBC_FP64:
        assert np.abs(fa - fb).max() <= 1e-9 * scale
        assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() <= 1e-12
        assert ctx.stats().touching_wall == ref.stats().touching_wall
    else:
        assert np.abs(fa - fb).max() <= 2e-3 * scale

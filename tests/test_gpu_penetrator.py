"""Penetrator paths on the GPU against the FP64 oracle: the triangulated
60-degree cone driven down into a deposited clumped bed (resistance force vs
depth), and the gravity-after-contact sphere."""
import numpy as np
import pytest

import nbc
import orc

pytestmark = pytest.mark.gpu


def cone_case(nclumps=1500, R=0.012, steps=600):
    bed = nbc.generate_bed(1, nclumps, seed=31, container=(R, 0, 0, 0))
    bed.params.isIntruder = 1
    bed.params.OutputInterval = steps
    top = bed.pos[:, 2].max() + bed.radius.max()
    Rc = 0.004
    Hc = Rc / np.tan(np.radians(30.0))  # 60-degree apex
    # apex 0.5 mm above the highest sphere; 0.1 m/s would need 8e4 steps per
    # 4 mm, so the test drives faster to reach the bed within a few hundred steps
    pene = nbc.cone_penetrator(Rc, Hc, (0.0, 0.0, top + Hc + 0.5e-3), (0.0, 0.0, -2.0), nseg=48)
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
    assert np.abs(fb[:, 2]).max() > 0, "cone never reached the bed"
    # reaction on the penetrator resists the downward motion
    assert fb[-1, 2] > 0
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
    top = bed.pos[:, 2].max() + bed.radius.max()
    Rp = 0.003
    pene = nbc.sphere_penetrator(Rp, (0.0, 0.0, top + Rp + 1e-4), (0.0, 0.0, -0.5), mass=2e-3,
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

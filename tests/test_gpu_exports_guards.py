"""GPU parity of the per-contact exports, the cylinder penetrator and the
invalid-input guard against the FP64 oracle, through the C ABI.

* Force-chain exports (src/dem_force_particle.cu:238-268): per front slot the
  normal force -nf and the contact point x_I + Ap n, after a recorded step.
  Bars: FP64 1e-12 of the largest normal force / 1e-12 m; FP32-mixed 1e-4 of
  the largest normal force and 1e-4 of the largest radius for the point
  (its arm Ap is an FP32 quantity).
* Cylinder penetrator (src/dem_contact_penetrator.cu:42-83,
  src/dem_force_penetrator_clump.cu:4-182), driven fast enough that its wall
  list is rebuilt many times: reaction history FP64 1e-9, mixed 2e-3 of the
  largest (the usual penetrator bars), per-sphere penetrator forces of the
  last step FP64 1e-12 / mixed 1e-4.
* Invalid input (src/dem_force_cube_wall_clump.cu:191-203): refused at
  upload with NBC_E_NONFINITE before any device copy; state that turns
  non-finite during a run is reported with the same sorted index as the oracle.
"""
import numpy as np
import pytest

import nbc
import orc
from util import check_mixed_history, cube_wall, single_sphere_clumps

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[case] {msg}", flush=True)


def record_last(ctx, ref, steps):
    if steps > 1:
        ctx.step(steps - 1)
        ref.step(steps - 1)
    ctx.set_record(True)
    ctx.step(1)
    ref.step(1)
    ctx.set_record(False)


def check_exports(ctx, ref, prec):
    pa, pb = ctx.pairs(), ref.pairs()
    assert np.array_equal(pa["foff"], pb["foff"]) and np.array_equal(pa["fj"], pb["fj"])
    fa, ca = ctx.pair_exports()
    fb, cb = ref.pair_exports()
    touching = np.linalg.norm(fb, axis=1) > 0
    assert touching.sum() > 20, "too few touching pairs to compare"
    # the same pairs are touching (zero rows elsewhere)
    assert np.array_equal(np.linalg.norm(fa, axis=1) > 0, touching)
    scale = np.abs(fb).max()
    rmax = ref.spheres()["radius"].max()
    ef = np.abs(fa - fb).max() / scale
    ec = np.abs(ca - cb)[touching].max()
    print(f"  exports: {touching.sum()} touching pairs, Fcn rel err {ef:.2e}, Cpt abs err {ec:.2e} m", flush=True)
    if prec == nbc.NBC_FP64:
        assert ef <= 1e-12 and ec <= 1e-12
    else:
        assert ef <= 1e-4 and ec <= 1e-4 * rmax


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_pair_exports_clump_loop(prec):
    banner(f"pair exports, clump loop, precision {prec}")
    bed = nbc.generate_bed(1, 500, seed=23)
    # settle the deposited bed on the GPU (FP32-mixed, larger step), then both
    # sides start from the downloaded contact-rich state
    dt = bed.params.StepSize
    bed.params.StepSize = 5e-6
    so = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    sctx = nbc.Context(bed.finalized_params(so), so)
    sctx.upload(bed, bed.wall)
    sctx.step(30000)
    bed = nbc.download_bed(sctx, bed)
    sctx.close()
    bed.params.StepSize = dt
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, bed.wall)
    record_last(ctx, ref, 50)
    check_exports(ctx, ref, prec)
    if prec == nbc.NBC_MIXED:
        check_mixed_history(ctx, ref, bed.params.Frcoe)


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_pair_exports_demmain_cube(prec):
    banner(f"pair exports, DEMmain cube loop, precision {prec}")
    bed = nbc.generate_bed(3, 1500, seed=9, container=(0.015, 0, 0, 0))
    part = nbc.Particles.from_bed(bed)
    R, H = bed.wall.radius, bed.wall.height
    wall = cube_wall((R, R, H), (0.0, 0.0, H))
    opt = nbc.make_options(precision=prec, driver=nbc.NBC_DRIVER_CUBE)
    ctx = nbc.Context(part.finalized_params(opt), opt)
    ctx.upload_particles(part, wall)
    ref = orc.Oracle(orc.finalize_params(part), opt)
    ref.upload_particles(part, wall)
    record_last(ctx, ref, 60)
    check_exports(ctx, ref, prec)


def cylinder_pene_case():
    bed = nbc.generate_bed(1, 600, seed=4, container=(0.01, 0, 0, 0))
    bed.params.isIntruder = 1
    bed.params.OutputInterval = 500
    bed.params.StepSize = 1e-6
    bed.params.Thresh = 5e-6  # wall-list allowance 40 um: a new wall list every ~10 steps at 4 m/s
    Rp, Hp = 0.002, 0.003
    rho = np.hypot(bed.pos[:, 0], bed.pos[:, 1])
    m = rho < Rp + bed.radius
    z_top = (bed.pos[m, 2] + bed.radius[m]).max()
    # flat bottom face 0.5 mm above the highest sphere under it, driven down at 4 m/s
    pene = nbc.cylinder_penetrator(Rp, Hp, (0.0, 0.0, z_top + Hp + 5e-4), (0.0, 0.0, -4.0))
    return bed, pene


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_cylinder_penetrator_parity(prec):
    banner(f"cylinder penetrator driven at 4 m/s, wall list refreshes, precision {prec}")
    bed, pene = cylinder_pene_case()
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall, pene)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, bed.wall, pene)
    record_last(ctx, ref, 500)
    fa, fb = ctx.penetrator()["cache_f"], ref.penetrator()["cache_f"]
    assert np.count_nonzero(fb[:, 2]) > 150, "penetrator out of contact"
    assert fb[:, 2].sum() > 0, "the reaction resists the descent"
    scale = np.abs(fb).max()
    pa, pb = ctx.sphere_forces(3)[0], ref.sphere_forces(3)[0]
    pscale = np.abs(pb).max()
    assert pscale > 0
    e_hist, e_last = np.abs(fa - fb).max() / scale, np.abs(pa - pb).max() / pscale
    print(f"  reaction history rel err {e_hist:.2e}, last-step penetrator forces rel err {e_last:.2e}, "
          f"contacts {ref.stats().touching_wall}", flush=True)
    if prec == nbc.NBC_FP64:
        assert e_hist <= 1e-9 and e_last <= 1e-12
        assert ctx.stats().touching_wall == ref.stats().touching_wall
        assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() <= 1e-12
    else:
        assert e_hist <= 2e-3 and e_last <= 1e-4
    # per-sphere penetrator exports (yi file: Fn, Ft, Tn, Tt) of the recorded step
    xa, xb = ctx.pene_exports(), ref.pene_exports()
    tol = 1e-12 if prec == nbc.NBC_FP64 else 1e-4
    for a, b in zip(xa, xb):
        assert np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


# ---------------------------------------------------------------------------
def small_bed():
    centres = [[0.002 * k - 0.004, 0.0, 0.005] for k in range(5)]
    return single_sphere_clumps(centres, 5e-4, isWallForce=1, lo=(-0.01, -0.01, 0.0), size=(0.02, 0.02, 0.02))


@pytest.mark.parametrize("bad", ["nan_pos", "inf_local", "zero_radius", "neg_mass", "nan_vel", "nan_quat",
                                 "zero_inertia", "clump_id"])
def test_invalid_input_refused_before_device_copy(bad):
    banner(f"invalid input refused at upload: {bad}")
    wall = cube_wall((0.01, 0.01, 0.01), (0.0, 0.0, 0.01))
    good = small_bed()
    opt = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CUBE)
    ctx = nbc.Context(good.finalized_params(opt), opt)
    ctx.upload(good, wall)
    before = ctx.spheres()["pos"].copy(), ctx.clumps()["vel"].copy()
    badbed = small_bed()
    k = 3
    if bad == "nan_pos":
        badbed.pos[k, 0] = np.nan
    elif bad == "inf_local":
        badbed.local[k, 2] = np.inf
    elif bad == "zero_radius":
        badbed.radius[k] = 0.0
    elif bad == "neg_mass":
        badbed.mass[k] = -1e-6
    elif bad == "nan_vel":
        badbed.cvel[k, 1] = np.nan
    elif bad == "nan_quat":
        badbed.cquat[k, 3] = np.nan
    elif bad == "zero_inertia":
        badbed.cprop[k, 1] = 0.0
    else:
        badbed.clump_id[k] = 99
    with pytest.raises(nbc.NbcError) as e:
        ctx.upload(badbed, wall)
    assert e.value.code == nbc.NBC_E_NONFINITE
    assert "FATAL: Invalid input" in str(e.value) and f"index {k}" in str(e.value)
    # nothing reached the device: the previous upload is intact and still steps
    assert np.array_equal(ctx.spheres()["pos"], before[0]) and np.array_equal(ctx.clumps()["vel"], before[1])
    ctx.step(2)
    ctx.close()


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_state_turning_nonfinite_reports_the_oracle_index(prec):
    """Two spheres of different clumps at the same centre: the contact normal
    is 0/0, the forces NaN, and the next step's clump velocities NaN. The
    wall kernel's check then stops the run; GPU and oracle name the same
    (smallest) sorted index. Every kernel still ends normally."""
    banner(f"state turning non-finite during the run, precision {prec}")
    centres = [[0.002 * k - 0.004, 0.0, 0.005] for k in range(5)] + [[0.0, 0.0, 0.005]]
    bed = single_sphere_clumps(centres, 5e-4, isWallForce=1, lo=(-0.01, -0.01, 0.0), size=(0.02, 0.02, 0.02))
    wall = cube_wall((0.01, 0.01, 0.01), (0.0, 0.0, 0.01))
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CUBE)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, wall)
    with pytest.raises(nbc.NbcError) as eo:
        ref.step(3)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, wall)
    with pytest.raises(nbc.NbcError) as eg:
        ctx.step(3)
    assert eo.value.code == eg.value.code == nbc.NBC_E_NONFINITE
    io = int(str(eo.value).split("sorted particle index ")[1].split(".")[0])
    ig = int(str(eg.value).split("sorted particle index ")[1].split(".")[0])
    print(f"  oracle index {io}, GPU index {ig}", flush=True)
    assert io == ig
    ctx.close()

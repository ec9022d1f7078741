"""GPU parity of the HIP timestep against the FP64 oracle, through the C ABI.

Bars: integer/index outputs (cell hashes, sorted order, cell bounds, pair
lists, back-slot resolution, touching sets) bit-exact; FP64 form forces and
state within 1e-12 relative to the largest value (the kernels evaluate the
reference formulas in the same order with contraction off); FP32-mixed form
forces within 1e-4 relative to the largest contact force.
"""
import numpy as np
import pytest

import nbc
import orc
from util import check_mixed_history

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-4


def banner(msg):
    print(f"[case] {msg}", flush=True)


def pair_setup(bed, opt, wall=True, pene=None):
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall if wall else None, pene)
    ref = orc.Oracle(orc.finalize_params(bed, opt.cell_size), opt)
    ref.upload(bed, bed.wall if wall else None, pene)
    return ctx, ref


def relerr(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def assert_lists_equal(ctx, ref):
    ga, gb = ctx.grid(), ref.grid()
    for k in ("cell", "perm", "beg", "end"):
        assert np.array_equal(ga[k], gb[k]), f"grid {k} differs"
    pa, pb = ctx.pairs(), ref.pairs()
    for k in ("foff", "fj", "fflag", "boff", "bi", "bslot"):
        assert np.array_equal(pa[k], pb[k]), f"pairs {k} differs"
    sa, sb = ctx.spheres(), ref.spheres()
    assert np.array_equal(sa["fid"], sb["fid"])
    assert np.array_equal(sa["cid"], sb["cid"])


@pytest.fixture(scope="module")
def c1_small():
    return nbc.generate_bed(1, 400, seed=21)


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_first_step_cylinder(c1_small, prec):
    banner(f"first step, cylinder wall, precision {prec}, N={c1_small.n}")
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx, ref = pair_setup(c1_small, opt)
    ctx.set_record(True)
    ctx.step(1)
    ref.step(1)
    assert_lists_equal(ctx, ref)
    sa, sb = ctx.stats(), ref.stats()
    assert sa.touching_ss == sb.touching_ss and sb.touching_ss > 0
    assert sa.touching_wall == sb.touching_wall and sb.touching_wall > 0
    tol = FP64_TOL if prec == nbc.NBC_FP64 else MIXED_TOL
    for kind in (1, 2, 0):
        fa, ta = ctx.sphere_forces(kind)
        fb, tb = ref.sphere_forces(kind)
        assert relerr(fa, fb) <= tol, f"force kind {kind}: {relerr(fa, fb):.3e}"
        assert relerr(ta, tb) <= tol, f"torque kind {kind}: {relerr(ta, tb):.3e}"
    ca, cb = ctx.clumps(), ref.clumps()
    assert relerr(ca["F"], cb["F"]) <= tol
    assert relerr(ca["T"], cb["T"]) <= tol
    # positions and orientations after the first drift are FP64 in both forms
    assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() == 0.0
    assert np.abs(ca["quat"] - cb["quat"]).max() == 0.0
    if prec == nbc.NBC_FP64:
        assert relerr(ctx.pair_history(), ref.pair_history()) <= tol
    else:
        check_mixed_history(ctx, ref, c1_small.params.Frcoe)



def test_fp64_bitexact_report(c1_small):
    """Records how close the FP64 form is to bit equality after one step."""
    banner("fp64 exactness report")
    opt = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx, ref = pair_setup(c1_small, opt)
    ctx.set_record(True)
    ctx.step(1)
    ref.step(1)
    fa, ta = ctx.sphere_forces(0)
    fb, tb = ref.sphere_forces(0)
    print(f"fp64 max |dF| = {np.abs(fa - fb).max():.3e}, max |dT| = {np.abs(ta - tb).max():.3e}, "
          f"bit-exact F: {np.array_equal(fa, fb)}, T: {np.array_equal(ta, tb)}", flush=True)


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_multi_step_with_rebuilds(prec):
    banner(f"200 steps with rebuilds, precision {prec}")
    bed = nbc.generate_bed(1, 300, seed=5)
    bed.params.Thresh = 2e-6  # force frequent rebuilds
    bed.params.StepSize = 1e-5  # about 20 um of free fall in 200 steps
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx, ref = pair_setup(bed, opt)
    ctx.step(200)
    ref.step(200)
    sa, sb = ctx.stats(), ref.stats()
    assert sb.rebuilds > 2
    if prec == nbc.NBC_FP64:
        assert sa.rebuilds == sb.rebuilds
        assert_lists_equal(ctx, ref)
        pa, pb = ctx.spheres()["pos"], ref.spheres()["pos"]
        assert np.abs(pa - pb).max() <= 1e-12
    else:
        # chaotic divergence is tiny over 200 steps; orders and sets still agree
        assert abs(sa.rebuilds - sb.rebuilds) <= 1
        ca, cb = ctx.clumps(), ref.clumps()
        assert np.abs(ca["pos"] - cb["pos"]).max() < 1e-9


def test_mixed_displacement_since_rebuild():
    """The FP32-mixed clump loop keeps each sphere's position at the last
    rebuild and forms the displacement as x - x_rebuild, where the oracle
    accumulates x - x_old every step. Through 80 steps with a rebuild every
    10-30 steps the rebuild count stays the oracle's after every 10 steps, and the
    reported displacement (by file index) agrees with the oracle's to 1e-9 m,
    the bar the clump positions meet over such a run."""
    banner("mixed form: displacement since the last rebuild against the oracle's accumulated one")
    bed = nbc.generate_bed(1, 300, seed=5)
    bed.params.Thresh = 5e-7  # a rebuild every 10-30 steps of the fall
    bed.params.StepSize = 1e-5
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx, ref = pair_setup(bed, opt)
    seen = 0.0
    for k in range(8):
        ctx.step(10)
        ref.step(10)
        sa, sb = ctx.stats(), ref.stats()
        assert sa.rebuilds == sb.rebuilds, k
        a, b = ctx.spheres(), ref.spheres()
        ca, cb = a["cum"][np.argsort(a["fid"])], b["cum"][np.argsort(b["fid"])]
        assert np.abs(ca - cb).max() <= 1e-9, k
        seen = max(seen, np.abs(cb).max())
    assert ref.stats().rebuilds >= 4 and seen > 1e-7


def test_cube_wall_and_sphere_penetrator():
    banner("cube wall + sphere penetrator, fp64 and mixed")
    bed = nbc.generate_bed(0, 96, seed=3, container=(0.016, 0.016, 0.024, 0))
    # push a sphere penetrator into the top layer
    k = int(np.argmax(bed.pos[:, 2]))
    x, y, z = bed.pos[k]
    Rp = 0.004
    pene = nbc.sphere_penetrator(Rp, (x, y, z + bed.radius[k] + Rp - 2e-5), (0.0, 0.0, -0.5))
    bed.params.isIntruder = 1
    bed.params.OutputInterval = 50
    for prec, tol in ((nbc.NBC_FP64, FP64_TOL), (nbc.NBC_MIXED, MIXED_TOL)):
        opt = nbc.make_options(precision=prec)
        ctx, ref = pair_setup(bed, opt, pene=pene)
        ctx.step(39)
        ref.step(39)
        ctx.set_record(True)
        ctx.step(1)
        ref.step(1)
        assert_lists_equal(ctx, ref)
        pa, pb = ctx.penetrator(), ref.penetrator()
        assert np.abs(pa["pos"] - pb["pos"]).max() == 0.0
        assert np.abs(pb["cache_f"]).max() > 0.0, "penetrator never touched the bed"
        assert relerr(pa["cache_f"], pb["cache_f"]) <= max(tol, 1e-10)
        fa, _ = ctx.sphere_forces(3)
        fb, _ = ref.sphere_forces(3)
        assert relerr(fa, fb) <= tol
        ea, eb = ctx.pene_exports(), ref.pene_exports()
        for x, y in zip(ea, eb):
            assert relerr(x, y) <= max(tol, 1e-10)
        sa, sb = ctx.stats(), ref.stats()
        assert sa.touching_wall == sb.touching_wall


def test_grid_rebuild_large():
    banner("grid + list rebuild on 40k clumps (bit-exact)")
    bed = nbc.generate_bed(1, 40000, seed=9)
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx, ref = pair_setup(bed, opt)
    ctx.rebuild()
    ref.rebuild()
    assert_lists_equal(ctx, ref)


def test_sqrt_div_rounding_on_device_matches_host():
    """The FP64 touching predicate relies on correctly rounded sqrt/div on the
    device; a one-step run of a dense bed exercises both, compared bitwise."""
    banner("fp64 pair lists on a dense bed")
    bed = nbc.generate_bed(1, 2000, seed=13)
    opt = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx, ref = pair_setup(bed, opt)
    ctx.set_record(True)
    ctx.step(3)
    ref.step(3)
    assert_lists_equal(ctx, ref)
    assert ctx.stats().touching_ss == ref.stats().touching_ss


@pytest.mark.parametrize("horizon", [-1.0, 0.0, 20.0])
@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_fast_collisions_between_rebuilds(prec, horizon):
    """Clumps thrown at each other at up to 0.5 m/s: pairs close from outside
    the contact kernel's near range (gap 2 * Thresh at the last rebuild) and
    collide between rebuilds. Touching sets, histories and states must follow
    the oracle, which scans every list pair every step, with the fixed near
    skin (horizon -1), the per-sphere skin of the default 500-step horizon
    (0) and a 20-step horizon that asks for many near-list refreshes."""
    banner(f"fast collisions between rebuilds, precision {prec}, near horizon {horizon}")
    bed = nbc.generate_bed(1, 300, seed=8)
    rng = np.random.default_rng(8)
    bed.cvel[:] = rng.uniform(-0.5, 0.5, bed.cvel.shape)
    bed.cangv[:] = rng.uniform(-50.0, 50.0, bed.cangv.shape)
    bed.params.StepSize = 1e-5  # up to 5 um per step: a rebuild every ~50 steps
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER, near_horizon=horizon)
    ctx, ref = pair_setup(bed, opt)
    seen = 0
    for k in range(20):
        ctx.step(20)
        ref.step(20)
        sa, sb = ctx.stats(), ref.stats()
        seen = max(seen, sb.touching_ss)
        if prec == nbc.NBC_FP64:
            assert (sa.touching_ss, sa.touching_wall, sa.rebuilds) == (sb.touching_ss, sb.touching_wall, sb.rebuilds), k
    sb = ref.stats()
    assert sb.rebuilds >= 4 and seen > 0
    # the contact kernel's near lists were refreshed between rebuilds too
    if horizon != 0.0:
        assert ctx.stats().near_refreshes > 0
    if prec == nbc.NBC_FP64:
        assert_lists_equal(ctx, ref)
        assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() <= 1e-12
        ha, hb = ctx.pair_history(), ref.pair_history()
        assert relerr(ha, hb) <= FP64_TOL
    else:
        ca, cb = ctx.clumps(), ref.clumps()
        assert np.abs(ca["pos"] - cb["pos"]).max() < 1e-7

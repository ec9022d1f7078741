"""GPU parity of the hierarchical grid (3 levels, polydisperse clumped bed,
radii 0.2-2.0 mm) against the FP64 oracle, through the C ABI.

Bars as in the uniform-grid parity tests: cell hashes, sorted order, cell
bounds, pair rows (same-level partners, then cross-level partners in
ascending index), back-slot resolution and touching sets bit-exact; FP64
forces and state within 1e-12; FP32-mixed forces within 1e-4 relative to the
largest contact force.
"""
import numpy as np
import pytest

import nbc
import orc

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-4


def banner(msg):
    print(f"[case] {msg}", flush=True)


def setup(bed, prec, levels=3, nested=False):
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CUBE, hgrid_levels=levels, hgrid_nested=int(nested))
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    ref = orc.Oracle(orc.finalize_params(bed, 0.0, levels, nested=nested), opt)
    ref.upload(bed, bed.wall)
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
    assert np.array_equal(ctx.spheres()["fid"], ref.spheres()["fid"])


@pytest.fixture(scope="module")
def c2_small():
    return nbc.generate_bed(2, 1500, seed=5, container=(0.025, 0, 0, 0))


@pytest.mark.parametrize("nested", [False, True])
@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_hgrid_first_step(c2_small, prec, nested):
    banner(f"hgrid first step, precision {prec}, nested {nested}, N={c2_small.n}")
    ctx, ref = setup(c2_small, prec, nested=nested)
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
    assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() == 0.0
    ctx.close()
    ref.close()


@pytest.mark.parametrize("nested", [False, True])
@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_hgrid_steps_with_rebuilds(c2_small, prec, nested):
    """Random clump velocities up to 0.5 m/s at dt 2e-6 s: a rebuild every few
    tens of steps, so cross-level histories are carried over many times (with
    nested levels, also those of coarse owners that sort after their
    partner)."""
    banner(f"hgrid 400 steps with rebuilds, precision {prec}, nested {nested}")
    bed = nbc.generate_bed(2, 1500, seed=5, container=(0.025, 0, 0, 0))
    rng = np.random.default_rng(5)
    bed.cvel[:] = rng.uniform(-0.5, 0.5, bed.cvel.shape)
    bed.cangv[:] = rng.uniform(-100.0, 100.0, bed.cangv.shape)
    bed.params.StepSize = 2e-6
    ctx, ref = setup(bed, prec, nested=nested)
    seen = 0
    for k in range(20):
        ctx.step(20)
        ref.step(20)
        sa, sb = ctx.stats(), ref.stats()
        seen = max(seen, sb.touching_ss)
        if prec == nbc.NBC_FP64:
            assert (sa.touching_ss, sa.touching_wall, sa.rebuilds) == (sb.touching_ss, sb.touching_wall, sb.rebuilds), k
    sb = ref.stats()
    assert sb.rebuilds >= 3 and seen > 0
    if prec == nbc.NBC_FP64:
        assert_lists_equal(ctx, ref)
        assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() <= 1e-12
        assert relerr(ctx.pair_history(), ref.pair_history()) <= FP64_TOL
    else:
        ca, cb = ctx.clumps(), ref.clumps()
        assert np.abs(ca["pos"] - cb["pos"]).max() < 1e-7
    ctx.close()
    ref.close()


def test_hgrid_overflow_stops_before_use(c2_small):
    banner("hgrid overflow, BranchLmtLen 8")
    bed = nbc.generate_bed(2, 1500, seed=5, container=(0.025, 0, 0, 0))
    bed.params.BranchLmtLen = 8
    ctx, ref = setup(bed, nbc.NBC_MIXED)
    for o in (ctx, ref):
        with pytest.raises(nbc.NbcError) as e:
            o.step(1)
        assert e.value.code == nbc.NBC_E_NEIGHBOR_OVERFLOW
    ctx.close()
    ref.close()

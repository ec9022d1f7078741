"""GPU parity of a batch of independent beds (one context, one launch set per
step) against one FP64 oracle per bed, through the C ABI.

A rebuild flagged by any bed rebuilds every bed, so each bed's oracle is told
to rebuild at the steps where the batch did. Bars per bed: sorted order, cell
bounds, pair rows, back slots bit-exact (indices taken relative to the bed's
range); FP64 state within 1e-12; FP32-mixed forces within 1e-4 relative.
"""
import numpy as np
import pytest

import nbc
import orc

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[case] {msg}", flush=True)


def make_beds(speed=0.0):
    beds = []
    for nc, seed, R in ((300, 11, 0.012), (180, 12, 0.009), (420, 13, 0.015)):
        b = nbc.generate_bed(1, nc, seed=seed, container=(R, 0, 0, 0))
        if speed:
            rng = np.random.default_rng(seed)
            b.cvel[:] = rng.uniform(-speed, speed, b.cvel.shape)
            b.cangv[:] = rng.uniform(-50.0, 50.0, b.cangv.shape)
            b.params.StepSize = 1e-5
        beds.append(b)
    return beds


def setup(beds, prec):
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.BatchContext([b.finalized_params(opt) for b in beds], opt)
    ctx.upload(beds)
    refs = []
    for b in beds:
        r = orc.Oracle(orc.finalize_params(b), opt)
        r.upload(b, b.wall)
        refs.append(r)
    return ctx, refs


def step_together(ctx, refs, n):
    """Step the batch and the per-bed oracles n steps, one at a time, forcing
    the oracles to rebuild where the batch rebuilt."""
    for _ in range(n):
        r0 = ctx.stats().rebuilds
        ctx.step(1)
        if ctx.stats().rebuilds > r0:
            for r in refs:
                r.request_rebuild()
        for r in refs:
            r.step(1)


def check_bed(ctx, ref, b, so, co, cell_off, fp64, all_grid, all_pairs, all_sph, lists=True):
    s0, s1 = so[b], so[b + 1]
    n = s1 - s0
    sb = ref.spheres()
    # positions by file index (the FP32-mixed form may rebuild, and so sort,
    # at other steps than the FP64 oracle)
    xa, xb = np.zeros((n, 3)), np.zeros((n, 3))
    xa[all_sph["fid"][s0:s1] - s0] = all_sph["pos"][s0:s1]
    xb[sb["fid"]] = sb["pos"]
    d = np.abs(xa - xb).max()
    assert d <= (1e-12 if fp64 else 1e-7), f"bed {b} positions differ by {d:.3e}"
    if not lists:
        return
    gb = ref.grid()
    cn = len(gb["beg"])
    assert np.array_equal(all_grid["cell"][s0:s1] - cell_off, gb["cell"]), f"bed {b} cells"
    assert np.array_equal(all_grid["perm"][s0:s1] - s0, gb["perm"]), f"bed {b} order"
    for k in ("beg", "end"):
        a = all_grid[k][cell_off:cell_off + cn].copy()
        a[a >= 0] -= s0
        assert np.array_equal(a, gb[k]), f"bed {b} cell {k}"
    pa, pb = all_pairs, ref.pairs()
    f0, f1 = pa["foff"][s0], pa["foff"][s1]
    assert np.array_equal(pa["foff"][s0:s1 + 1] - f0, pb["foff"]), f"bed {b} foff"
    assert np.array_equal(pa["fj"][f0:f1] - s0, pb["fj"]), f"bed {b} fj"
    assert np.array_equal(pa["fflag"][f0:f1], pb["fflag"]), f"bed {b} flags"
    g0, g1 = pa["boff"][s0], pa["boff"][s1]
    assert np.array_equal(pa["boff"][s0:s1 + 1] - g0, pb["boff"]), f"bed {b} boff"
    assert np.array_equal(pa["bi"][g0:g1] - s0, pb["bi"]), f"bed {b} bi"
    bs = pa["bslot"][g0:g1].copy()
    bs[bs >= 0] -= f0
    assert np.array_equal(bs, pb["bslot"]), f"bed {b} back slots"
    assert np.array_equal(all_sph["fid"][s0:s1] - s0, sb["fid"])
    assert np.array_equal(all_sph["cid"][s0:s1] - co[b], sb["cid"])


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_batch_first_step_matches_each_bed(prec):
    beds = make_beds()
    banner(f"batch of {len(beds)} beds, first step, precision {prec}, N={sum(b.n for b in beds)}")
    ctx, refs = setup(beds, prec)
    ctx.set_record(True)
    step_together(ctx, refs, 1)
    so, co = ctx.ranges()
    assert list(so) == [0] + list(np.cumsum([b.n for b in beds]))
    g, p, s = ctx.grid(), ctx.pairs(), ctx.spheres()
    cell_off = 0
    for b, r in enumerate(refs):
        check_bed(ctx, r, b, so, co, cell_off, prec == nbc.NBC_FP64, g, p, s)
        cell_off += r.params.CellNum
    st = ctx.stats()
    assert st.touching_ss == sum(r.stats().touching_ss for r in refs) > 0
    assert st.touching_wall == sum(r.stats().touching_wall for r in refs) > 0
    tol = 1e-12 if prec == nbc.NBC_FP64 else 1e-4
    for kind in (1, 2, 0):
        fa, ta = ctx.sphere_forces(kind)
        for b, r in enumerate(refs):
            fb, tb = r.sphere_forces(kind)
            scale_f, scale_t = max(np.abs(fb).max(), 1e-300), max(np.abs(tb).max(), 1e-300)
            assert np.abs(fa[so[b]:so[b + 1]] - fb).max() <= tol * scale_f, f"bed {b} force kind {kind}"
            assert np.abs(ta[so[b]:so[b + 1]] - tb).max() <= tol * scale_t, f"bed {b} torque kind {kind}"
    ctx.close()
    for r in refs:
        r.close()


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_batch_steps_with_rebuilds(prec):
    beds = make_beds(speed=0.5)
    banner(f"batch of {len(beds)} beds, 150 steps with rebuilds, precision {prec}")
    ctx, refs = setup(beds, prec)
    step_together(ctx, refs, 150)
    st = ctx.stats()
    assert st.rebuilds >= 3
    if prec == nbc.NBC_FP64:
        assert st.touching_ss == sum(r.stats().touching_ss for r in refs)
        assert st.touching_wall == sum(r.stats().touching_wall for r in refs)
    so, co = ctx.ranges()
    g, p, s = ctx.grid(), ctx.pairs(), ctx.spheres()
    cell_off = 0
    for b, r in enumerate(refs):
        check_bed(ctx, r, b, so, co, cell_off, prec == nbc.NBC_FP64, g, p, s, lists=prec == nbc.NBC_FP64)
        cell_off += r.params.CellNum
    cl = ctx.clumps()
    for b, r in enumerate(refs):
        d = np.abs(cl["pos"][co[b]:co[b + 1]] - r.clumps()["pos"]).max()
        assert d <= (1e-12 if prec == nbc.NBC_FP64 else 1e-7)
    ctx.close()
    for r in refs:
        r.close()


def test_batch_rejects_mixed_materials():
    beds = make_beds()
    beds[1].params.Frcoe = 0.3
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    with pytest.raises(nbc.NbcError) as e:
        nbc.BatchContext([b.finalized_params(opt) for b in beds], opt)
    assert e.value.code == nbc.NBC_E_BADINPUT

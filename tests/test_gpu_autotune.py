"""Autotuner on the GPU: it measures every candidate, and the options it picks
without NBC_TUNE_CELL / NBC_TUNE_HGRID leave every FP64 result bit-identical
to the default options (near skin, chunk size, radix bits, block size and
graph replays change no arithmetic); the batch autotuner likewise."""
import numpy as np
import pytest

import nbc

pytestmark = pytest.mark.gpu


def test_autotune_picks_result_preserving_options():
    print("[case] autotune on a 600-clump C1 bed", flush=True)
    bed = nbc.generate_bed(1, 600, seed=17)
    rng = np.random.default_rng(17)
    bed.cvel[:] = rng.uniform(-0.3, 0.3, bed.cvel.shape)
    bed.params.StepSize = 5e-6
    base = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CYLINDER)
    best, entries = nbc.autotune(bed, base, wall=bed.wall, steps=60)
    assert len(entries) == 5 + 4 + 3 + 2 + 2
    assert all(ms > 0 for _, ms in entries)
    # a sweep keeps the incoming value (0: default) unless a candidate beats it
    assert best.near_skin in (0.0, 0.125, 0.25, 0.5, 1.0) and best.batch_steps in (0, 2, 4, 8, 16)
    assert best.cell_size == 0.0 and best.hgrid_levels == 0
    out = []
    for opt in (base, best):
        ctx = nbc.Context(bed.finalized_params(opt), opt)
        ctx.upload(bed, bed.wall)
        ctx.step(150)
        st = ctx.stats()
        out.append((ctx.spheres()["pos"], ctx.clumps()["vel"], st.rebuilds, st.touching_ss))
        ctx.close()
    (pa, va, ra, ta), (pb, vb, rb, tb) = out
    assert ra == rb and ta == tb and ra >= 2
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)


def test_autotune_cell_and_hgrid_sweeps():
    print("[case] autotune with cell-size and hierarchical-grid sweeps, C2-style bed", flush=True)
    bed = nbc.generate_bed(2, 800, seed=6, container=(0.02, 0, 0, 0))
    base = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CUBE)
    best, entries = nbc.autotune(bed, base, wall=bed.wall, steps=40, flags=nbc.NBC_TUNE_CELL | nbc.NBC_TUNE_HGRID)
    assert len(entries) == 5 + 4 + 3 + 2 + 2 + 3 + 5
    assert all(ms > 0 for _, ms in entries)
    assert best.hgrid_levels in (0, 2, 3) and best.hgrid_nested in (0, 1)
    # the picked options run
    ctx = nbc.Context(bed.finalized_params(best), best)
    ctx.upload(bed, bed.wall)
    ctx.step(20)
    assert ctx.stats().touching_ss > 0
    ctx.close()


def test_autotune_batch_picks_result_preserving_options():
    print("[case] batch autotune on three C1-style beds", flush=True)
    beds = [nbc.generate_bed(1, nc, seed=sd, container=(R, 0, 0, 0))
            for nc, sd, R in ((300, 31, 0.012), (200, 32, 0.01), (400, 33, 0.015))]
    rng = np.random.default_rng(3)
    for b in beds:
        b.cvel[:] = rng.uniform(-0.3, 0.3, b.cvel.shape)
        b.params.StepSize = 5e-6
    base = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CYLINDER)
    best, entries = nbc.autotune_batch(beds, base, steps=40)
    assert len(entries) == 5 + 4 + 3 + 2 + 2
    assert all(ms > 0 for _, ms in entries)
    assert best.cell_size == 0.0 and best.hgrid_levels == 0
    out = []
    for opt in (base, best):
        ctx = nbc.BatchContext([b.finalized_params(opt) for b in beds], opt)
        ctx.upload(beds)
        ctx.step(150)
        st = ctx.stats()
        out.append((ctx.spheres()["pos"], ctx.clumps()["vel"], st.rebuilds, st.touching_ss))
        ctx.close()
    (pa, va, ra, ta), (pb, vb, rb, tb) = out
    assert ra == rb and ta == tb and ra >= 2
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)

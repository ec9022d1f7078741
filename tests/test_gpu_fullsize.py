"""Parity at full size: the headline bed (200,000 clumps from 8 templates of
2-8 spheres, 1,124,673 spheres, lognormal radii, cylinder R 60 mm), settled
on the GPU until it is contact-rich, and the polydisperse C2 bed (1,929,826
spheres in 571,429 clumps, radii 0.2-2 mm, cube, nested 3-level grid) as
deposited; one step of the FP64 and the FP32-mixed forms against the FP64
oracle from the identical state.

This is where the at-scale paths are exercised: int32 CSR row totals of
~13 M list pairs, partner indices above 2^20 in the 29-bit pair word, history
offsets past 2^24 entries, the radix sort over 1.1 M keys and the wall list.

Bars: sorted order, cell bounds, front/back rows, back slots and touching
counts bit-exact; per-sphere total force and torque
(src/dem_force_clump.cu:165-191 + walls) and clump force and torque
(src/dem_clumpsphere.cu:110-151) within 1e-12 (FP64) / 1e-4 (mixed) of the
largest value; clump state after the step FP64 1e-12.

The oracle runs in its compact slot mode (bit-identical to the faithful
slot-padded mode, tests/test_oracle_pins_more.py) so one step takes seconds.
"""
import numpy as np
import pytest

import nbc
import orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def settled_c1():
    print("[case] full-size C1 bed: generate and settle 20000 steps of 5 us on the GPU", flush=True)
    bed = nbc.generate_bed(1, 200000, seed=1)
    bed.params.BranchLmtLen = 128
    dt = bed.params.StepSize
    bed.params.StepSize = 5e-6
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    for _ in range(4):
        ctx.step(5000)
        s = ctx.stats()
        print(f"  settle step {s.step}: touching {s.touching_ss / bed.n:.3f}/sphere", flush=True)
    out = nbc.download_bed(ctx, bed)
    ctx.close()
    out.params.StepSize = dt
    return out


@pytest.fixture(scope="module")
def oracle_step(settled_c1):
    bed = settled_c1
    orc.set_threads(16)
    opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER)
    ref = orc.Oracle(orc.finalize_params(bed), opt, compact=True)
    ref.upload(bed, bed.wall)
    ref.step(1)
    out = dict(grid=ref.grid(), pairs=ref.pairs(), stats=ref.stats(), forces=ref.sphere_forces(0),
               clumps=ref.clumps(), spheres=ref.spheres())
    ref.close()
    print(f"[case] oracle step done: {out['stats'].touching_ss} touching pairs, {out['stats'].front_pairs} list pairs",
          flush=True)
    return out


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_full_size_first_step(settled_c1, oracle_step, prec):
    bed, o = settled_c1, oracle_step
    print(f"[case] full-size C1 first step, precision {prec}, N={bed.n}", flush=True)
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    ctx.set_record(True)
    ctx.step(1)
    ga = ctx.grid()
    for k in ("cell", "perm", "beg", "end"):
        assert np.array_equal(ga[k], o["grid"][k]), f"grid {k} differs"
    pa = ctx.pairs()
    for k in ("foff", "fj", "fflag", "boff", "bi", "bslot"):
        assert np.array_equal(pa[k], o["pairs"][k]), f"pairs {k} differs"
    assert np.array_equal(ctx.spheres()["fid"], o["spheres"]["fid"])
    sa, sb = ctx.stats(), o["stats"]
    assert sa.touching_ss == sb.touching_ss and sa.touching_wall == sb.touching_wall
    assert sb.touching_ss > 0.05 * bed.n
    tol = 1e-12 if prec == nbc.NBC_FP64 else 1e-4
    fa, ta = ctx.sphere_forces(0)
    fb, tb = o["forces"]
    ca, cb = ctx.clumps(), o["clumps"]
    errs = dict(sphere_F=relerr(fa, fb), sphere_T=relerr(ta, tb), clump_F=relerr(ca["F"], cb["F"]),
                clump_T=relerr(ca["T"], cb["T"]), clump_vel=relerr(ca["vel"], cb["vel"]),
                clump_angv=relerr(ca["angv"], cb["angv"]))
    print("  " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k in ("sphere_F", "sphere_T", "clump_F", "clump_T"):
        assert errs[k] <= tol, k
    # the drift before the contacts is FP64 in both forms
    assert np.abs(ca["pos"] - cb["pos"]).max() <= 1e-15
    if prec == nbc.NBC_FP64:
        assert errs["clump_vel"] <= 1e-12 and errs["clump_angv"] <= 1e-12
    ctx.close()


@pytest.fixture(scope="module")
def c2_full():
    print("[case] full-size C2 bed (1.93 M spheres, 571,429 clumps of 1-6, radii 0.2-2 mm)", flush=True)
    return nbc.generate_bed(2, 0, seed=1)


@pytest.fixture(scope="module")
def c2_oracle_step(c2_full):
    bed = c2_full
    orc.set_threads(16)
    opt = nbc.make_options(wall_type=nbc.NBC_WALL_CUBE, hgrid_levels=3, hgrid_nested=1)
    ref = orc.Oracle(orc.finalize_params(bed, 0.0, 3, nested=True), opt, compact=True)
    ref.upload(bed, bed.wall)
    ref.step(1)
    out = dict(grid=ref.grid(), pairs=ref.pairs(), stats=ref.stats(), forces=ref.sphere_forces(0),
               clumps=ref.clumps(), spheres=ref.spheres())
    ref.close()
    print(f"[case] C2 oracle step done: {out['stats'].touching_ss} touching pairs, {out['stats'].front_pairs} list "
          "pairs", flush=True)
    return out


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_c2_full_size_first_step(c2_full, c2_oracle_step, prec):
    """The C2 bed at full size on the nested 3-level grid (the configuration
    the C2 generator writes): 21-bit cell keys over 1.44 M cells, cross-level
    entries at scale, rows up to BranchLmtLen / 2 = 128 per side."""
    bed, o = c2_full, c2_oracle_step
    print(f"[case] full-size C2 first step, precision {prec}, N={bed.n}", flush=True)
    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CUBE, hgrid_levels=3, hgrid_nested=1)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    ctx.set_record(True)
    ctx.step(1)
    ga = ctx.grid()
    for k in ("cell", "perm", "beg", "end"):
        assert np.array_equal(ga[k], o["grid"][k]), f"grid {k} differs"
    pa = ctx.pairs()
    for k in ("foff", "fj", "fflag", "boff", "bi", "bslot"):
        assert np.array_equal(pa[k], o["pairs"][k]), f"pairs {k} differs"
    sa, sb = ctx.stats(), o["stats"]
    assert sa.touching_ss == sb.touching_ss and sa.touching_wall == sb.touching_wall
    assert sb.touching_ss > 0.05 * bed.n
    tol = 1e-12 if prec == nbc.NBC_FP64 else 1e-4
    fa, ta = ctx.sphere_forces(0)
    fb, tb = o["forces"]
    ca, cb = ctx.clumps(), o["clumps"]
    errs = dict(sphere_F=relerr(fa, fb), sphere_T=relerr(ta, tb), clump_F=relerr(ca["F"], cb["F"]),
                clump_T=relerr(ca["T"], cb["T"]))
    print("  " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= tol, k
    assert np.abs(ca["pos"] - cb["pos"]).max() <= 1e-15
    ctx.close()

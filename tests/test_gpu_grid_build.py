"""nbc_grid_build (DEMGrid::NeighborSearch on device buffers,
src/dem_grid.cu:649-678) against the FP64 oracle's grid on the same
positions: cell keys, sorted order and cell bounds bit-exact; at the full
C1 size against a context's rebuild (itself bit-exact to the oracle) and the
sort's invariants."""
import ctypes as C

import numpy as np
import pytest

import nbc
import orc
from hiprt import DeviceArray

pytestmark = pytest.mark.gpu


def pos4(bed):
    return np.ascontiguousarray(np.column_stack([bed.pos, bed.radius]))


def grid_build(params, p4, fn="nbc_grid_build"):
    n = len(p4)
    d_pos = DeviceArray.from_host(p4)
    out = [DeviceArray((n,), np.uint32), DeviceArray((n,), np.int32), DeviceArray((params.CellNum,), np.int32),
           DeviceArray((params.CellNum,), np.int32)]
    rc = getattr(nbc.lib(), fn)(C.byref(params), d_pos.ptr, n, *[o.ptr for o in out], None)
    if rc:
        raise nbc.NbcError(rc, nbc.lib().nbc_last_error(None).decode())
    res = dict(cell=out[0].host().astype(np.int32), perm=out[1].host(), beg=out[2].host(), end=out[3].host())
    for d in [d_pos] + out:
        d.free()
    return res


def oracle_grid(bed, params, levels=0, nested=False):
    o = orc.Oracle(orc.finalize_params(bed, 0.0, levels, nested=nested))
    o.upload(bed, None)
    o.rebuild()
    g = o.grid()
    o.close()
    return g


def check_sort_invariants(g, n, cellnum):
    cell, perm = g["cell"], g["perm"]
    assert np.all(np.diff(cell) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(n))
    same = cell[1:] == cell[:-1]
    assert np.all(perm[1:][same] > perm[:-1][same]), "stable within a cell"
    counts = np.bincount(cell, minlength=cellnum)
    occ = counts > 0
    assert np.array_equal(g["end"][occ] - g["beg"][occ], counts[occ])
    assert np.all(g["beg"][~occ] == -1) and np.all(g["end"][~occ] == -1)


@pytest.mark.parametrize("kind", ["uniform", "hgrid", "hgrid-nested"])
def test_grid_build_matches_oracle(kind):
    print(f"[case] nbc_grid_build vs oracle, {kind}", flush=True)
    if kind == "uniform":
        bed = nbc.generate_bed(1, 600, seed=21)
        opt, levels = nbc.make_options(), 0
    else:
        bed = nbc.generate_bed(2, 2000, seed=3)
        opt, levels = nbc.make_options(hgrid_levels=3, hgrid_nested=int(kind == "hgrid-nested")), 3
    params = bed.finalized_params(opt)
    g = grid_build(params, pos4(bed))
    r = oracle_grid(bed, params, levels, nested=kind == "hgrid-nested")
    for k in ("cell", "perm", "beg", "end"):
        assert np.array_equal(g[k], r[k]), k
    check_sort_invariants(g, bed.n, params.CellNum)


@pytest.mark.timeout(600)
def test_grid_build_full_size_matches_context_rebuild():
    print("[case] nbc_grid_build at 1.12M spheres vs a context rebuild", flush=True)
    bed = nbc.generate_bed(1, 200000, seed=1)
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    params = bed.finalized_params(opt)
    g = grid_build(params, pos4(bed))
    check_sort_invariants(g, bed.n, params.CellNum)
    ctx = nbc.Context(params, opt)
    ctx.upload(bed, bed.wall)
    ctx.rebuild()
    r = ctx.grid()
    ctx.close()
    for k in ("cell", "perm", "beg", "end"):
        assert np.array_equal(g[k], r[k]), k


def test_grid_build_rejects_bad_arguments():
    print("[case] nbc_grid_build argument checks", flush=True)
    bed = nbc.generate_bed(1, 50, seed=2)
    params = bed.finalized_params(nbc.make_options())
    p4 = pos4(bed)
    d = [DeviceArray((bed.n,), np.uint32), DeviceArray((bed.n,), np.int32), DeviceArray((params.CellNum,), np.int32),
         DeviceArray((params.CellNum,), np.int32)]
    rc = nbc.lib().nbc_grid_build(C.byref(params), p4.ctypes.data_as(C.c_void_p), bed.n, *[x.ptr for x in d], None)
    assert rc == nbc.NBC_E_BADINPUT and "device memory" in nbc.lib().nbc_last_error(None).decode()
    raw = nbc.Params()
    C.memmove(C.byref(raw), C.byref(bed.params), C.sizeof(nbc.Params))
    dp = DeviceArray.from_host(p4)
    rc = nbc.lib().nbc_grid_build(C.byref(raw), dp.ptr, bed.n, *[x.ptr for x in d], None)
    assert rc == nbc.NBC_E_BADINPUT and "finalize" in nbc.lib().nbc_last_error(None).decode()
    rc = nbc.lib().nbc_grid_build(C.byref(params), dp.ptr, 0, *[x.ptr for x in d], None)
    assert rc == nbc.NBC_E_BADINPUT
    for x in d + [dp]:
        x.free()

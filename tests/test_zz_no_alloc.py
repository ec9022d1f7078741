"""nbc_grid_build_noalloc (the grid build without any device allocation)
against nbc_grid_build, bit for bit, on two calls in a row, the second with
a smaller bed, so that whatever the first call left in the shared scratch
array cannot leak into the second; plus the size limit."""
import ctypes as C

import numpy as np
import pytest

import nbc
from hiprt import DeviceArray

pytestmark = pytest.mark.gpu


def run(fn, params, p4):
    n = len(p4)
    d_pos = DeviceArray.from_host(p4)
    out = [DeviceArray((n,), np.uint32), DeviceArray((n,), np.int32), DeviceArray((params.CellNum,), np.int32),
           DeviceArray((params.CellNum,), np.int32)]
    rc = getattr(nbc.lib(), fn)(C.byref(params), d_pos.ptr, n, *[o.ptr for o in out], None)
    assert rc == 0, nbc.lib().nbc_last_error(None).decode()
    res = [o.host() for o in out]
    for d in [d_pos] + out:
        d.free()
    return res


def bed_pos4(nclumps, seed):
    bed = nbc.generate_bed(1, nclumps, seed=seed)
    return bed, np.ascontiguousarray(np.column_stack([bed.pos, bed.radius]))


@pytest.mark.timeout(600)
def test_noalloc_matches_allocating_build_on_two_calls():
    # first call: about 1.12M spheres (the headline bed), then a 60k-sphere bed
    for nclumps, seed in ((200000, 1), (11000, 5)):
        bed, p4 = bed_pos4(nclumps, seed)
        print(f"[case] grid build no-alloc vs allocating, n={bed.n}", flush=True)
        params = bed.finalized_params(nbc.make_options())
        a = run("nbc_grid_build", params, p4)
        b = run("nbc_grid_build_noalloc", params, p4)
        for x, y, name in zip(a, b, ("cell", "perm", "beg", "end")):
            assert np.array_equal(x, y), name
        assert np.all(np.diff(a[0].astype(np.int64)) >= 0)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    print("[case] grid build no-alloc size limit", flush=True)
    bed, p4 = bed_pos4(50, 2)
    params = bed.finalized_params(nbc.make_options())
    n = 2097152 + 1
    d = [DeviceArray((1,), np.float64), DeviceArray((1,), np.uint32), DeviceArray((1,), np.int32),
         DeviceArray((1,), np.int32), DeviceArray((1,), np.int32)]
    # rejected before any buffer is touched
    rc = nbc.lib().nbc_grid_build_noalloc(C.byref(params), *[x.ptr for x in d[:1]], n, *[x.ptr for x in d[1:]], None)
    assert rc == nbc.NBC_E_BADINPUT
    assert "exceed the scratch limit" in nbc.lib().nbc_last_error(None).decode()
    for x in d:
        x.free()

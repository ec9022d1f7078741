"""Pins of the hierarchical grid (level table, level-major cell hash, pair
rows) in the FP64 oracle and the library's parameter set-up.

Independent statements used: a KD-tree (scipy) for the 1.5 (ri + rj) pair set,
the level rule r <= cell / 3 that makes a 27-cell search complete, numpy for
the hash formula, and the uniform grid's own pair set for the same positions.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial import cKDTree

import nbc
import orc
from util import make_params


@pytest.fixture(scope="module")
def c2_small():
    return nbc.generate_bed(2, 1500, seed=5, container=(0.025, 0, 0, 0))


def opt3():
    return nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CUBE, hgrid_levels=3)


def kd_pairs(pos, rad, mask=None):
    """All i < j with |pi - pj| < 1.5 (ri + rj), by KD-tree."""
    tree = cKDTree(pos)
    cand = tree.query_pairs(3.0 * rad.max() * (1 + 1e-9), output_type="ndarray")
    i, j = cand[:, 0], cand[:, 1]
    d2 = ((pos[i] - pos[j]) ** 2).sum(1)
    lim = 1.5 * (rad[i] + rad[j])
    keep = d2 < lim * lim
    if mask is not None:
        keep &= mask[i] & mask[j]
    a, b = np.minimum(i, j)[keep], np.maximum(i, j)[keep]
    return set(zip(a.tolist(), b.tolist()))


def test_c2_generator_radii_and_box(c2_small):
    b = c2_small
    assert b.radius.min() >= 0.2e-3 * (1 - 1e-6) and b.radius.max() <= 2.0e-3 * (1 + 1e-6)
    # log-uniform spread: each decade third holds a comparable share of clumps
    lr = np.log10(b.radius / 0.2e-3)
    h, _ = np.histogram(lr, bins=3, range=(0, 1))
    assert h.min() > 0.15 * h.sum()
    sizes = np.bincount(b.clump_id)
    assert sizes.min() >= 1 and sizes.max() <= 6
    assert b.params.Frcoe == 0.6 and b.params.GridLevels == 3
    lo = np.array(b.params.WorldOrigin[:])
    hi = lo + np.array(b.params.WorldSize[:])
    assert np.all(b.pos - b.radius[:, None] >= lo - 1e-9) and np.all(b.pos + b.radius[:, None] <= hi + 1e-9)
    assert b.wall.type == nbc.NBC_WALL_CUBE


def test_level_table_library_equals_oracle_and_covers_radii(c2_small):
    b = c2_small
    p = b.finalized_params(opt3())
    q = orc.finalize_params(b, 0.0, 3)
    assert bytes(p) == bytes(q)
    L = p.Levels
    assert L == 3
    cell = np.array(p.LevelCell[:L])
    assert np.all(np.diff(cell) < 0), "level 0 is the coarsest"
    assert cell[0] == p.CellSize[0] == 3.0 * b.radius.max()
    # level of each sphere by the stated rule, and r <= cell / 3 on its level
    rmin = np.array(p.LevelRmin[:L])
    lev = np.array([next(t for t in range(L) if t == L - 1 or r > rmin[t]) for t in [0] for r in b.radius])
    assert np.all(3.0 * b.radius <= cell[lev] * (1 + 1e-15))
    assert set(np.unique(lev)) == {0, 1, 2}
    off, tot = 0, 0
    for t in range(L):
        dims = np.array(p.LevelDim[t][:])
        assert p.LevelOff[t] == off
        assert np.all(dims * cell[t] >= np.array(p.WorldSize[:]))
        off += int(np.prod(dims))
    assert p.CellNum == off + 1


def test_level_table_rejects_bad_requests(c2_small):
    p = nbc.Params()
    C.memmove(C.byref(p), C.byref(c2_small.params), C.sizeof(nbc.Params))
    o = nbc.make_options(hgrid_levels=9)
    assert nbc.lib().nbc_finalize_params(C.byref(p), nbc.ptr(c2_small.radius), c2_small.n, C.byref(o)) == \
        nbc.NBC_E_BADINPUT
    o = nbc.make_options(hgrid_levels=3, cell_size=0.01)
    assert nbc.lib().nbc_finalize_params(C.byref(p), nbc.ptr(c2_small.radius), c2_small.n, C.byref(o)) == \
        nbc.NBC_E_BADINPUT


def test_hash_formula_two_sizes():
    # two big and two small single-sphere clumps at hand-picked points
    pos = np.array([[0.0101, 0.0202, 0.0303], [0.031, 0.007, 0.012], [0.0112, 0.0213, 0.0301], [0.05, 0.05, 0.05]])
    rad = np.array([2e-3, 2e-3, 0.2e-3, 0.2e-3])
    n = len(pos)
    p = make_params(n, n, lo=(0.0, 0.0, 0.0), size=(0.06, 0.06, 0.06))
    m = 2650 * 4 / 3 * np.pi * rad ** 3
    prop = np.column_stack([0.4 * m * rad ** 2] * 3 + [m])
    bed = nbc.Bed(p, np.arange(n, dtype=np.int32), pos, np.zeros((n, 3)), m, rad, pos.copy(), np.zeros((n, 3)),
                  np.tile([0, 0, 0, 1.0], (n, 1)), np.zeros((n, 3)), prop, None)
    q = orc.finalize_params(bed, 0.0, 2)
    o = orc.Oracle(q)
    o.upload(bed, None)
    o.rebuild()
    O = np.array(q.WorldOrigin[:])
    want = []
    for x, r in zip(pos, rad):
        t = 0 if r > q.LevelRmin[0] else 1
        c = np.floor((x - O) / q.LevelCell[t]).astype(int)
        d = q.LevelDim[t]
        want.append(q.LevelOff[t] + (c[2] * d[1] + c[1]) * d[0] + c[0])
    assert sorted(want) == o.grid()["cell"].tolist()
    # the big spheres (level 0) sort first
    assert np.all(o.spheres()["radius"][:2] == 2e-3)


def test_hgrid_rows_match_kdtree_and_uniform_grid(c2_small):
    b = c2_small
    o = orc.Oracle(orc.finalize_params(b, 0.0, 3), opt3())
    o.upload(b, b.wall)
    o.rebuild()
    s = o.spheres()
    g = o.grid()
    assert np.all(np.diff(g["cell"]) >= 0)
    pr = o.pairs()
    own = np.repeat(np.arange(b.n), np.diff(pr["foff"]))
    rows = set(zip(own.tolist(), pr["fj"].tolist()))
    assert all(i < j for i, j in rows), "the owner of a pair is its lower sorted index"
    assert rows == kd_pairs(s["pos"], s["radius"])
    # back rows are the transpose, every back entry resolves to its owner's slot
    bown = np.repeat(np.arange(b.n), np.diff(pr["boff"]))
    assert set(zip(pr["bi"].tolist(), bown.tolist())) == rows
    assert np.all(pr["bslot"] >= 0)
    assert np.array_equal(pr["fj"][pr["bslot"]], bown)
    # row order: same-level partners first, then finer partners ascending
    q = o.params
    rmin = np.array(q.LevelRmin[:3])
    lev = np.array([0 if r > rmin[0] else (1 if r > rmin[1] else 2) for r in s["radius"]])
    nx = 0
    for i in range(b.n):
        fj = pr["fj"][pr["foff"][i]:pr["foff"][i + 1]]
        cross = fj[lev[fj] != lev[i]]
        nsame = len(fj) - len(cross)
        assert np.all(lev[fj[:nsame]] == lev[i]) and np.all(np.diff(cross) > 0)
        nx += len(cross)
    assert nx > 0
    # intra-clump flags
    assert np.array_equal(pr["fflag"] & 1, (s["cid"][own] == s["cid"][pr["fj"]]).astype(np.uint8))
    # the same pairs, by file index, as the uniform grid on the same bed
    u = orc.Oracle(orc.finalize_params(b), opt3())
    u.upload(b, b.wall)
    u.rebuild()
    su, pu = u.spheres(), u.pairs()
    ou = np.repeat(np.arange(b.n), np.diff(pu["foff"]))
    fid_u = {tuple(sorted(x)) for x in zip(su["fid"][ou].tolist(), su["fid"][pu["fj"]].tolist())}
    fid_h = {tuple(sorted(x)) for x in zip(s["fid"][own].tolist(), s["fid"][pr["fj"]].tolist())}
    assert fid_u == fid_h


def test_hgrid_overflow_is_reported(c2_small):
    b = c2_small
    saved = b.params.BranchLmtLen
    b.params.BranchLmtLen = 8
    try:
        o = orc.Oracle(orc.finalize_params(b, 0.0, 3), opt3())
        o.upload(b, b.wall)
        with pytest.raises(nbc.NbcError) as e:
            o.rebuild()
        assert e.value.code == nbc.NBC_E_NEIGHBOR_OVERFLOW
    finally:
        b.params.BranchLmtLen = saved


def test_hgrid_dynamics_agree_with_uniform_grid(c2_small):
    # same pair set, different pair order: FP64 sums agree to rounding
    b = c2_small
    res = []
    for lv in (0, 3):
        o = orc.Oracle(orc.finalize_params(b, 0.0, lv), opt3())
        o.upload(b, b.wall)
        o.step(30)
        s = o.spheres()
        pos = np.zeros_like(s["pos"])
        pos[s["fid"]] = s["pos"]
        st = o.stats()
        res.append((pos, st.touching_ss, st.touching_wall))
    (pa, ta, wa), (pb, tb, wb) = res
    assert ta == tb and wa == wb and ta > 0
    scale = np.abs(pa).max()
    assert np.abs(pa - pb).max() <= 1e-12 * scale

"""Pins of the nested hierarchical grid (power-of-two levels numbered
coarse-cell-major, cross-level pairs owned by the coarser sphere) in the FP64
oracle and the library's parameter set-up.

Independent statements used: the level rule (the finest level whose cell is
at least 3 r), the key written out with numpy from the coarse cell
floor((x - O) / s0) and the level cell floor((x - O) / s_t), a KD-tree (scipy)
for the 1.5 (ri + rj) pair set, and the uniform grid and the level-major
hierarchical grid on the same bed, which hold the same pairs and, through
rebuilds that carry cross-level histories, the same dynamics to rounding.
"""
import ctypes as C

import numpy as np
import pytest

import nbc
import orc
from test_oracle_hgrid import kd_pairs
from util import make_params


@pytest.fixture(scope="module")
def c2_small():
    return nbc.generate_bed(2, 1500, seed=5, container=(0.025, 0, 0, 0))


def opt_nested(levels=3, prec=nbc.NBC_FP64):
    return nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CUBE, hgrid_levels=levels, hgrid_nested=1)


def levels_of(radius, cell):
    """Finest level whose cell is at least 3 r (level 0 takes the rest)."""
    L = len(cell)
    return np.array([max([t for t in range(L) if 3.0 * r <= cell[t] * (1 + 1e-15)] + [0]) for r in radius])


@pytest.mark.parametrize("L", [2, 3, 4])
def test_nested_table_library_equals_oracle(c2_small, L):
    b = c2_small
    p = b.finalized_params(opt_nested(L))
    q = orc.finalize_params(b, 0.0, L, nested=True)
    assert bytes(p) == bytes(q)
    assert p.Levels == L and p.LevelNested == 1
    s0 = 3.0 * b.radius.max()
    assert p.LevelCell[0] == s0 == p.CellSize[0]
    for t in range(L):
        assert p.LevelCell[t] == s0 / 2 ** t
        assert list(p.LevelDim[t]) == [g << t for g in p.GridDim]
        assert p.LevelOff[t] == sum(8 ** u for u in range(t))
    assert p.LevelBlock == sum(8 ** t for t in range(L))
    assert p.CellNum == int(np.prod(list(p.GridDim))) * p.LevelBlock + 1
    # the level rule of the table: r > LevelRmin[t] stays on t, and every
    # sphere's cell is at least 3 r
    cell = np.array(p.LevelCell[:L])
    rmin = np.array(p.LevelRmin[:L])
    lev = np.array([next(t for t in range(L) if t == L - 1 or r > rmin[t]) for r in b.radius])
    assert np.array_equal(lev, levels_of(b.radius, cell))
    assert np.all(3.0 * b.radius <= cell[lev] * (1 + 1e-15))


def test_nested_table_rejects_bad_requests(c2_small):
    b = c2_small
    for o in (opt_nested(5), nbc.make_options(hgrid_nested=1),
              nbc.make_options(hgrid_levels=3, hgrid_nested=1, cell_size=0.01)):
        p = nbc.Params()
        C.memmove(C.byref(p), C.byref(b.params), C.sizeof(nbc.Params))
        assert nbc.lib().nbc_finalize_params(C.byref(p), nbc.ptr(b.radius), b.n, C.byref(o)) == nbc.NBC_E_BADINPUT


def test_nested_hash_of_hand_placed_spheres():
    # big, medium and small single-sphere clumps at hand-picked points; two
    # small ones share a coarse cell with a big one
    pos = np.array([[0.0101, 0.0202, 0.0303], [0.031, 0.007, 0.012], [0.0112, 0.0213, 0.0301],
                    [0.05, 0.05, 0.05], [0.0115, 0.0195, 0.0322], [0.0290, 0.0101, 0.0147]])
    rad = np.array([2e-3, 2e-3, 0.2e-3, 0.2e-3, 0.7e-3, 0.6e-3])
    n = len(pos)
    p = make_params(n, n, lo=(0.0, 0.0, 0.0), size=(0.06, 0.06, 0.06))
    m = 2650 * 4 / 3 * np.pi * rad ** 3
    prop = np.column_stack([0.4 * m * rad ** 2] * 3 + [m])
    bed = nbc.Bed(p, np.arange(n, dtype=np.int32), pos, np.zeros((n, 3)), m, rad, pos.copy(), np.zeros((n, 3)),
                  np.tile([0, 0, 0, 1.0], (n, 1)), np.zeros((n, 3)), prop, None)
    q = orc.finalize_params(bed, 0.0, 3, nested=True)
    o = orc.Oracle(q)
    o.upload(bed, None)
    o.rebuild()
    O = np.array(q.WorldOrigin[:])
    s0 = q.LevelCell[0]
    G = np.array(q.GridDim[:])
    lev = levels_of(rad, np.array(q.LevelCell[:3]))
    assert lev.tolist() == [0, 0, 2, 2, 1, 1]
    want = []
    for x, t in zip(pos, lev):
        coarse = np.floor((x - O) / s0).astype(int)
        fine = np.floor((x - O) / (s0 / 2 ** t)).astype(int)
        sub = fine - coarse * 2 ** t
        assert np.all((sub >= 0) & (sub < 2 ** t))
        key = ((coarse[2] * G[1] + coarse[1]) * G[0] + coarse[0]) * q.LevelBlock
        key += sum(8 ** u for u in range(t)) + (sub[2] * 2 ** t + sub[1]) * 2 ** t + sub[0]
        want.append(key)
    g = o.grid()
    assert sorted(want) == g["cell"].tolist()
    # spheres of one coarse cell are contiguous, coarser levels first: file
    # spheres 0, 4 and 2 share a coarse cell
    fid = o.spheres()["fid"].tolist()
    assert np.all(np.diff(g["cell"] // q.LevelBlock) >= 0)
    k = fid.index(0)
    assert fid[k:k + 3] == [0, 4, 2]


def test_nested_rows_match_kdtree_and_uniform_grid(c2_small):
    b = c2_small
    q = orc.finalize_params(b, 0.0, 3, nested=True)
    o = orc.Oracle(q, opt_nested())
    o.upload(b, b.wall)
    o.rebuild()
    s, g, pr = o.spheres(), o.grid(), o.pairs()
    assert np.all(np.diff(g["cell"]) >= 0)
    lev = levels_of(s["radius"], np.array(q.LevelCell[:3]))
    own = np.repeat(np.arange(b.n), np.diff(pr["foff"]))
    fj = pr["fj"]
    same = lev[own] == lev[fj]
    assert np.all(own[same] < fj[same]), "same level: the lower sorted index owns the pair"
    assert np.all(lev[own[~same]] < lev[fj[~same]]), "two levels: the coarser sphere owns the pair"
    assert np.any(own[~same] > fj[~same]), "some owners sort after their partner"
    rows = {(min(a, c), max(a, c)) for a, c in zip(own.tolist(), fj.tolist())}
    assert len(rows) == len(own) and rows == kd_pairs(s["pos"], s["radius"])
    bown = np.repeat(np.arange(b.n), np.diff(pr["boff"]))
    assert set(zip(pr["bi"].tolist(), bown.tolist())) == set(zip(own.tolist(), fj.tolist()))
    assert np.all(pr["bslot"] >= 0) and np.array_equal(fj[pr["bslot"]], bown)
    for i in range(b.n):
        row = fj[pr["foff"][i]:pr["foff"][i + 1]]
        cross = row[lev[row] != lev[i]]
        assert np.all(lev[row[:len(row) - len(cross)]] == lev[i]) and np.all(np.diff(cross) > 0)
    u = orc.Oracle(orc.finalize_params(b), opt_nested())
    u.upload(b, b.wall)
    u.rebuild()
    su, pu = u.spheres(), u.pairs()
    ou = np.repeat(np.arange(b.n), np.diff(pu["foff"]))
    fid_u = {tuple(sorted(x)) for x in zip(su["fid"][ou].tolist(), su["fid"][pu["fj"]].tolist())}
    fid_h = {tuple(sorted(x)) for x in zip(s["fid"][own].tolist(), s["fid"][fj].tolist())}
    assert fid_u == fid_h


def test_nested_overflow_is_reported(c2_small):
    b = c2_small
    saved = b.params.BranchLmtLen
    b.params.BranchLmtLen = 8
    try:
        o = orc.Oracle(orc.finalize_params(b, 0.0, 3, nested=True), opt_nested())
        o.upload(b, b.wall)
        with pytest.raises(nbc.NbcError) as e:
            o.rebuild()
        assert e.value.code == nbc.NBC_E_NEIGHBOR_OVERFLOW
    finally:
        b.params.BranchLmtLen = saved


def test_nested_dynamics_through_rebuilds_agree_with_other_grids():
    """Clumps thrown at up to 0.5 m/s, a rebuild forced every 8 steps on top
    of the ones the motion asks for: cross-level pairs whose coarse owner sorts
    after its partner carry their histories through every rebuild. The
    uniform, level-major and nested grids hold the same pairs, so the state
    agrees to rounding; a history lost or negated at a rebuild would show at
    the size of the tangential force."""
    bed = nbc.generate_bed(2, 600, seed=8, container=(0.02, 0, 0, 0))
    rng = np.random.default_rng(8)
    bed.cvel[:] = rng.uniform(-0.5, 0.5, bed.cvel.shape)
    bed.cangv[:] = rng.uniform(-100.0, 100.0, bed.cangv.shape)
    bed.params.StepSize = 2e-6
    res = []
    for kind in ("uniform", "level-major", "nested"):
        q = orc.finalize_params(bed, 0.0, 0 if kind == "uniform" else 3, nested=kind == "nested")
        o = orc.Oracle(q, opt_nested())
        o.upload(bed, bed.wall)
        cross_late = 0
        for k in range(12):
            o.step(8)
            o.request_rebuild()
            if kind == "nested":
                s, pr, h = o.spheres(), o.pairs(), o.pair_history()
                lev = levels_of(s["radius"], np.array(q.LevelCell[:3]))
                own = np.repeat(np.arange(bed.n), np.diff(pr["foff"]))
                late = (lev[own] != lev[pr["fj"]]) & (own > pr["fj"]) & (np.abs(h).max(axis=1) > 0)
                cross_late += int(late.sum())
        c = o.clumps()
        st = o.stats()
        res.append((c["pos"], c["vel"], c["angv"], st.touching_ss, st.rebuilds))
        if kind == "nested":
            assert cross_late > 0, "no history of a late-sorting coarse owner was carried"
        o.close()
    base = res[0]
    assert base[4] >= 12
    for r in res[1:]:
        assert r[3] == base[3] and r[4] == base[4]
        for a, b in zip(r[:3], base[:3]):
            assert np.abs(a - b).max() <= 1e-10 * max(np.abs(b).max(), 1e-30)

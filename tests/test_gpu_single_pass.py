"""Single-pass neighbour search (rows stored entry-major during the count
pass, then compacted) against the FP64 oracle: after every rebuild the sorted
order, cell bounds, rows and back slots are bit-exact, whether the rows came
from the single pass or, when a row outgrew the capacity sized from the
previous rebuild, from the multi-pass search (the first rebuild starts with
32 entries per row, which a polydisperse bed on the uniform grid outgrows),
and with the multi-pass search forced (option single_pass = -1). Rebuilds
are forced between steps as well as taken when the motion asks. Uniform grid
(monodisperse-like C1 bed and polydisperse C2 bed), level-major and nested
hierarchical grids, FP64."""
import numpy as np
import pytest

import nbc
import orc
from test_gpu_hgrid import assert_lists_equal

pytestmark = pytest.mark.gpu


def first_row_difference(ctx, ref):
    pa, pb = ctx.pairs(), ref.pairs()
    for key in ("foff", "boff"):
        d = np.nonzero(pa[key] != pb[key])[0]
        if len(d):
            i = max(int(d[0]) - 1, 0)
            return (f"{key} first differs after sphere {i}: gpu {pa[key][i:i + 3].tolist()} "
                    f"oracle {pb[key][i:i + 3].tolist()}")
    return None


@pytest.mark.parametrize("single_pass", [0, -1])
@pytest.mark.parametrize("kind", ["uniform", "uniform-polydisperse", "hgrid", "hgrid-nested"])
def test_single_pass_rows_bit_exact_through_rebuilds(kind, single_pass):
    print(f"[case] single-pass search, {kind}, single_pass option {single_pass}", flush=True)
    if kind == "uniform":
        bed = nbc.generate_bed(1, 800, seed=12)
        levels, nested, wall = 0, False, nbc.NBC_WALL_CYLINDER
        bed.params.StepSize = 5e-6
    else:
        bed = nbc.generate_bed(2, 1500, seed=5, container=(0.025, 0, 0, 0))
        levels, nested, wall = (0 if kind == "uniform-polydisperse" else 3), kind == "hgrid-nested", nbc.NBC_WALL_CUBE
        bed.params.StepSize = 2e-6
    rng = np.random.default_rng(4)
    bed.cvel[:] = rng.uniform(-0.3, 0.3, bed.cvel.shape)
    opt = nbc.make_options(precision=nbc.NBC_FP64, wall_type=wall, hgrid_levels=levels, hgrid_nested=int(nested),
                           single_pass=single_pass)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    ref = orc.Oracle(orc.finalize_params(bed, 0.0, levels, nested=nested), opt)
    ref.upload(bed, bed.wall)
    longest = 0
    for k in range(5):
        ctx.rebuild()
        ref.rebuild()
        st = ctx.stats()
        print(f"  rebuild {k}: rebuilds {st.rebuilds}, single-pass {st.single_pass_rebuilds}, "
              f"pos diff {np.abs(ctx.spheres()['pos'] - ref.spheres()['pos']).max():.3e}", flush=True)
        diff = first_row_difference(ctx, ref)
        assert diff is None, diff
        assert_lists_equal(ctx, ref)
        pr = ref.pairs()
        longest = max(longest, int(np.diff(pr["foff"]).max()), int(np.diff(pr["boff"]).max()))
        ctx.step(15)
        ref.step(15)
    st = ctx.stats()
    assert st.rebuilds == ref.stats().rebuilds
    if single_pass < 0:
        assert st.single_pass_rebuilds == 0
    else:
        assert st.single_pass_rebuilds >= 3, "later rebuilds take the single pass"
        if levels == 0 and longest > 32:
            # a 2 mm sphere among 0.2 mm ones lists dozens of partners: the
            # first rebuild (32 entries per row) fills with the passes, the
            # later ones fit the capacity sized from it
            assert st.single_pass_rebuilds == st.rebuilds - 1
    assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() <= 1e-12
    ctx.close()
    ref.close()


def test_cell_search_spheres_on_cell_faces():
    """The uniform-grid search runs one wave per cell and hands a sphere to the
    per-sphere loop when its cell coordinates, computed from its position with
    the inverse cell size, differ from its hashed cell (computed with a
    division). Spheres are placed on cell-face values where the two roundings
    disagree (a 4.3 mm cell makes such values common); the rows stay
    bit-exact to the oracle, which makes the same two computations."""
    print("[case] uniform-grid search with spheres on cell faces", flush=True)
    bed = nbc.generate_bed(0, 600, seed=21)
    cs_opt = 0.0043
    opt = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CUBE, cell_size=cs_opt)
    P = bed.finalized_params(opt)
    O, cs, ic = (np.array(v[:3]) for v in (P.WorldOrigin, P.CellSize, P.i_CellSize))

    def split(x, a):
        return np.floor((x - O[a]) / cs[a]) != np.floor((x - O[a]) * ic[a])

    moved = 0
    for k in range(bed.nc):
        # the middle sphere of a C0 clump sits at the clump centre (offset 0)
        m = np.nonzero((bed.clump_id == k) & np.all(bed.local == 0.0, axis=1))[0]
        if len(m) != 1:
            continue
        for a in (0, 1):
            base = O[a] + np.round((bed.cpos[k, a] - O[a]) / cs[a]) * cs[a]
            xs = [base]
            up = dn = base
            for _ in range(50):
                up, dn = np.nextafter(up, np.inf), np.nextafter(dn, -np.inf)
                xs += [up, dn]
            hit = [x for x in xs if split(x, a)]
            if hit:
                bed.cpos[k, a] = hit[0]
                bed.pos[m[0], a] = hit[0]
                moved += 1
                break
    assert moved >= 50, moved
    ctx = nbc.Context(P, opt)
    ctx.upload(bed, bed.wall)
    ref = orc.Oracle(orc.finalize_params(bed, cs_opt), opt)
    ref.upload(bed, bed.wall)
    for k in range(2):
        ctx.rebuild()
        ref.rebuild()
        pos = ctx.spheres()["pos"]
        nsplit = int(np.count_nonzero(split(pos[:, 0], 0) | split(pos[:, 1], 1)))
        print(f"  rebuild {k}: {nsplit} spheres on split cell faces", flush=True)
        assert nsplit >= 50
        diff = first_row_difference(ctx, ref)
        assert diff is None, diff
        assert_lists_equal(ctx, ref)
    assert ctx.stats().single_pass_rebuilds >= 1
    ctx.close()
    ref.close()

"""The oracle side of batched beds: a bed rebuilt whenever another bed of its
batch needs it follows the same physics (rebuilds only re-sort the spheres,
and the Verlet skin keeps every touching pair listed), so its state agrees
with the bed stepped alone to rounding and its touching sets are identical."""
import numpy as np

import nbc
import orc


def test_extra_rebuilds_leave_the_physics_unchanged():
    bed = nbc.generate_bed(1, 250, seed=31, container=(0.011, 0, 0, 0))
    rng = np.random.default_rng(31)
    bed.cvel[:] = rng.uniform(-0.2, 0.2, bed.cvel.shape)
    bed.params.StepSize = 2e-6
    opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER)
    a = orc.Oracle(orc.finalize_params(bed), opt)
    b = orc.Oracle(orc.finalize_params(bed), opt)
    a.upload(bed, bed.wall)
    b.upload(bed, bed.wall)
    forced = 0
    for k in range(120):
        if k % 7 == 3:
            b.request_rebuild()
            forced += 1
        a.step(1)
        b.step(1)
        sa, sb = a.stats(), b.stats()
        assert (sa.touching_ss, sa.touching_wall) == (sb.touching_ss, sb.touching_wall), k
    assert b.stats().rebuilds >= a.stats().rebuilds + forced - 1
    pa, pb = a.spheres(), b.spheres()
    xa, xb = np.zeros_like(pa["pos"]), np.zeros_like(pb["pos"])
    xa[pa["fid"]] = pa["pos"]
    xb[pb["fid"]] = pb["pos"]
    assert np.abs(xa - xb).max() <= 1e-12 * np.abs(xa).max()
    assert a.stats().touching_ss > 0

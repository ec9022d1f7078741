"""Long-run statistics against the FP64 oracle, over a horizon where DEM is
chaotic and individual trajectories part ways.

Workload: the C1-style loose bed (600 clumps of 2-8 spheres, 3,198 spheres,
cylinder R 10 mm, solid fraction 0.50) is shaken (clump velocities N(0,
0.15 m/s), spins N(0, 50 rad/s), seed 4) and stepped 24,000 times at dt
2e-6 s, with checkpoints every 1,000 steps. The shake's kinetic energy decays
through the contacts over the first ~5,000 steps; then the bed collapses under
gravity (height 86.8 -> 83.3 mm, solid fraction 0.500 -> 0.534), which
raises the KE again.

The oracle against itself sets the scale: a 1e-12 relative change in one
clump's velocity leaves KE within 1e-11 for 9,000 steps, then grows about
e-fold per 1,000 steps (KE differs by 3e-3 at step 23,000); a 1e-7 change of
every velocity, the size of the FP32 form's rounding, saturates that growth
within the horizon at KE 4e-3, height and solid fraction 2e-5 and touching
totals 4 %. Stated bands:
* FP64 GPU form: KE within 1e-9 (relative) of the oracle for the first
  16,000 steps, which a 1e-16 difference per step would take that long to
  reach; after that, the FP32-mixed bands (measured: 1.6e-15 over the whole
  horizon, since the summation orders agree);
* KE at every checkpoint within 3 % of the oracle's (measured, FP32-mixed:
  0.65 %);
* bed height (99th percentile of sphere tops) and solid fraction at every
  checkpoint within 0.1 % (measured: 1.7e-6);
* sphere-sphere contacts summed over all checkpoints within 20 % (each
  checkpoint holds only 10-150 touching pairs, so single counts are noise;
  measured: 7 %).
"""
import numpy as np
import pytest

import nbc
import orc

pytestmark = pytest.mark.gpu

STEPS, EVERY, EXACT = 24000, 1000, 16
KE_BAND, PACK_BAND, TOUCH_BAND, FP64_BAND = 0.03, 0.001, 0.20, 1e-9
CONE_WIN_BAND, CONE_CUM_BAND = 0.25, 0.20


def shaken_bed():
    bed = nbc.generate_bed(1, 600, seed=4, container=(0.01, 0, 0, 0))
    bed.params.StepSize = 2e-6
    rng = np.random.default_rng(4)
    bed.cvel[:] = rng.normal(0.0, 0.15, bed.cvel.shape)
    bed.cangv[:] = rng.normal(0.0, 50.0, bed.cangv.shape)
    return bed


def run(bed, o):
    """KE, bed height, solid fraction and touching pairs at each checkpoint."""
    vol = (4.0 / 3.0 * np.pi * bed.radius ** 3).sum()
    area = np.pi * bed.wall.radius ** 2
    out = []
    for _ in range(STEPS // EVERY):
        o.step(EVERY)
        st, s = o.stats(), o.spheres()
        h = np.percentile(s["pos"][:, 2] + s["radius"], 99) - bed.params.WorldOrigin[2]
        out.append((st.kinetic_energy, h, vol / (area * h), st.touching_ss))
    return np.array(out)


@pytest.mark.timeout(900)
def test_shaken_bed_statistics_match_the_oracle():
    print("[case] shaken C1-style bed, 3198 spheres, 24000 steps: FP64 oracle, mixed and FP64 GPU", flush=True)
    bed = shaken_bed()
    opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER)
    orc.set_threads(16)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, bed.wall)
    o_ = run(bed, ref)
    ref.close()
    print(f"[case] oracle done: KE {o_[0, 0]:.3e} -> {o_[-1, 0]:.3e}, solid fraction {o_[0, 2]:.4f} -> {o_[-1, 2]:.4f}",
          flush=True)
    forms = {}
    for prec in (nbc.NBC_MIXED, nbc.NBC_FP64):
        opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
        ctx = nbc.Context(bed.finalized_params(opt), opt)
        ctx.upload(bed, bed.wall)
        forms[prec] = run(bed, ctx)
        ctx.close()
    assert o_[4, 0] < 0.7 * o_[0, 0] and o_[-1, 2] > o_[0, 2] + 0.02, "the shaken bed neither calmed nor compacted"
    for prec, g in forms.items():
        rel = np.abs(g - o_) / np.abs(o_)
        tsum = (g[:, 3].sum(), o_[:, 3].sum())
        print(f"[case] {'mixed' if prec == nbc.NBC_MIXED else 'FP64'}: max KE dev {rel[:, 0].max():.2e} "
              f"(first {EXACT} checkpoints {rel[:EXACT, 0].max():.2e}), height {rel[:, 1].max():.2e}, "
              f"solid fraction {rel[:, 2].max():.2e}, touching {tsum[0]} / {tsum[1]}", flush=True)
        if prec == nbc.NBC_FP64:
            assert rel[:EXACT, 0].max() <= FP64_BAND
        assert rel[:, 0].max() <= KE_BAND
        assert rel[:, 1].max() <= PACK_BAND and rel[:, 2].max() <= PACK_BAND
        assert abs(tsum[0] - tsum[1]) <= TOUCH_BAND * tsum[1]


def cone_apex_clearance(bed, Rc, alpha):
    """Lowest apex height at which no sphere of the bed touches the lateral
    surface of a cone of half-angle alpha on the z axis."""
    rho = np.hypot(bed.pos[:, 0], bed.pos[:, 1])
    m = rho < Rc + bed.radius
    return (bed.pos[m, 2] - (rho[m] * np.cos(alpha) - bed.radius[m]) / np.sin(alpha)).max()


def cone_forces(o, bins):
    """Mean penetrator force of each 1,000-step bin, and the bins' per-step forces."""
    means, steps = [], []
    for _ in range(bins):
        o.step(1000)
        f = o.penetrator()["cache_f"].copy()
        means.append(f.mean(0))
        steps.append(f)
    return np.array(means), np.array(steps)


@pytest.mark.timeout(900)
def test_cone_drive_force_curve_matches_the_oracle():
    """Resistance-force curve of a 60-degree cone (R 3 mm, 48 segments, wall
    friction 0.4) driven at 0.25 m/s for 16,000 steps of 2e-6 s (8 mm) into
    the C1-style bed after it settled under gravity (150,000 GPU steps; bed
    top 48.3 mm, 220 wall contacts). A few spheres carry the load through
    force chains, so the curve is rough: the oracle against itself, with the
    cone speed changed by -1e-7, 1e-9 and 3e-8 (relative), gives 1,000-step
    bin means up to 21 % apart at depth, 2 mm window means up to 8.2 % and
    total impulses up to 7.0 %; the first 6,000 steps agree within 2e-4.
    Stated bands, about three times that spread:
    * FP64 GPU form: every per-step force within 1e-9 of the curve's largest
      force over the first 8,000 steps;
    * FP32-mixed form: per-step forces within 2e-3 of the largest over the
      first 2,000 steps, before the oracle's own spread passes 1e-3;
    * both forms: each 2 mm window's mean axial force within CONE_WIN_BAND
      and the total axial impulse within CONE_CUM_BAND of the oracle's.
    """
    print("[case] cone drive into the settled C1-style bed: 16000 steps, mixed and FP64 GPU against the oracle",
          flush=True)
    bed = nbc.generate_bed(1, 600, seed=4, container=(0.01, 0, 0, 0))
    bed.params.StepSize = 2e-6
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    ctx.step(150000)
    bed = nbc.download_bed(ctx, bed)
    ctx.close()
    bed.params.isIntruder = 1
    bed.params.OutputInterval = 1000
    Rc = 0.003
    Hc = Rc / np.tan(np.radians(30.0))
    apex = cone_apex_clearance(bed, Rc, np.radians(30.0)) + 20e-6
    pene = nbc.cone_penetrator(Rc, Hc, (0.0, 0.0, apex + Hc), (0.0, 0.0, -0.25), nseg=48)
    print(f"[case] settled: apex {apex * 1e3:.3f} mm", flush=True)
    opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER, pene_frcoe=0.4)
    orc.set_threads(16)
    ref = orc.Oracle(orc.finalize_params(bed), opt)
    ref.upload(bed, bed.wall, pene)
    mo, so = cone_forces(ref, 16)
    ref.close()
    scale = np.abs(so).max()
    print(f"[case] oracle done: bin mean Fz {mo[0, 2]:.3e} -> {mo[-1, 2]:.3e} N, largest force {scale:.3e} N", flush=True)
    assert mo[-1, 2] > 100 * mo[0, 2] > 0, "the cone did not meet growing resistance"
    wo, co = mo[:, 2].reshape(4, 4).mean(1), mo[:, 2].sum()
    for prec in (nbc.NBC_FP64, nbc.NBC_MIXED):
        opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER, pene_frcoe=0.4)
        ctx = nbc.Context(bed.finalized_params(opt), opt)
        ctx.upload(bed, bed.wall, pene)
        mg, sg = cone_forces(ctx, 16)
        ctx.close()
        step_dev = np.abs(sg - so).reshape(16, -1).max(1) / scale
        wg, cg = mg[:, 2].reshape(4, 4).mean(1), mg[:, 2].sum()
        wrel, crel = np.abs(wg - wo) / np.abs(wo), abs(cg - co) / abs(co)
        name = "FP64" if prec == nbc.NBC_FP64 else "mixed"
        print(f"[case] {name}: per-step deviation by bin {np.array2string(step_dev, precision=1)}; "
              f"2 mm windows {np.array2string(wrel, precision=2)}; total impulse {crel:.2e}", flush=True)
        if prec == nbc.NBC_FP64:
            assert step_dev[:8].max() <= 1e-9
        else:
            assert step_dev[:2].max() <= 2e-3
        assert wrel.max() <= CONE_WIN_BAND and crel <= CONE_CUM_BAND

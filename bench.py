"""Benchmark of the clumped-sphere DEM timestep on one MI355X.

Workload (the configuration the headline metric is quoted on): 200,000
clumps from 8 random templates of 2-8 sub-spheres (lognormal radii, median
1 mm, sigma 0.15, clipped to [0.6, 1.6] mm), about 1.12M spheres, deposited in
a cylinder wall of radius 60 mm; E = 1e8 Pa, nu = 0.3, rho = 2650 kg/m^3,
e = 0.5, mu = 0.5, dt = 5e-7 s, uniform grid. A "step" is one full timestep
of the whole bed (kinematics, rebuild when flagged, contacts, walls, clump
reduction, integration) with all state resident in HBM.

value: sphere-steps/s of the FP32-mixed form; the plain FP64 form runs the
same steps beside it. roofline: the contact kernel (k_contacts_f32, which also
takes the wall-list spheres' wall contacts in extra blocks), the dominant
kernel, with algorithmic bytes per launch
    N * 72 B + C_ss * 80 B + C_w * 72 B
(per-array accounting of the force inputs/outputs; C_ss / C_w = touching
sphere-sphere / wall contacts counted on the device each step), timed with HIP
events on the library stream over a second timed region that follows the
throughput region (the events themselves add gaps between kernels, so
`value` comes from a region without them). cpu_baseline: the FP64 oracle (the reference's
slot-padded dataflow restated in C++ with OpenMP) on a bounded sample of the
same bed.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nbody-clump_amd"))

HBM_PEAK = 8.0e12
AMORT_STEPS = 1000


def run_form(bed, precision, steps, warmup, prof_steps):
    import nbc

    opt = nbc.make_options(precision=precision, wall_type=nbc.NBC_WALL_CYLINDER)
    params = bed.finalized_params(opt)
    ctx = nbc.Context(params, opt)
    ctx.upload(bed, bed.wall)
    ctx.step(warmup)
    s0 = ctx.stats()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)  # returns after a stream synchronize
    t1 = time.perf_counter()
    s1 = ctx.stats()
    # amortised window: 1,000 further steps whatever --steps is, so that the
    # rebuilds and near-list refreshes a longer run meets are in the figure
    ctx.sync()
    ta = time.perf_counter()
    ctx.step(AMORT_STEPS)
    tb = time.perf_counter()
    sa = ctx.stats()
    # kernel durations: HIP events around every kernel class, recorded on the
    # stream each kernel runs on, over a second timed region of the same run
    # (the events add gaps between kernels, so the throughput region runs
    # without them)
    ctx.profile(True)
    ctx.step(max(prof_steps, 1))
    ctx.sync()
    s2 = ctx.stats()
    k3_ms, k3_n = ctx.profile_get(1)
    out = dict(seconds=t1 - t0, k3_ms=k3_ms, k3_n=k3_n, rebuilds=s1.rebuilds - s0.rebuilds,
               prof_rebuilds=s2.rebuilds - sa.rebuilds, kinetic_energy=s1.kinetic_energy,
               amortised=dict(steps=AMORT_STEPS, ms_per_step=(tb - ta) / AMORT_STEPS * 1e3,
                              rebuilds=sa.rebuilds - s1.rebuilds, near_refreshes=sa.near_refreshes - s1.near_refreshes),
               touching_ss=0.5 * (s1.touching_ss + s2.touching_ss), touching_wall=0.5 * (s1.touching_wall + s2.touching_wall),
               front_pairs=s2.front_pairs, near_pairs=s2.near_pairs,
               near_refreshes=s1.near_refreshes - s0.near_refreshes)
    for cls in (0, 2, 3, 4, 5, 6):
        out[f"cls{cls}_ms"], out[f"cls{cls}_n"] = ctx.profile_get(cls)
    ctx.close()
    return out


def settle(bed, steps, dt, wall_type=None):
    """Let the deposited bed collapse under gravity (FP32-mixed form, larger
    step) so the benchmark runs on a contact-rich packing."""
    import nbc

    opt = nbc.make_options(precision=nbc.NBC_MIXED,
                           wall_type=nbc.NBC_WALL_CYLINDER if wall_type is None else wall_type)
    saved = bed.params.StepSize
    bed.params.StepSize = dt
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    chunk = 5000
    for done in range(0, steps, chunk):
        ctx.step(min(chunk, steps - done))
        s = ctx.stats()
        print(f"[bench] settle step {s.step}: touching {s.touching_ss / bed.n:.3f}/sphere, wall {s.touching_wall}, "
              f"KE {s.kinetic_energy:.4e} J, rebuilds {s.rebuilds}", file=sys.stderr, flush=True)
    out = nbc.download_bed(ctx, bed)
    ctx.close()
    bed.params.StepSize = saved
    out.params.StepSize = saved
    return out


def run_c2(a):
    """Polydisperse bed on the hierarchical grid (not the headline line): the
    deposited C2 bed (1.93M spheres in 571K clumps of 1-6, radii log-uniform
    0.2-2.0 mm, cube section 150 mm), FP32-mixed, 3-level grid next to the
    uniform grid on the same bed; steady steps and forced rebuilds timed
    separately."""
    import nbc

    if a.bed_cache and os.path.exists(a.bed_cache):
        bed = nbc.Bed.load(a.bed_cache)
    else:
        t0 = time.perf_counter()
        bed = nbc.generate_bed(2, a.nclumps if a.nclumps != 200000 else 0, seed=1)
        print(f"[bench] C2 bed: {bed.n} spheres, {bed.nc} clumps in {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        if a.bed_cache:
            bed.save(a.bed_cache)
    out = {}
    grids = (("hgrid3_nested", 3, 1), ("hgrid3", 3, 0), ("uniform", 0, 0))
    for key, levels, nested in [g for g in grids if g[0] in a.c2_grids.split(",")]:
        opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CUBE, hgrid_levels=levels,
                               hgrid_nested=nested)
        ctx = nbc.Context(bed.finalized_params(opt), opt)
        ctx.upload(bed, bed.wall)
        ctx.step(a.warmup)
        s0 = ctx.stats()
        ctx.profile(True)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(a.steps)
        t1 = time.perf_counter()
        s1 = ctx.stats()
        k3_ms, k3_n = ctx.profile_get(1)
        k2_ms, k2_n = ctx.profile_get(0)
        k4_ms, k4_n = ctx.profile_get(2)
        r_ms0, r_n0 = ctx.profile_get(3)
        for _ in range(3):
            ctx.rebuild()
        ctx.sync()
        r_ms, r_n = ctx.profile_get(3)
        out[key] = dict(sphere_steps_per_s=bed.n * a.steps / (t1 - t0), ms_per_step=(t1 - t0) / a.steps * 1e3,
                        k_contacts_us=k3_ms / max(k3_n, 1) * 1e3, k_sphere_pose_us=k2_ms / max(k2_n, 1) * 1e3,
                        k_clump_after_us=k4_ms / max(k4_n, 1) * 1e3, rebuilds_in_steps=s1.rebuilds - s0.rebuilds,
                        rebuild_ms=(r_ms - r_ms0) / max(r_n - r_n0, 1), touching_ss_per_sphere=s1.touching_ss / bed.n,
                        list_pairs_per_sphere=s1.front_pairs / bed.n, kinetic_energy_J=s1.kinetic_energy)
        print(f"[bench] C2 {key}: {out[key]}", file=sys.stderr, flush=True)
        ctx.close()
    h = out.get("hgrid3_nested") or next(iter(out.values()))
    print(json.dumps({
        "metric": "sphere-steps/s, C2 polydisperse bed on the hierarchical grid (FP32-mixed)",
        "value": h["sphere_steps_per_s"], "unit": "sphere-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": h["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 forces / f64 integration", "data": "synthetic (deterministic generator, seed 1)",
        "config": {"workload": "C2 bed: 571,429 clumps of 1-6 spheres, radii log-uniform 0.2-2.0 mm, cube section "
                               "150 mm, mu 0.6, dt 2e-7 s, 3-level nested hierarchical grid, as deposited",
                   "n_spheres": bed.n, "n_clumps": bed.nc},
        "forms": out,
    }))


def run_c5(a):
    """DEMmain's single-sphere cylinder loop (RunSimulation, not the headline
    line): 1,000,000 single spheres (radii as C1) deposited in a cylinder of
    radius 60 mm, settled, then stepped at dt 5e-7 s in the FP32-mixed and
    FP64 forms; plus the glue loop (RunSimulation_Split) with every tenth
    sphere fixed and every tenth driven."""
    import numpy as np

    import nbc

    if a.bed_cache and os.path.exists(a.bed_cache):
        part = nbc.Particles.load(a.bed_cache)
    else:
        t0 = time.perf_counter()
        part = nbc.Particles.from_bed(nbc.generate_bed(3, a.nclumps if a.nclumps != 200000 else 1000000, seed=1))
        print(f"[bench] C5 bed: {part.n} spheres in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if a.settle_steps > 0:
            opt = nbc.make_options(precision=nbc.NBC_MIXED, driver=nbc.NBC_DRIVER_CYLINDER)
            saved = part.params.StepSize
            part.params.StepSize = a.settle_dt
            ctx = nbc.Context(part.finalized_params(opt), opt)
            ctx.upload_particles(part, part.wall)
            for done in range(0, a.settle_steps, 5000):
                ctx.step(min(5000, a.settle_steps - done))
                st = ctx.stats()
                print(f"[bench] settle step {st.step}: touching {st.touching_ss / part.n:.3f}/sphere, "
                      f"KE {st.kinetic_energy:.4e} J", file=sys.stderr, flush=True)
            cl = ctx.clumps()
            for k in ("pos", "vel", "quat", "angv"):
                getattr(part, k)[:] = cl[k]
            ctx.close()
            part.params.StepSize = saved
        if a.bed_cache:
            part.save(a.bed_cache)
    forms = {}
    for name, driver, prec in (("cylinder_mixed", nbc.NBC_DRIVER_CYLINDER, nbc.NBC_MIXED),
                               ("cylinder_fp64", nbc.NBC_DRIVER_CYLINDER, nbc.NBC_FP64),
                               ("split_mixed", nbc.NBC_DRIVER_SPLIT, nbc.NBC_MIXED)):
        if prec == nbc.NBC_FP64 and a.no_fp64:
            continue
        p2 = nbc.Particles(part.params, part.type.copy(), part.pos, part.vel, part.quat, part.angv, part.prop,
                           part.radius, part.wall)
        if driver == nbc.NBC_DRIVER_SPLIT:
            p2.type[:] = np.where(np.arange(part.n) % 10 == 3, 2, np.where(np.arange(part.n) % 10 == 7, 1, 0))
        opt = nbc.make_options(precision=prec, driver=driver)
        ctx = nbc.Context(p2.finalized_params(opt), opt)
        ctx.upload_particles(p2, p2.wall)
        ctx.step(a.warmup)
        s0 = ctx.stats()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(a.steps)
        t1 = time.perf_counter()
        s1 = ctx.stats()
        ctx.profile(True)
        ctx.step(max(a.prof_steps, 1))
        ctx.sync()
        k3_ms, k3_n = ctx.profile_get(1)
        forms[name] = dict(sphere_steps_per_s=part.n * a.steps / (t1 - t0), ms_per_step=(t1 - t0) / a.steps * 1e3,
                           k_contacts_us=k3_ms / max(k3_n, 1) * 1e3, rebuilds=s1.rebuilds - s0.rebuilds,
                           touching_ss_per_sphere=s1.touching_ss / part.n, touching_wall=s1.touching_wall)
        print(f"[bench] C5 {name}: {forms[name]}", file=sys.stderr, flush=True)
        ctx.close()
    m = forms["cylinder_mixed"]
    print(json.dumps({
        "metric": "sphere-steps/s, DEMmain single-sphere cylinder loop (FP32-mixed)",
        "value": m["sphere_steps_per_s"], "unit": "sphere-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 forces / f64 integration", "data": "synthetic (deterministic generator, seed 1)",
        "config": {"workload": "C5 bed: 1,000,000 single spheres, lognormal radii (median 1 mm), cylinder R 60 mm, "
                               "dt 5e-7 s, settled; DEMmain cylinder loop",
                   "n_spheres": part.n},
        "forms": forms,
    }))


def run_c0(a):
    """The reference's small CPU-runnable case (not the headline line): 2,000
    clumps of 3 collinear spheres (6,000 spheres) in a 40 x 40 x 80 mm cube,
    settled under gravity until the spheres rest on each other, then --steps
    steps (20,000 in the configuration) at dt 1e-6 s: FP64 and FP32-mixed on
    the GPU next to the FP64 oracle on the host cores over the same steps,
    in its faithful (slot-padded) and compact modes."""
    import nbc

    bed = nbc.generate_bed(0, 2000, seed=1)
    t0 = time.perf_counter()
    bed = settle(bed, 100000, 5e-6, wall_type=nbc.NBC_WALL_CUBE)
    print(f"[bench] C0 settled 100000 steps of 5e-6 s in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    out = {}
    for prec, name in ((nbc.NBC_FP64, "fp64"), (nbc.NBC_MIXED, "mixed")):
        opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CUBE)
        ctx = nbc.Context(bed.finalized_params(opt), opt)
        ctx.upload(bed, bed.wall)
        ctx.step(a.warmup)
        s0 = ctx.stats()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(a.steps)
        t1 = time.perf_counter()
        st = ctx.stats()
        out[name] = dict(sphere_steps_per_s=bed.n * a.steps / (t1 - t0), us_per_step=(t1 - t0) / a.steps * 1e6,
                         rebuilds=st.rebuilds - s0.rebuilds, touching_ss=st.touching_ss,
                         touching_wall=st.touching_wall, kinetic_energy_J=st.kinetic_energy)
        print(f"[bench] C0 {name}: {out[name]}", file=sys.stderr, flush=True)
        ctx.close()
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc

        orc.set_threads(os.cpu_count() or 1)
        for compact, name in ((False, "cpu_oracle_fp64_faithful"), (True, "cpu_oracle_fp64_compact")):
            o = orc.Oracle(orc.finalize_params(bed), nbc.make_options(wall_type=nbc.NBC_WALL_CUBE), compact=compact)
            o.upload(bed, bed.wall)
            o.step(a.warmup)
            t0 = time.perf_counter()
            o.step(a.steps)
            dt_cpu = time.perf_counter() - t0
            st = o.stats()
            out[name] = dict(sphere_steps_per_s=bed.n * a.steps / dt_cpu, seconds=dt_cpu, cores=orc.threads(),
                             steps=a.steps, touching_ss=st.touching_ss)
            print(f"[bench] C0 {name}: {out[name]}", file=sys.stderr, flush=True)
            o.close()
    m = out["mixed"]
    print(json.dumps({
        "metric": "sphere-steps/s, C0 small bed (FP32-mixed)", "value": m["sphere_steps_per_s"],
        "unit": "sphere-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": m["us_per_step"] / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 forces / f64 integration", "data": "synthetic (deterministic generator, seed 1)",
        "config": {"workload": "C0: 2,000 clumps of 3 collinear spheres, cube 40x40x80 mm, settled, dt 1e-6 s",
                   "n_spheres": bed.n}, "forms": out}))


def run_c3(a):
    """Penetrator case (not the headline line): a C1-style bed of 200,000
    clumps deposited and settled in a cylinder of radius 80 mm, then a
    triangulated 60-degree cone of base radius 10 mm driven down at 0.1 m/s
    from first contact for --pene-depth (40 mm: 800,000 steps of 5e-7 s),
    penetrator friction 0.4, FP32-mixed. Reports the throughput of the first
    --steps steps and the resistance force versus depth: the FP64 reaction on
    the cone recorded every step (src/dem_integrate_penetrator.cu:97-111),
    averaged in 0.5 mm depth bins."""
    import numpy as np

    import nbc

    if a.bed_cache and os.path.exists(a.bed_cache):
        bed = nbc.Bed.load(a.bed_cache)
    else:
        bed = nbc.generate_bed(1, a.nclumps, seed=3, container=(0.08, 0, 0, 0))
        bed.params.BranchLmtLen = 128
        if a.settle_steps > 0:
            bed = settle(bed, a.settle_steps, a.settle_dt)
        if a.bed_cache:
            bed.save(a.bed_cache)
    chunk = 20000
    bed.params.isIntruder = 1
    bed.params.OutputInterval = chunk
    dt = bed.params.StepSize
    v = 0.1
    Rc = 0.010
    alpha = np.radians(30.0)
    Hc = Rc / np.tan(alpha)
    pos = bed.pos.reshape(-1, 3)
    rho = np.hypot(pos[:, 0], pos[:, 1])
    m = rho < Rc + bed.radius
    gap0 = 20e-6
    apex = (pos[m, 2] - (rho[m] * np.cos(alpha) - bed.radius[m]) / np.sin(alpha)).max() + gap0
    pene = nbc.cone_penetrator(Rc, Hc, (0.0, 0.0, apex + Hc), (0.0, 0.0, -v), nseg=64)
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER, pene_frcoe=0.4)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall, pene)
    total = int(round((a.pene_depth + gap0) / (v * dt)))
    fz = np.zeros(total)
    done = 0
    t_steps = None

    def run_chunk(k):
        nonlocal done
        ctx.step(k)
        cf = ctx.penetrator()["cache_f"][:, 2]
        steps = np.arange(done, done + k)
        fz[done:done + k] = cf[steps % chunk]
        done += k

    ctx.sync()
    t0 = time.perf_counter()
    run_chunk(a.steps)  # the timed window starts at first contact
    t1 = time.perf_counter()
    st = ctx.stats()
    t_steps = t1 - t0
    print(f"[bench] C3 timed {a.steps} steps: {t_steps / a.steps * 1e3:.4f} ms/step", file=sys.stderr, flush=True)
    tc = time.perf_counter()
    while done < total:
        run_chunk(min(chunk - done % chunk, total - done))
        if done % (10 * chunk) == 0 or done == total:
            s2 = ctx.stats()
            print(f"[bench] C3 drive: depth {(done * v * dt - gap0) * 1e3:.2f} mm, last-chunk mean Fz "
                  f"{fz[max(done - chunk, 0):done].mean():.3f} N, rebuilds {s2.rebuilds}, near refreshes "
                  f"{s2.near_refreshes}, {time.perf_counter() - tc:.0f} s", file=sys.stderr, flush=True)
    drive_s = time.perf_counter() - tc
    s3 = ctx.stats()
    ctx.close()
    depth = np.arange(1, total + 1) * v * dt - gap0  # apex depth below first contact after each step
    edges = np.arange(0.0, a.pene_depth + 1e-12, 0.5e-3)
    curve = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (depth > lo) & (depth <= hi)
        if np.any(sel):
            curve.append([round(hi * 1e3, 3), float(fz[sel].mean()), float(fz[sel].max())])
    res = {"metric": "sphere-steps/s, C3 cone penetrator in a settled 1M-sphere bed (FP32-mixed)",
           "value": bed.n * a.steps / t_steps, "unit": "sphere-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": 0, "ms_per_step": t_steps / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 forces / f64 integration", "data": "synthetic (deterministic generator)",
           "config": {"workload": "C3: 200,000 clumps (C1 templates) settled in a cylinder R 80 mm, 60-degree cone "
                                  "R 10 mm with 128 triangles at 0.1 m/s from first contact, friction 0.4, dt 5e-7 s",
                      "n_spheres": bed.n, "triangles": int(pene.ntri)},
           "penetrator": {"depth_m": float(depth[-1]), "steps": total, "drive_seconds": drive_s + t_steps,
                          "ms_per_step_whole_drive": (drive_s + t_steps) / total * 1e3,
                          "rebuilds": s3.rebuilds, "near_refreshes": s3.near_refreshes,
                          "contacts_at_end": s3.touching_wall,
                          "force_vs_depth": {"columns": ["depth_mm (bin upper edge)", "mean Fz N", "max Fz N"],
                                             "rows": curve}}}
    print(json.dumps(res))


def c4_beds(nbeds, seed0=1000, threads=16):
    """nbeds C1-style beds, clump counts uniform in [2000, 60000], cylinder
    radius scaled with count (60 mm at 200,000 clumps), seed seed0 + bed;
    generated in parallel threads (the generator releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import nbc

    rng = np.random.default_rng(seed0)
    counts = rng.integers(2000, 60001, nbeds)

    def one(b):
        R = 0.06 * (counts[b] / 200000.0) ** (1.0 / 3.0)
        return nbc.generate_bed(1, int(counts[b]), seed=seed0 + b, container=(R, 0, 0, 0))

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, range(nbeds)))


def c4_bed_error(beds, picks, steps=10):
    """Per-bed error of the FP32-mixed batch against the FP64 oracle of each
    picked bed: the whole batch and the picked beds' oracles advance `steps`
    steps together (an oracle rebuilds wherever the batch did), the last step
    recorded; per-sphere totals, clump force and torque and positions are
    compared in file order."""
    import numpy as np

    import nbc

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    orc.set_threads(os.cpu_count() or 1)
    opt = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.BatchContext([b.finalized_params(opt) for b in beds], opt)
    ctx.upload(beds)
    refs = {}
    for b in picks:
        o = orc.Oracle(orc.finalize_params(beds[b]), nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER), compact=True)
        o.upload(beds[b], beds[b].wall)
        refs[b] = o
    for k in range(steps):
        r0 = ctx.stats().rebuilds
        if k == steps - 1:
            ctx.set_record(True)
        ctx.step(1)
        rebuilt = ctx.stats().rebuilds > r0
        for o in refs.values():
            if rebuilt and k > 0:
                o.request_rebuild()
            o.step(1)
    so, co = ctx.ranges()
    sp = ctx.spheres()
    fa, ta = ctx.sphere_forces(0)
    cl = ctx.clumps()
    out = {}
    for b, o in refs.items():
        s0, s1 = so[b], so[b + 1]
        n = s1 - s0
        loc = sp["fid"][s0:s1] - s0
        F, T, X = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        F[loc], T[loc], X[loc] = fa[s0:s1], ta[s0:s1], sp["pos"][s0:s1]
        sb = o.spheres()
        fb, tb = o.sphere_forces(0)
        Fb, Tb, Xb = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        Fb[sb["fid"]], Tb[sb["fid"]], Xb[sb["fid"]] = fb, tb, sb["pos"]
        cb = o.clumps()
        c0, c1 = co[b], co[b + 1]
        rel = lambda x, y: float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))
        out[f"bed{b}"] = dict(n_spheres=int(n), sphere_F=rel(F, Fb), sphere_T=rel(T, Tb),
                              clump_F=rel(cl["F"][c0:c1], cb["F"]), clump_T=rel(cl["T"][c0:c1], cb["T"]),
                              pos_abs_m=float(np.abs(X - Xb).max()), touching_ss_oracle=int(o.stats().touching_ss))
        o.close()
    ctx.close()
    return out


def run_c4(a):
    """Batched beds (not the headline line): nbeds independent C1-style beds
    advanced by one context (one launch set per step), FP32-mixed and FP64.
    The options come from the batch autotuner (nbc_autotune_batch) run on every
    eighth bed; the full batch is also timed with its steady chunks as HIP
    graph replays and without; the per-bed error of the mixed batch against
    the FP64 oracle is reported for the smallest, the median and the largest
    bed; next to it the first 16 beds stepped as separate single-bed contexts."""
    import numpy as np

    import nbc

    t0 = time.perf_counter()
    beds = c4_beds(a.nbeds)
    N = sum(b.n for b in beds)
    print(f"[bench] C4: {a.nbeds} beds, {N} spheres, {sum(b.nc for b in beds)} clumps in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    base = nbc.make_options(precision=nbc.NBC_MIXED, wall_type=nbc.NBC_WALL_CYLINDER)
    t0 = time.perf_counter()
    sample = beds[::8]
    tuned, table = nbc.autotune_batch(sample, base, steps=100, flags=nbc.NBC_TUNE_CELL)
    tune = dict(sample_beds=len(sample), sample_spheres=sum(b.n for b in sample), seconds=time.perf_counter() - t0,
                best=dict(near_skin=tuned.near_skin, batch_steps=tuned.batch_steps, radix_bits=tuned.radix_bits,
                          block_size=tuned.block_size, graphs=tuned.graphs, cell_size=tuned.cell_size),
                candidates=[dict(near_skin=o.near_skin, batch_steps=o.batch_steps, radix_bits=o.radix_bits,
                                 block_size=o.block_size, graphs=o.graphs, cell_size=o.cell_size, ms_per_step=ms)
                            for o, ms in table])
    print(f"[bench] C4 autotune: {tune['best']} in {tune['seconds']:.1f} s", file=sys.stderr, flush=True)

    def timed(opt, name):
        ctx = nbc.BatchContext([b.finalized_params(opt) for b in beds], opt)
        ctx.upload(beds)
        ctx.step(a.warmup)
        s0 = ctx.stats()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(a.steps)
        t1 = time.perf_counter()
        s1 = ctx.stats()
        ctx.profile(True)
        ctx.step(max(a.prof_steps // 3, 1))
        ctx.sync()
        k3_ms, k3_n = ctx.profile_get(1)
        rb_ms, rb_n = ctx.profile_get(3)
        r = dict(sphere_steps_per_s=N * a.steps / (t1 - t0), ms_per_step=(t1 - t0) / a.steps * 1e3,
                 k_contacts_us=k3_ms / max(k3_n, 1) * 1e3, rebuilds_in_steps=s1.rebuilds - s0.rebuilds,
                 rebuild_ms=rb_ms / max(rb_n, 1), graph_replays=s1.graph_replays - s0.graph_replays,
                 touching_ss_per_sphere=s1.touching_ss / N, kinetic_energy_J=s1.kinetic_energy)
        print(f"[bench] C4 {name}: {r}", file=sys.stderr, flush=True)
        ctx.close()
        return r

    out = {}
    out["batch_mixed_tuned"] = timed(tuned, "batch_mixed_tuned")
    for g, name in ((1, "batch_mixed_graphs_on"), (-1, "batch_mixed_graphs_off")):
        o = nbc.Options.from_buffer_copy(tuned)
        o.graphs = g
        out[name] = timed(o, name)
    out["batch_mixed_default"] = timed(base, "batch_mixed_default")
    out["batch_fp64"] = timed(nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CYLINDER), "batch_fp64")
    # the same kind of beds one context each, stepped one after another
    sub = beds[:16]
    ns = sum(b.n for b in sub)
    ctxs = []
    for b in sub:
        c = nbc.Context(b.finalized_params(base), base)
        c.upload(b, b.wall)
        c.step(a.warmup)
        ctxs.append(c)
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(a.steps)
    for c in ctxs:
        c.sync()
    t1 = time.perf_counter()
    out["separate_contexts_mixed_16_beds"] = dict(sphere_steps_per_s=ns * a.steps / (t1 - t0), n_spheres=ns)
    for c in ctxs:
        c.close()
    settling = c4_settling(beds, tuned, a.c4_settle_steps) if a.c4_settle_steps > 0 else None
    err = None
    if not a.no_cpu:
        sizes = np.array([b.n for b in beds])
        order = np.argsort(sizes)
        picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
        err = c4_bed_error(beds, picks)
        print(f"[bench] C4 per-bed oracle error: {err}", file=sys.stderr, flush=True)
    m = out["batch_mixed_tuned"]
    print(json.dumps({
        "metric": "sphere-steps/s, C4 batched beds (FP32-mixed, one context, autotuned)",
        "value": m["sphere_steps_per_s"], "unit": "sphere-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 forces / f64 integration", "data": "synthetic (deterministic generator)",
        "config": {"workload": f"C4: {a.nbeds} C1-style beds, clump counts uniform in [2000, 60000], cylinder "
                               "radius scaled with count, as deposited, dt 5e-7 s",
                   "n_spheres": N, "n_beds": a.nbeds},
        "autotune": tune, "forms": out, "settling_rebuilds": settling, "oracle_error_per_bed": err,
    }))


def c4_settling(beds, opt, steps, pre=10000, dt=5e-6, sample_every=8):
    """Cost of rebuilding every bed whenever one bed needs it, on beds that
    are settling: the deposited batch falls for `pre` steps of dt (ten times
    the timed step) until its beds move fast enough to rebuild often, then
    `steps` further steps are timed with the rebuilds timed beside them
    (profile class 3). Every `sample_every`-th bed is stepped alone through
    the same pre-fall and window to count the rebuilds it needs by itself.
    The saving bound of per-bed rebuilds is the share of the union's rebuild
    work (sphere count x rebuilds) that the beds' own rebuilds would not
    need, estimated on the sample."""
    import nbc

    old = [b.params.StepSize for b in beds]
    for b in beds:
        b.params.StepSize = dt
    try:
        N = sum(b.n for b in beds)
        ctx = nbc.BatchContext([b.finalized_params(opt) for b in beds], opt)
        ctx.upload(beds)
        chunk = 2000
        for done in range(0, pre, chunk):
            ctx.step(min(chunk, pre - done))
            s = ctx.stats()
            print(f"[bench] C4 settling: step {s.step}, {s.rebuilds} union rebuilds", file=sys.stderr, flush=True)
        ctx.profile(True)
        s0 = ctx.stats()
        r_ms0, r_n0 = ctx.profile_get(3)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        t1 = time.perf_counter()
        s1 = ctx.stats()
        rb_ms, rb_n = ctx.profile_get(3)
        rb_ms, rb_n = rb_ms - r_ms0, rb_n - r_n0
        ctx.close()
        union = s1.rebuilds - s0.rebuilds
        own, work_own, work_union = [], 0.0, 0.0
        for b in beds[::sample_every]:
            c = nbc.Context(b.finalized_params(opt), opt)
            c.upload(b, b.wall)
            c.step(pre)
            r0 = c.stats().rebuilds
            c.step(steps)
            own.append(c.stats().rebuilds - r0)
            c.close()
            work_own += b.n * own[-1]
            work_union += b.n * union
        frac = work_own / max(work_union, 1.0)
        res = dict(dt=dt, pre_steps=pre, steps=steps, ms_per_step=(t1 - t0) / steps * 1e3, union_rebuilds=union,
                   rebuild_ms_each=rb_ms / max(rb_n, 1), rebuild_share_of_time=rb_ms / ((t1 - t0) * 1e3),
                   sampled_beds=len(own), own_rebuilds_mean=sum(own) / len(own), own_rebuilds_max=max(own),
                   per_bed_rebuild_work_fraction=frac, per_bed_saving_bound_ms_per_step=rb_ms * (1 - frac) / steps,
                   n_spheres=N)
        print(f"[bench] C4 settling rebuilds: {res}", file=sys.stderr, flush=True)
        return res
    finally:
        for b, s in zip(beds, old):
            b.params.StepSize = s


def profiled_traffic():
    """HBM bytes per k_contacts_f32 launch on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (each scaled by what the
    counter reads for a copy of known size). None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_contacts_f32_c1_mixed.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_probe(mib):
    """A plain device-to-device copy of known size: reading its FETCH_SIZE and
    WRITE_SIZE counters gives the factor that converts those counters into HBM
    bytes for the contact kernel."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n * 4, "launches": 5}))


def _relerr(a, b):
    import numpy as np

    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def oracle_error(bed, ref, prec, steps):
    """Per-output error of one GPU form against the FP64 oracle `ref`, both
    advanced `steps` steps from the same bed (the last one recorded): the
    outputs of src/dem_force_clump.cu:165-191 (per-sphere totals),
    src/dem_clumpsphere.cu:110-151 (clump F, T), the clump state, the pair
    histories and the contact set."""
    import numpy as np

    import nbc

    opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER)
    ctx = nbc.Context(bed.finalized_params(opt), opt)
    ctx.upload(bed, bed.wall)
    ctx.step(steps - 1)
    ctx.set_record(True)
    ctx.step(1)
    sa, sb = ctx.spheres(), ref.spheres()
    ia, ib = np.argsort(sa["fid"]), np.argsort(sb["fid"])  # file order
    fa, ta = ctx.sphere_forces(0)
    fb, tb = ref.sphere_forces(0)
    ca, cb = ctx.clumps(), ref.clumps()
    pa, pb = ctx.pairs(), ref.pairs()
    lists_equal = all(np.array_equal(pa[k], pb[k]) for k in ("foff", "fj", "boff", "bi", "bslot")) and \
        np.array_equal(sa["fid"], sb["fid"])
    st_a, st_b = ctx.stats(), ref.stats()
    out = dict(sphere_F=_relerr(fa[ia], fb[ib]), sphere_T=_relerr(ta[ia], tb[ib]),
               clump_F=_relerr(ca["F"], cb["F"]), clump_T=_relerr(ca["T"], cb["T"]),
               clump_vel=_relerr(ca["vel"], cb["vel"]), clump_angv=_relerr(ca["angv"], cb["angv"]),
               clump_pos_abs_m=float(np.abs(ca["pos"] - cb["pos"]).max()),
               clump_quat_abs=float(np.abs(ca["quat"] - cb["quat"]).max()),
               lists_bitexact=bool(lists_equal), touching_ss=[int(st_a.touching_ss), int(st_b.touching_ss)],
               touching_wall=[int(st_a.touching_wall), int(st_b.touching_wall)],
               rebuilds=[int(st_a.rebuilds), int(st_b.rebuilds)])
    if lists_equal:
        ha, hb = ctx.pair_history(), ref.pair_history()
        xa, _ = ctx.pair_exports()
        xb, _ = ref.pair_exports()
        out["pair_history"] = _relerr(ha, hb)
        out["contact_set_equal"] = bool(np.array_equal(np.linalg.norm(xa, axis=1) > 0, np.linalg.norm(xb, axis=1) > 0))
    ctx.close()
    return out


def cpu_baseline(bed, sample_steps, with_error=True):
    """The FP64 oracle on the host cores: the reference's slot-padded dataflow
    ("faithful") over `sample_steps` steps after the list-building step, with
    one forced rebuild halfway (src/dem_simulation_clump.cu:124-175), then the
    compact slot mode (CSR pair range) over the same number of steps without
    it. The compact run is the FP64 reference for the per-output error of both
    GPU forms on the same steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    import nbc

    orc.set_threads(os.cpu_count() or 1)
    p = orc.finalize_params(bed)
    opt = nbc.make_options(wall_type=nbc.NBC_WALL_CYLINDER)
    o = orc.Oracle(p, opt)
    o.upload(bed, bed.wall)
    o.step(1)  # first step builds the lists
    half = sample_steps // 2
    t0 = time.perf_counter()
    o.step(half)
    o.request_rebuild()
    o.step(sample_steps - half)
    dt = time.perf_counter() - t0
    forced = o.stats().rebuilds
    o.close()
    oc = orc.Oracle(p, opt, compact=True)
    oc.upload(bed, bed.wall)
    oc.step(1)
    t0 = time.perf_counter()
    oc.step(sample_steps)
    dtc = time.perf_counter() - t0
    res = dict(value=bed.n * sample_steps / dt, unit="sphere-steps/s", cores=orc.threads(), kind="port",
               sample=f"FP64 oracle, faithful mode (the reference's slot-padded dataflow, OpenMP), same "
                      f"{bed.n}-sphere bed, {sample_steps} steps after the first (list-building) step with one "
                      f"forced rebuild halfway ({forced} rebuilds in all), BranchLmtLen "
                      f"{bed.params.BranchLmtLen}, {dt:.1f} s",
               compact=dict(value=bed.n * sample_steps / dtc, unit="sphere-steps/s",
                            sample=f"compact slot mode (CSR pair range), same steps without the forced rebuild, "
                                   f"{dtc:.1f} s"))
    err = None
    if with_error:
        err = {"steps": 1 + sample_steps,
               "reference": "FP64 oracle (compact mode) from the same timed bed"}
        for prec, name in ((nbc.NBC_FP64, "fp64"), (nbc.NBC_MIXED, "mixed")):
            err[name] = oracle_error(bed, oc, prec, 1 + sample_steps)
            print(f"[bench] oracle error {name}: {err[name]}", file=sys.stderr, flush=True)
    oc.close()
    return res, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--nclumps", type=int, default=200000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=25,
                    help="steps of the CPU baseline sample (about 10 s of oracle time on 16 cores at 1.1M spheres)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--settle-steps", type=int, default=60000)
    ap.add_argument("--settle-dt", type=float, default=5e-6)
    ap.add_argument("--bed-cache", default=None,
                    help="load the settled bed from this .npz if it exists, else build it and save it there")
    ap.add_argument("--workload", choices=("c0", "c1", "c2", "c3", "c4", "c5"), default="c1",
                    help="c1: the headline bed (default); c0: the small cube bed; c2: polydisperse bed on the "
                         "hierarchical grid; c3: cone penetrator in a settled bed; c4: batched beds; c5: DEMmain "
                         "single-sphere loops")
    ap.add_argument("--nbeds", type=int, default=256, help="beds of the c4 workload")
    ap.add_argument("--c2-grids", default="hgrid3_nested,hgrid3,uniform",
                    help="c2: grids to time (hgrid3_nested, hgrid3, uniform)")
    ap.add_argument("--c4-settle-steps", type=int, default=1000,
                    help="c4: steps of the settling window that measures the union-rebuild cost (0: skip)")
    ap.add_argument("--pene-depth", type=float, default=0.04, help="c3: cone travel from first contact (m)")
    ap.add_argument("--prof-steps", type=int, default=300,
                    help="steps of the second timed region, with HIP events around every kernel")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_probe_mib > 0:
        copy_probe(a.copy_probe_mib)
        return
    import nbc

    if nbc.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.workload == "c2":
        run_c2(a)
        return
    if a.workload == "c4":
        run_c4(a)
        return
    if a.workload == "c0":
        run_c0(a)
        return
    if a.workload == "c3":
        run_c3(a)
        return
    if a.workload == "c5":
        run_c5(a)
        return
    if a.bed_cache and os.path.exists(a.bed_cache):
        bed = nbc.Bed.load(a.bed_cache)
        print(f"[bench] bed loaded from {a.bed_cache}: {bed.n} spheres, {bed.nc} clumps", file=sys.stderr, flush=True)
    else:
        bed = nbc.generate_bed(1, a.nclumps, seed=1)
        bed.params.BranchLmtLen = 128
        print(f"[bench] bed: {bed.n} spheres, {bed.nc} clumps", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        if a.settle_steps > 0:
            bed = settle(bed, a.settle_steps, a.settle_dt)
        print(f"[bench] settled {a.settle_steps} steps of {a.settle_dt} s in {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        if a.bed_cache:
            bed.save(a.bed_cache)
    mixed = run_form(bed, nbc.NBC_MIXED, a.steps, a.warmup, a.prof_steps)
    print(f"[bench] mixed: {mixed}", file=sys.stderr, flush=True)
    fp64 = None
    if not a.no_fp64:
        fp64 = run_form(bed, nbc.NBC_FP64, a.steps, a.warmup, a.prof_steps)
        print(f"[bench] fp64: {fp64}", file=sys.stderr, flush=True)
    N = bed.n
    value = N * a.steps / mixed["seconds"]
    def avg(form, key):
        return form[key + "_ms"] / max(form[key + "_n"], 1) * 1e-3

    t_contacts = avg(mixed, "k3")
    k3_avg = t_contacts
    k3_bytes = N * 72 + mixed["touching_ss"] * 80 + mixed["touching_wall"] * 72
    achieved = k3_bytes / k3_avg
    # whole-step algorithmic bytes (B_step, mixed): N*236 + Nc*336 + C_ss*80 + C_w*72
    step_bytes = N * 236 + bed.nc * 336 + mixed["touching_ss"] * 80 + mixed["touching_wall"] * 72
    res = {
        "metric": "sphere-steps/s @1M clumped spheres, FP32-mixed vs FP64; % HBM peak",
        "value": value,
        "unit": "sphere-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": mixed["seconds"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 integration",
        "data": "synthetic (deterministic generator, seed 1)",
        "config": {
            "workload": "C1 bed: 200,000 clumps from 8 templates of 2-8 spheres, lognormal radii, "
                        "cylinder R 60 mm, dt 5e-7 s, uniform grid; deposited, then settled for "
                        f"{a.settle_steps} steps of {a.settle_dt} s before timing",
            "n_spheres": N, "n_clumps": bed.nc, "touching_ss_per_sphere": mixed["touching_ss"] / N,
            "list_pairs_per_sphere": mixed["front_pairs"] / N, "rebuilds_in_timed_steps": mixed["rebuilds"],
            "near_entries_per_sphere": mixed["near_pairs"] / N, "near_refreshes_in_timed_steps": mixed["near_refreshes"],
            "kinetic_energy_J": mixed["kinetic_energy"],
        },
        # the same run continued for 1,000 steps: rebuilds and near-list
        # refreshes included, so short timed windows do not flatter the rate
        "amortised": {"mixed": dict(mixed["amortised"], sphere_steps_per_s=N / mixed["amortised"]["ms_per_step"] * 1e3),
                      "fp64": (dict(fp64["amortised"], sphere_steps_per_s=N / fp64["amortised"]["ms_per_step"] * 1e3)
                               if fp64 else None)},
        "roofline": {
            "kernel": "k_contacts_f32",
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": profiled_traffic(),
            "traffic_source": "profiles/k_contacts_f32_c1_mixed.md (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                              "copy-calibrated)",
            "avg_launch_us": k3_avg * 1e6, "algorithmic_bytes_per_launch": k3_bytes,
            "avg_us": {"k_contacts (+ wall-list blocks)": t_contacts * 1e6,
                       "k_sphere_pose": avg(mixed, "cls0") * 1e6,
                       "k_clump_after (+ next step's k_clump_before)": avg(mixed, "cls2") * 1e6},
        },
        "step_hbm_frac": step_bytes * a.steps / mixed["seconds"] / HBM_PEAK,
    }
    if fp64:
        v64 = N * a.steps / fp64["seconds"]
        res["fp64_value"] = v64
        res["speedup_vs_fp64"] = value / v64
    if not a.no_cpu:
        try:
            res["cpu_baseline"], res["oracle_error"] = cpu_baseline(bed, a.cpu_steps)
        except Exception as e:  # report, do not hide
            res["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(res))


if __name__ == "__main__":
    main()

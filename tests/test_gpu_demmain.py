"""GPU parity of DEMmain's single-sphere loops (nbc_options.driver: CUBE =
RunSimulation2, CYLINDER = RunSimulation, SPLIT = RunSimulation_Split with glue
types, HGRID = RunSimulation_HGrid) against the FP64 oracle, through the C ABI.

Bars: cell hashes, sorted order, cell bounds, pair rows and back slots
bit-exact; FP64 forces, positions and histories within 1e-12; FP32-mixed
forces within 1e-4 relative to the largest, positions within 1e-7 m.
"""
import ctypes as C
import os

import numpy as np
import pytest

import nbc
import orc
from util import cube_wall

pytestmark = pytest.mark.gpu

DRIVERS = [nbc.NBC_DRIVER_CUBE, nbc.NBC_DRIVER_CYLINDER, nbc.NBC_DRIVER_SPLIT, nbc.NBC_DRIVER_HGRID]


def banner(msg):
    print(f"[case] {msg}", flush=True)


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make_part(driver, speed=0.0, seed=7, n=3000):
    bed = nbc.generate_bed(3, n, seed=seed, container=(0.02, 0, 0, 0))
    part = nbc.Particles.from_bed(bed)
    rng = np.random.default_rng(seed)
    if speed:
        part.vel[:] = rng.uniform(-speed, speed, part.vel.shape)
        part.angv[:] = rng.uniform(-100.0, 100.0, part.angv.shape)
        part.params.StepSize = 2e-6
        part.params.Thresh = 1e-4
    if driver == nbc.NBC_DRIVER_SPLIT:
        part.type[:] = rng.choice([0, 1, 2], size=part.n, p=[0.8, 0.1, 0.1])
    wall = part.wall
    if driver == nbc.NBC_DRIVER_CUBE:
        R, H = bed.wall.radius, bed.wall.height
        wall = cube_wall((R, R, H), (0.0, 0.0, H))
    return part, wall


def setup(part, wall, driver, prec):
    levels = 3 if driver == nbc.NBC_DRIVER_HGRID else 0
    opt = nbc.make_options(precision=prec, driver=driver, hgrid_levels=levels)
    ctx = nbc.Context(part.finalized_params(opt), opt)
    ctx.upload_particles(part, wall)
    ref = orc.Oracle(orc.finalize_params(part, 0.0, levels), opt)
    ref.upload_particles(part, wall)
    return ctx, ref


def assert_lists_equal(ctx, ref):
    ga, gb = ctx.grid(), ref.grid()
    for k in ("cell", "perm", "beg", "end"):
        assert np.array_equal(ga[k], gb[k]), f"grid {k} differs"
    pa, pb = ctx.pairs(), ref.pairs()
    for k in ("foff", "fj", "fflag", "boff", "bi", "bslot"):
        assert np.array_equal(pa[k], pb[k]), f"pairs {k} differs"
    assert np.array_equal(ctx.spheres()["fid"], ref.spheres()["fid"])


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
@pytest.mark.parametrize("driver", DRIVERS)
def test_first_step(driver, prec):
    part, wall = make_part(driver)
    banner(f"DEMmain driver {driver} first step, precision {prec}, N={part.n}")
    ctx, ref = setup(part, wall, driver, prec)
    ctx.set_record(True)
    ctx.step(1)
    ref.step(1)
    assert_lists_equal(ctx, ref)
    sa, sb = ctx.stats(), ref.stats()
    assert sa.touching_ss == sb.touching_ss > 0
    assert sa.touching_wall == sb.touching_wall > 0
    tol = 1e-12 if prec == nbc.NBC_FP64 else 1e-4
    for kind in (1, 2, 0):
        fa, ta = ctx.sphere_forces(kind)
        fb, tb = ref.sphere_forces(kind)
        assert relerr(fa, fb) <= tol, f"force kind {kind}: {relerr(fa, fb):.3e}"
        assert relerr(ta, tb) <= tol, f"torque kind {kind}: {relerr(ta, tb):.3e}"
    assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() == 0.0
    ctx.close()
    ref.close()


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
@pytest.mark.parametrize("driver", DRIVERS)
def test_steps_with_rebuilds(driver, prec):
    """Random velocities up to 0.3 m/s and spins up to 100 rad/s at dt 2e-6 s
    and Thresh 0.1 mm: several rebuilds in 300 steps, so the history exchange (CYLINDER, SPLIT) or
    inheritance (CUBE, HGRID) and the glue kinematics are exercised."""
    part, wall = make_part(driver, speed=0.3)
    banner(f"DEMmain driver {driver} 300 steps with rebuilds, precision {prec}")
    ctx, ref = setup(part, wall, driver, prec)
    for k in range(15):
        ctx.step(20)
        ref.step(20)
        if prec == nbc.NBC_FP64:
            sa, sb = ctx.stats(), ref.stats()
            assert (sa.touching_ss, sa.touching_wall, sa.rebuilds) == (sb.touching_ss, sb.touching_wall,
                                                                        sb.rebuilds), k
    sb = ref.stats()
    assert sb.rebuilds >= 3
    ca, cb = ctx.clumps(), ref.clumps()
    if prec == nbc.NBC_FP64:
        assert_lists_equal(ctx, ref)
        assert np.abs(ctx.spheres()["pos"] - ref.spheres()["pos"]).max() <= 1e-12
        assert relerr(ctx.pair_history(), ref.pair_history()) <= 1e-12
        for k in ("pos", "vel", "quat", "angv"):
            assert relerr(ca[k], cb[k]) <= 1e-12, k
    else:
        assert np.abs(ca["pos"] - cb["pos"]).max() < 1e-7
    if driver == nbc.NBC_DRIVER_SPLIT:
        fixed = part.type == 2
        assert np.array_equal(ca["pos"][fixed], part.pos[fixed])
        assert np.array_equal(ca["vel"][fixed], part.vel[fixed])
    ctx.close()
    ref.close()


@pytest.mark.parametrize("driver", [nbc.NBC_DRIVER_CUBE, nbc.NBC_DRIVER_SPLIT])
def test_particles_view(driver):
    part, wall = make_part(driver, speed=0.3)
    banner(f"DEMmain driver {driver} nbc_get_particles")
    ctx, _ = setup(part, wall, driver, nbc.NBC_FP64)
    ctx.step(60)
    v = ctx.particles()
    cl = ctx.clumps()
    fid = v["fid"]
    assert np.array_equal(v["sid"], np.arange(part.n))
    assert not np.array_equal(fid, np.arange(part.n))  # the bed was sorted
    assert np.array_equal(v["pos"], ctx.spheres()["pos"])
    assert np.array_equal(v["pos"], cl["pos"][fid])
    assert np.array_equal(v["vel"], cl["vel"][fid]) and np.array_equal(v["quat"], cl["quat"][fid])
    assert np.array_equal(v["prop"], part.prop[fid]) and np.array_equal(v["radius"], part.radius[fid])
    # the cube loop leaves Type in file order; the others sort it
    want = part.type if driver == nbc.NBC_DRIVER_CUBE else part.type[fid]
    assert np.array_equal(v["type"], want)
    ctx.close()


def test_run_simulation_cylinder_loop(tmp_path):
    """nbc_run_simulation on DEMmain input files: world.par, the 21-column
    ph.00000000.bt, Rocks/CylindWall.bt and nbc.par Driver = cylinder; the
    OutPart/ph.<step> files hold the oracle's state."""
    part, wall = make_part(nbc.NBC_DRIVER_CYLINDER, speed=0.1, n=800)
    banner(f"DEMmain run loop, cylinder driver, N={part.n}")
    spec = nbc.BedSpec(3, 800, 7, nbc.D4(0.02, 0, 0, 0))
    out = nbc.BedOut()
    assert nbc.lib().nbc_generate_bed(C.byref(spec), C.byref(out)) == 0
    try:
        out.params.EndStep, out.params.OutputInterval, out.params.StepSize = 40, 20, part.params.StepSize
        out.params.Thresh = part.params.Thresh
        assert nbc.lib().nbc_save_bed_text(str(tmp_path).encode(), C.byref(out), None) == 0
    finally:
        nbc.lib().nbc_free_bed(C.byref(out))
    rows = np.column_stack([np.arange(part.n), np.zeros(part.n), part.type, part.pos, part.vel, part.quat,
                            part.angv, part.prop, part.radius])
    np.savetxt(tmp_path / "ph.00000000.bt", rows, fmt=["%d", "%d", "%d"] + ["%.17g"] * 18)
    (tmp_path / "nbc.par").write_text("Driver = cylinder\n")
    # the file round trip is exact (%.17g)
    t = np.zeros(part.n, np.int32)
    a = [np.zeros((part.n, k)) for k in (3, 3, 4, 3, 4)] + [np.zeros(part.n)]
    assert nbc.lib().nbc_load_particles(str(tmp_path / "ph.00000000.bt").encode(), part.n, nbc.ptr(t),
                                        *[nbc.ptr(x) for x in a]) == 0
    for x, y in zip(a, (part.pos, part.vel, part.quat, part.angv, part.prop, part.radius)):
        assert np.array_equal(x, y)
    assert nbc.lib().nbc_run_simulation(str(tmp_path).encode(), 0, 0) == 0
    opt = nbc.make_options(driver=nbc.NBC_DRIVER_CYLINDER)
    ref = orc.Oracle(orc.finalize_params(part), opt)
    ref.upload_particles(part, wall)
    for step in (20, 40):
        ref.step(20)
        raw = open(tmp_path / "OutPart" / f"ph.{step:08d}.bin", "rb").read()
        n = int(np.frombuffer(raw[:4], np.int32)[0])
        assert n == part.n
        ints = np.frombuffer(raw[4:4 + 12 * n], np.int32).reshape(3, n)
        dbl = np.frombuffer(raw[4 + 12 * n:], np.float64)
        assert len(dbl) == n * (3 + 3 + 4 + 3 + 4 + 1)
        pos = dbl[:3 * n].reshape(n, 3)
        s = ref.spheres()
        assert np.array_equal(ints[0], np.arange(n)) and np.array_equal(ints[1], s["fid"])
        assert np.abs(pos - s["pos"]).max() <= 1e-12
        radius = dbl[-n:]
        assert np.array_equal(radius, s["radius"])
    ref.close()


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
@pytest.mark.parametrize("driver", [nbc.NBC_DRIVER_CUBE, nbc.NBC_DRIVER_CYLINDER])
def test_penetrator(driver, prec):
    """A sphere penetrator pressed into the top of the bed at 1 m/s: the
    DEMmain penetrator law and the drift without its term, 300 steps."""
    part, wall = make_part(driver)
    top = int(np.argmax(part.pos[:, 2]))
    R = 3e-3
    pp = part.pos[top] + np.array([0.0, 0.0, part.radius[top] + R - 1e-5])
    pene = nbc.sphere_penetrator(R, tuple(pp), (0.0, 0.0, -1.0))
    part.params.isIntruder = 1
    banner(f"DEMmain driver {driver} with a sphere penetrator, precision {prec}")
    opt = nbc.make_options(precision=prec, driver=driver)
    ctx = nbc.Context(part.finalized_params(opt), opt)
    ctx.upload_particles(part, wall, pene)
    ref = orc.Oracle(orc.finalize_params(part), opt)
    ref.upload_particles(part, wall, pene)
    seen = 0
    for k in range(15):
        ctx.step(20)
        ref.step(20)
        pa, pb = ctx.penetrator(), ref.penetrator()
        seen = max(seen, np.abs(pb["cache_f"]).max())
        if prec == nbc.NBC_FP64:
            assert ctx.stats().touching_wall == ref.stats().touching_wall, k
            assert relerr(pa["cache_f"], pb["cache_f"]) <= 1e-12, k
    assert seen > 0
    ca, cb = ctx.clumps(), ref.clumps()
    if prec == nbc.NBC_FP64:
        for k in ("pos", "vel", "angv", "F", "T"):
            assert relerr(ca[k], cb[k]) <= 1e-12, k
    else:
        assert np.abs(ca["pos"] - cb["pos"]).max() < 1e-7
        assert relerr(ctx.penetrator()["cache_f"], ref.penetrator()["cache_f"]) <= 1e-3
    ctx.close()
    ref.close()


def test_run_simulation_cube_loop_with_intruder(tmp_path):
    """The cube loop with isIntruder: ph, intruder and fc files every
    OutputInterval, Nei/Fcn/Cpt every two."""
    part, wall = make_part(nbc.NBC_DRIVER_CUBE, n=600)
    banner(f"DEMmain cube run loop with intruder, N={part.n}")
    top = int(np.argmax(part.pos[:, 2]))
    spec = nbc.BedSpec(3, 600, 7, nbc.D4(0.02, 0, 0, 0))
    out = nbc.BedOut()
    assert nbc.lib().nbc_generate_bed(C.byref(spec), C.byref(out)) == 0
    R = 3e-3
    pene = nbc.sphere_penetrator(R, tuple(part.pos[top] + np.array([0.0, 0.0, part.radius[top] + R - 1e-5])),
                                 (0.0, 0.0, -1.0))
    try:
        out.params.EndStep, out.params.OutputInterval, out.params.isIntruder = 40, 10, 1
        out.wall = wall
        assert nbc.lib().nbc_save_bed_text(str(tmp_path).encode(), C.byref(out), C.byref(pene)) == 0
    finally:
        nbc.lib().nbc_free_bed(C.byref(out))
    rows = np.column_stack([np.arange(part.n), np.zeros(part.n), part.type, part.pos, part.vel, part.quat,
                            part.angv, part.prop, part.radius])
    np.savetxt(tmp_path / "ph.00000000.bt", rows, fmt=["%d", "%d", "%d"] + ["%.17g"] * 18)
    (tmp_path / "nbc.par").write_text("Driver = cube\n")
    assert nbc.lib().nbc_run_simulation(str(tmp_path).encode(), 0, 0) == 0, nbc.lib().nbc_last_error(None)
    for step in (10, 20, 30, 40):
        assert os.path.getsize(tmp_path / "OutPart" / f"ph.{step:08d}.bin") == 4 + 12 * part.n + 8 * 18 * part.n
        assert os.path.getsize(tmp_path / "OutPart" / f"intruder.{step:08d}.bin") == 4 + 4 + 8 * 17
        assert (tmp_path / "OutForce" / f"fc.{step:08d}.bin").exists()
    for step in (20, 40):
        for k in ("Nei", "Fcn", "Cpt"):
            assert (tmp_path / "OutCnt" / f"{k}.{step:08d}.bin").exists()
    assert not (tmp_path / "OutCnt" / "Nei.00000010.bin").exists()
    # the intruder file holds the penetrator state of the oracle run
    opt = nbc.make_options(driver=nbc.NBC_DRIVER_CUBE)
    part.params.isIntruder = 1
    ref = orc.Oracle(orc.finalize_params(part), opt)
    ref.upload_particles(part, wall, pene)
    ref.step(40)
    raw = np.frombuffer(open(tmp_path / "OutPart" / "intruder.00000040.bin", "rb").read()[8:], np.float64)
    assert np.abs(raw[:3] - ref.penetrator()["pos"]).max() <= 1e-15
    ref.close()


def grid_bt_part(seed=11):
    """A DEMmain bed with three radius classes, its generated 3-level table
    written as grid.bt, and each particle's type a random level it fits."""
    part, wall = make_part(nbc.NBC_DRIVER_HGRID, speed=0.3, seed=seed)
    rng = np.random.default_rng(seed)
    part.radius[:] = part.radius * rng.choice([0.4, 0.7, 1.0], part.n)
    p = orc.finalize_params(part, 0.0, 3)
    ok = [[t for t in range(3) if 3.0 * r <= p.LevelCell[t]] for r in part.radius]
    part.type[:] = [rng.choice(o) for o in ok]
    return part, wall, p


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_hgrid_levels_from_grid_bt(prec):
    part, wall, p = grid_bt_part()
    assert set(part.type) == {0, 1, 2}
    banner(f"DEMmain hgrid loop, levels by type, precision {prec}, N={part.n}")
    opt = nbc.make_options(precision=prec, driver=nbc.NBC_DRIVER_HGRID, level_from_type=1)
    ctx = nbc.Context(p, opt)
    ctx.upload_particles(part, wall)
    ref = orc.Oracle(p, opt)
    ref.upload_particles(part, wall)
    ctx.step(1)
    ref.step(1)
    assert_lists_equal(ctx, ref)
    for k in range(15):
        ctx.step(20)
        ref.step(20)
        if prec == nbc.NBC_FP64:
            sa, sb = ctx.stats(), ref.stats()
            assert (sa.touching_ss, sa.rebuilds) == (sb.touching_ss, sb.rebuilds), k
    assert ref.stats().rebuilds >= 3
    ca, cb = ctx.clumps(), ref.clumps()
    if prec == nbc.NBC_FP64:
        assert_lists_equal(ctx, ref)
        assert relerr(ca["pos"], cb["pos"]) <= 1e-12 and relerr(ca["vel"], cb["vel"]) <= 1e-12
    else:
        assert np.abs(ca["pos"] - cb["pos"]).max() < 1e-7
    # a type that is no level, or a particle too large for its level, is refused
    bad = nbc.Particles(part.params, part.type.copy(), part.pos, part.vel, part.quat, part.angv, part.prop,
                        part.radius, part.wall)
    bad.type[0] = 3
    with pytest.raises(nbc.NbcError):
        ctx.upload_particles(bad, wall)
    bad.type[:] = 2
    with pytest.raises(nbc.NbcError):
        ctx.upload_particles(bad, wall)
    ctx.close()
    ref.close()

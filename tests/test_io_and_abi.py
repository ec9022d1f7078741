"""Drop-in boundary on the CPU: file formats of the reference, failure
semantics of its loaders, the binary output layout, and the C ABI itself
(every symbol of include/nbc.h is exported by libnbc.so). No GPU calls."""
import ctypes as C
import os
import re
import shutil

import numpy as np
import pytest

import nbc

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "reference_inputs")
ROOT = os.path.dirname(HERE)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "nbc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nbc_[a-z_0-9]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    L = C.CDLL(nbc.LIB_PATH)
    names = header_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding knows all of them
    assert set(names) <= set(nbc._SIGS), set(names) - set(nbc._SIGS)


def load_params(path):
    p = nbc.Params()
    ini = C.create_string_buffer(256)
    rc = nbc.lib().nbc_load_world_par(path.encode(), C.byref(p), ini, 256)
    return rc, p, ini.value.decode()


def test_world_par_fixture():
    rc, p, ini = load_params(os.path.join(FIX, "world.par"))
    assert rc == 0
    assert (p.ParticleNum, p.ClumpNum, p.BranchLmtLen) == (576, 144, 256)
    assert ini == "ph.00000000.bt"
    assert (p.StepSize, p.EndStep, p.StartStep, p.OutputInterval) == (1e-5, 200000, 0, 10000)
    assert (p.Ymod, p.Prat, p.Frcoe, p.NormRes, p.Beta, p.Gravity, p.Thresh) == (2e7, 0.2, 0.5, 0.2, 0.5, -9.81, 1e-3)
    assert list(p.WorldSize) == [0.4, 0.4, 1.0] and list(p.BoxUpLmt) == [0.4, 0.4, 1.0]
    assert (p.isIntruder, p.isWallForce, p.isSelfGravity, p.isQuaternion) == (1, 1, 0, 1)
    ln = np.log(0.2)
    assert p.Damping == pytest.approx(2 * np.sqrt(5 / 6) * ln / np.sqrt(ln * ln + 3.14159265359 ** 2), rel=1e-15)
    assert p.Damping < 0


def test_world_par_failures(tmp_path):
    src = open(os.path.join(FIX, "world.par")).read()
    bad = tmp_path / "missing.par"
    bad.write_text(src.replace("Cohesion    = 0.0", ""))
    assert load_params(str(bad))[0] == nbc.NBC_E_PARAM
    assert "missing" in nbc.lib().nbc_last_error(None).decode()
    bad.write_text(src + "\nUnknownKey = 3\n")
    assert load_params(str(bad))[0] == nbc.NBC_E_PARAM
    assert "Unrecoganized item UnknownKey" in nbc.lib().nbc_last_error(None).decode()
    assert load_params(str(tmp_path / "nope.par"))[0] == nbc.NBC_E_IO


def test_mesh_size_matches_reference_rule():
    rc, p, _ = load_params(os.path.join(FIX, "world.par"))
    r = np.full(576, 0.005)
    assert nbc.lib().nbc_finalize_params(C.byref(p), nbc.ptr(r), 576, None) == 0
    assert list(p.CellSize) == [0.015] * 3
    assert list(p.WorldOrigin) == [-0.015] * 3
    assert p.WorldSize[2] == pytest.approx(1.03)
    assert list(p.GridDim) == [int(np.ceil(0.43 / 0.015))] * 2 + [int(np.ceil(1.03 / 0.015))]
    assert p.CellNum == p.GridDim[0] * p.GridDim[1] * p.GridDim[2]


def test_phase_and_clump_fixtures():
    n, nc = 576, 144
    cid = np.zeros(n, np.int32)
    pos, loc, m, r = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    rc = nbc.lib().nbc_load_phase(os.path.join(FIX, "ph.00000000.bt").encode(), n, nbc.ptr(cid), nbc.ptr(pos),
                                  nbc.ptr(loc), nbc.ptr(m), nbc.ptr(r))
    assert rc == 0
    raw = np.loadtxt(os.path.join(FIX, "ph.00000000.bt"))
    assert np.array_equal(cid, raw[:, 2].astype(int))
    assert np.array_equal(pos, raw[:, 3:6]) and np.array_equal(r, raw[:, 10])
    cp, cv, cq, cw, pr = (np.zeros((nc, 3)), np.zeros((nc, 3)), np.zeros((nc, 4)), np.zeros((nc, 3)), np.zeros((nc, 4)))
    rc = nbc.lib().nbc_load_clumps(os.path.join(FIX, "ch.00000000.bt").encode(), nc, *[nbc.ptr(a) for a in (cp, cv, cq, cw, pr)])
    assert rc == 0
    rawc = np.loadtxt(os.path.join(FIX, "ch.00000000.bt"))
    assert np.array_equal(cq, rawc[:, 7:11]) and np.array_equal(pr, rawc[:, 14:18])
    # every sphere's file position equals its clump pose applied to its local offset
    from util import quat_to_matrix
    for i in range(0, n, 37):
        c = cid[i]
        x = cp[c] + quat_to_matrix(cq[c]) @ loc[i]
        assert np.allclose(x, pos[i], atol=1e-8)


def test_rocks_fixtures(tmp_path):
    w = nbc.Wall()
    assert nbc.lib().nbc_load_cube_wall(os.path.join(FIX, "Rocks", "CubeWall.bt").encode(), C.byref(w)) == 0
    assert list(w.side) == [0.2, 0.2, 0.5] and list(w.pos) == [0.2, 0.2, 0.5]
    c = nbc.Wall()
    assert nbc.lib().nbc_load_cylinder_wall(os.path.join(FIX, "Rocks", "CylindWall.bt").encode(), C.byref(c)) == 0
    assert (c.radius, c.height, list(c.dir)) == (0.05, 0.5, [0.0, 0.0, 1.0])
    p = nbc.Penetrator()
    assert nbc.lib().nbc_load_penetrator(os.path.join(FIX, "Rocks", "Penetrator.bt").encode(), C.byref(p)) == 0
    assert (p.shape, p.motion, p.radius, p.height) == (nbc.NBC_PENE_SPHERE, nbc.NBC_PENE_CONSTANT_VELOCITY, 0.02, 0.02)
    assert list(p.vel) == [0.0, 0.0, -0.05]
    g = nbc.Penetrator()
    assert nbc.lib().nbc_load_penetrator(os.path.join(FIX, "Rocks", "Penetrator_gravity.bt").encode(), C.byref(g)) == 0
    assert g.motion == nbc.NBC_PENE_GRAVITY and g.wait_for_contact == 0
    f = tmp_path / "p.bt"
    f.write_text("cylinder 0.01 0.05 0 0 1\n1 1 1 2\n0 0 0\n0 0 0\n0 0 0 1\n0 0 0\nmotion gravity_after_contact\n")
    q = nbc.Penetrator()
    assert nbc.lib().nbc_load_penetrator(str(f).encode(), C.byref(q)) == 0
    assert (q.shape, q.motion, q.wait_for_contact, q.prop[3]) == (nbc.NBC_PENE_CYLINDER, nbc.NBC_PENE_GRAVITY, 1, 2.0)
    f.write_text("torus 0.01\n")
    assert nbc.lib().nbc_load_penetrator(str(f).encode(), C.byref(q)) == nbc.NBC_E_BADINPUT
    f.write_text("sphere -1\n1 1 1 1\n0 0 0\n0 0 0\n0 0 0 1\n0 0 0\n")
    assert nbc.lib().nbc_load_penetrator(str(f).encode(), C.byref(q)) == nbc.NBC_E_BADINPUT
    f.write_text("sphere 0.01\n1 1 1 1\n0 0 0\n0 0 0\n0 0 0 1\n0 0 0\nmotion teleport\n")
    assert nbc.lib().nbc_load_penetrator(str(f).encode(), C.byref(q)) == nbc.NBC_E_BADINPUT
    f.write_text("cone 0.01 0.0173 32\n1 1 1 1\n0 0 0.1\n0 0 -0.1\n0 0 0 1\n0 0 0\n")
    k = nbc.Penetrator()
    assert nbc.lib().nbc_load_penetrator(str(f).encode(), C.byref(k)) == 0
    assert k.shape == nbc.NBC_PENE_CONE and k.ntri == 64
    tri = np.ctypeslib.as_array(k.tri, shape=(64, 3, 3))
    assert np.isclose(tri[:, :, 2].min(), -0.0173) and np.isclose(np.hypot(tri[:, :, 0], tri[:, :, 1]).max(), 0.01)
    nbc.lib().nbc_free_mesh(C.byref(k))


def test_binary_writers_layout(tmp_path):
    n = 5
    sid = np.arange(n, dtype=np.int32)
    fid = np.array([3, 1, 4, 0, 2], np.int32)
    cid = np.array([0, 0, 1, 1, 1], np.int32)
    pos = np.random.default_rng(0).random((n, 3))
    loc = np.random.default_rng(1).random((n, 3))
    m, r = np.arange(n) + 0.5, np.arange(n) + 0.25
    d = str(tmp_path / "OutPart")
    assert nbc.lib().nbc_write_particles_bin(d.encode(), 100, n, *[nbc.ptr(a) for a in (sid, fid, cid, pos, loc, m, r)]) == 0
    raw = open(os.path.join(d, "ph.00000100.bin"), "rb").read()
    assert len(raw) == 4 + 3 * 4 * n + 2 * 24 * n + 2 * 8 * n
    assert np.frombuffer(raw[:4], np.int32)[0] == n
    ints = np.frombuffer(raw[4:4 + 12 * n], np.int32).reshape(3, n)
    assert np.array_equal(ints[1], fid) and np.array_equal(ints[2], cid)
    dbl = np.frombuffer(raw[4 + 12 * n:], np.float64)
    assert np.array_equal(dbl[:3 * n].reshape(n, 3), pos) and np.array_equal(dbl[-n:], r)
    # clumps with the penetrator first (id -1, zero property)
    nc = 2
    cp, cv, cw = np.ones((nc, 3)), 2 * np.ones((nc, 3)), 3 * np.ones((nc, 3))
    cq, pr = 4 * np.ones((nc, 4)), 5 * np.ones((nc, 4))
    pp, pv, pq, pw = np.full(3, 6.0), np.full(3, 7.0), np.full(4, 8.0), np.full(3, 9.0)
    d2 = str(tmp_path / "OutClump")
    assert nbc.lib().nbc_write_clumps_bin(d2.encode(), 7, nc, *[nbc.ptr(a) for a in (cp, cv, cq, cw, pr, pp, pv, pq, pw)]) == 0
    raw = open(os.path.join(d2, "ch.00000007.bin"), "rb").read()
    hdr = np.frombuffer(raw[:16], np.int32)
    assert list(hdr) == [3, -1, 0, 1]
    body = np.frombuffer(raw[16:], np.float64)
    assert len(body) == 3 * 3 + 3 * 3 + 3 * 4 + 3 * 3 + 3 * 4
    assert list(body[:3]) == [6.0] * 3 and list(body[3:9]) == [1.0] * 6
    assert list(body[-12:-8]) == [0.0] * 4
    # force history: every 100th entry of the interval
    oi = 250
    f = np.arange(3 * oi, dtype=float).reshape(oi, 3)
    d3 = str(tmp_path / "OutForce")
    assert nbc.lib().nbc_write_force_history_bin(d3.encode(), 250, oi, nbc.ptr(f), nbc.ptr(-f)) == 0
    raw = open(os.path.join(d3, "fc.00000250.bin"), "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == 3
    assert list(np.frombuffer(raw[4:16], np.int32)) == [0, 100, 200]
    vals = np.frombuffer(raw[16:], np.float64).reshape(2, 3, 3)
    assert np.array_equal(vals[0], f[[0, 100, 200]]) and np.array_equal(vals[1], -f[[0, 100, 200]])
    # yi: only spheres with |Fn|^2 > 1e-28
    fn = np.zeros((n, 3))
    fn[1] = [0, 0, 1e-15]
    fn[3] = [1.0, 0, 0]
    assert nbc.lib().nbc_write_yi_bin(d3.encode(), 250, n, nbc.ptr(cid), nbc.ptr(fn), nbc.ptr(fn), nbc.ptr(fn), nbc.ptr(fn)) == 0
    raw = open(os.path.join(d3, "yi.00000250.bin"), "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == 1 and np.frombuffer(raw[4:8], np.int32)[0] == cid[3]
    # contacts
    B = 4
    nei = np.arange(n * B, dtype=np.int32)
    fc = np.ones((n * B, 3))
    d4 = str(tmp_path / "OutCnt")
    assert nbc.lib().nbc_write_contacts_bin(d4.encode(), 20, n, B, nbc.ptr(nei), nbc.ptr(fc), nbc.ptr(2 * fc)) == 0
    for pre, size in (("Nei", 8 + 4 * n * B), ("Fcn", 8 + 24 * n * B), ("Cpt", 8 + 24 * n * B)):
        raw = open(os.path.join(d4, f"{pre}.00000020.bin"), "rb").read()
        assert len(raw) == size and list(np.frombuffer(raw[:8], np.int32)) == [n, B]


def test_generator_is_deterministic_and_fp32_exact():
    a = nbc.generate_bed(1, 200, seed=42)
    b = nbc.generate_bed(1, 200, seed=42)
    c = nbc.generate_bed(1, 200, seed=43)
    assert np.array_equal(a.pos, b.pos) and not np.array_equal(a.pos, c.pos)
    for arr in (a.local, a.mass, a.radius):
        assert np.array_equal(arr.astype(np.float32).astype(np.float64), arr)
    assert a.radius.min() >= 0.6e-3 and a.radius.max() <= 1.6e-3
    # sub-sphere mass is an equal share of the clump mass
    for k in range(0, a.nc, 17):
        m = a.mass[a.clump_id == k]
        assert np.allclose(m.sum(), a.cprop[k, 3], rtol=1e-6)
    # deposited bed: no overlaps beyond rounding between different clumps
    from scipy.spatial import cKDTree
    t = cKDTree(a.pos)
    for i, j in t.query_pairs(2 * a.radius.max()):
        if a.clump_id[i] != a.clump_id[j]:
            assert np.linalg.norm(a.pos[i] - a.pos[j]) >= a.radius[i] + a.radius[j] - 1e-9
    # inside the cylinder
    assert np.all(np.hypot(a.pos[:, 0], a.pos[:, 1]) + a.radius <= a.wall.radius + 1e-9)


def test_save_bed_text_roundtrip(tmp_path):
    a = nbc.generate_bed(0, 40, seed=3)
    out = nbc.BedOut()
    spec = nbc.BedSpec(0, 40, 3, nbc.D4(0, 0, 0, 0))
    assert nbc.lib().nbc_generate_bed(C.byref(spec), C.byref(out)) == 0
    assert nbc.lib().nbc_save_bed_text(str(tmp_path).encode(), C.byref(out), None) == 0
    nbc.lib().nbc_free_bed(C.byref(out))
    rc, p, ini = load_params(str(tmp_path / "world.par"))
    assert rc == 0 and p.ParticleNum == a.n and p.ClumpNum == a.nc
    cid = np.zeros(a.n, np.int32)
    pos, loc, m, r = np.zeros((a.n, 3)), np.zeros((a.n, 3)), np.zeros(a.n), np.zeros(a.n)
    assert nbc.lib().nbc_load_phase(str(tmp_path / ini).encode(), a.n, *[nbc.ptr(x) for x in (cid, pos, loc, m, r)]) == 0
    assert np.array_equal(pos, a.pos) and np.array_equal(r, a.radius) and np.array_equal(cid, a.clump_id)


def test_autotune_argument_checks():
    bed = nbc.generate_bed(0, 20, seed=2)
    s, c = bed.spheres_in(), bed.clumps_in()
    best = nbc.Options()
    n = C.c_int32()
    p = bed.finalized_params()
    rc = nbc.lib().nbc_autotune(C.byref(p), None, C.byref(s), C.byref(c), None, None, 0, 10, 0, C.byref(best), None, 0,
                                C.byref(n))
    assert rc == nbc.NBC_E_BADINPUT and "before nbc_finalize_params" in nbc.lib().nbc_last_error(None).decode()
    if nbc.device_count() == 0:
        rc = nbc.lib().nbc_autotune(C.byref(bed.params), None, C.byref(s), C.byref(c), None, None, 0, 10, 0,
                                    C.byref(best), None, 0, C.byref(n))
        assert rc == nbc.NBC_E_NODEVICE


def test_nbc_par_keys(tmp_path):
    f = tmp_path / "nbc.par"
    f.write_text("Precision = mixed\nWallType = cylinder\nBlockSize = 128\nRadixBits = 6\nPairCapacity = 20\n"
                 "PeneFrcoe = 0.4\nCellSize = 0.006\nHGridLevels = 3\nNearSkin = 0.5\nGraphs = -1\n"
                 "BatchSteps = 8\nDriver = split\nHGridNested = 1\nSinglePass = -1\nNearHorizon = 250\n")
    o = nbc.Options()
    assert nbc.lib().nbc_load_options(str(f).encode(), C.byref(o)) == 0
    assert (o.precision, o.wall_type, o.block_size, o.radix_bits, o.pair_capacity) == \
        (nbc.NBC_MIXED, nbc.NBC_WALL_CYLINDER, 128, 6, 20)
    assert (o.pene_frcoe, o.cell_size, o.hgrid_levels, o.near_skin, o.graphs, o.batch_steps) == \
        (0.4, 0.006, 3, 0.5, -1, 8)
    assert o.driver == nbc.NBC_DRIVER_SPLIT and o.hgrid_nested == 1 and o.single_pass == -1
    assert o.near_horizon == 250.0
    for name, want in (("clump", 0), ("cube", 1), ("cylinder", 2), ("split", 3), ("hgrid", 4)):
        f.write_text(f"Driver = {name}\n")
        assert nbc.lib().nbc_load_options(str(f).encode(), C.byref(o)) == 0 and o.driver == want
    f.write_text("Driver = spheres\n")
    assert nbc.lib().nbc_load_options(str(f).encode(), C.byref(o)) == nbc.NBC_E_PARAM
    o2 = nbc.Options()
    assert nbc.lib().nbc_load_options(str(tmp_path / "absent.par").encode(), C.byref(o2)) == 0
    assert o2.precision == nbc.NBC_FP64 and o2.hgrid_levels == 0


def test_c2_bed_text_asks_for_the_hierarchical_grid(tmp_path):
    spec = nbc.BedSpec(2, 60, 1, nbc.D4(0.01, 0, 0, 0))
    out = nbc.BedOut()
    assert nbc.lib().nbc_generate_bed(C.byref(spec), C.byref(out)) == 0
    try:
        assert nbc.lib().nbc_save_bed_text(str(tmp_path).encode(), C.byref(out), None) == 0
    finally:
        nbc.lib().nbc_free_bed(C.byref(out))
    o = nbc.Options()
    assert nbc.lib().nbc_load_options(str(tmp_path / "nbc.par").encode(), C.byref(o)) == 0
    assert o.hgrid_levels == 3 and o.wall_type == nbc.NBC_WALL_CUBE
    rc, p, _ = load_params(str(tmp_path / "world.par"))
    assert rc == 0 and p.GridLevels == 3 and p.Frcoe == 0.6


def test_particle_text_and_binary_layout(tmp_path):
    # DEMmain ph text (21 columns) in, DEMSphereIo::SaveParticles_Bin layout out
    rng = np.random.default_rng(4)
    n = 7
    t = rng.integers(0, 3, n).astype(np.int32)
    cols = [rng.normal(size=(n, k)) for k in (3, 3, 4, 3, 4)] + [rng.uniform(1e-3, 2e-3, n)]
    rows = np.column_stack([np.arange(n), np.zeros(n), t] + cols)
    np.savetxt(tmp_path / "ph.txt", rows, fmt=["%d", "%d", "%d"] + ["%.17g"] * 18)
    t2 = np.zeros(n, np.int32)
    got = [np.zeros_like(c) for c in cols]
    assert nbc.lib().nbc_load_particles(str(tmp_path / "ph.txt").encode(), n, nbc.ptr(t2),
                                        *[nbc.ptr(g) for g in got]) == 0
    assert np.array_equal(t2, t) and all(np.array_equal(a, b) for a, b in zip(got, cols))
    # a short file is an I/O error naming the row
    assert nbc.lib().nbc_load_particles(str(tmp_path / "ph.txt").encode(), n + 1, nbc.ptr(np.zeros(n + 1, np.int32)),
                                        *[nbc.ptr(np.zeros((n + 1,) + g.shape[1:])) for g in got]) == nbc.NBC_E_IO
    assert b"row 7" in nbc.lib().nbc_last_error(None)
    sid, fid = np.arange(n, dtype=np.int32), rng.permutation(n).astype(np.int32)
    assert nbc.lib().nbc_write_spheres_bin(str(tmp_path).encode(), 5, n, nbc.ptr(sid), nbc.ptr(fid), nbc.ptr(t),
                                           *[nbc.ptr(c) for c in cols]) == 0
    raw = open(tmp_path / "ph.00000005.bin", "rb").read()
    assert len(raw) == 4 + 3 * 4 * n + 8 * n * 18
    assert np.frombuffer(raw[:4], np.int32)[0] == n
    ints = np.frombuffer(raw[4:4 + 12 * n], np.int32).reshape(3, n)
    assert np.array_equal(ints[0], sid) and np.array_equal(ints[1], fid) and np.array_equal(ints[2], t)
    dbl = np.frombuffer(raw[4 + 12 * n:], np.float64)
    assert np.array_equal(dbl, np.concatenate([c.ravel() for c in cols]))


def test_demmain_driver_argument_checks():
    bed = nbc.generate_bed(3, 50, seed=2, container=(0.01, 0, 0, 0))
    assert bed.n == bed.nc == 50 and np.all(bed.local == 0.0)
    part = nbc.Particles.from_bed(bed)
    assert np.array_equal(part.pos, bed.cpos) and np.all(part.type == 0)
    # solid spheres: I = 0.4 m r^2
    assert np.allclose(part.prop[:, 0], 0.4 * part.prop[:, 3] * part.radius ** 2, rtol=1e-6)
    with pytest.raises(ValueError):
        nbc.Particles.from_bed(nbc.generate_bed(0, 10, seed=1))
    if nbc.device_count() == 0:
        return
    opt = nbc.make_options(driver=nbc.NBC_DRIVER_HGRID)
    h = C.c_void_p()
    assert nbc.lib().nbc_create(C.byref(h), C.byref(part.finalized_params(opt)), C.byref(opt), 0) == \
        nbc.NBC_E_BADINPUT


def test_generator_cli_writes_a_run_directory(tmp_path):
    """nbc_generate (the command-line generator, the role of
    球簇生成器/generate_dem_bed.m) writes the same bed as nbc_generate_bed for
    the same kind, count and seed, a cone penetrator above the bed, and
    refuses to overwrite ph/ch files."""
    import subprocess

    exe = os.path.join(os.path.dirname(nbc.LIB_PATH), "nbc_generate")
    d = tmp_path / "run"
    r = subprocess.run([exe, "--kind", "1", "--clumps", "150", "--seed", "9", "--cone", "0.004", "0.0069", "16",
                        "0.1", "--out", str(d)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ref = nbc.generate_bed(1, 150, seed=9)
    p = nbc.Params()
    ini = C.create_string_buffer(256)
    assert nbc.lib().nbc_load_world_par(str(d / "world.par").encode(), C.byref(p), ini, 256) == 0
    assert (p.ParticleNum, p.ClumpNum, p.isIntruder) == (ref.n, ref.nc, 1)
    n = p.ParticleNum
    cid = np.zeros(n, np.int32)
    pos, loc, m, rad = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    assert nbc.lib().nbc_load_phase(str(d / "ph.00000000.bt").encode(), n,
                                    *[nbc.ptr(a) for a in (cid, pos, loc, m, rad)]) == 0
    assert np.array_equal(cid, ref.clump_id) and np.array_equal(pos, ref.pos) and np.array_equal(rad, ref.radius)
    pe = nbc.Penetrator()
    assert nbc.lib().nbc_load_penetrator(str(d / "Rocks" / "Penetrator.bt").encode(), C.byref(pe)) == 0
    assert pe.shape == nbc.NBC_PENE_CONE and pe.vel[2] == -0.1 and pe.ntri == 2 * 16
    # the cone's apex (z = -H in its frame) sits 20 um above the highest sphere under it
    rho = np.hypot(ref.pos[:, 0] - ref.wall.pos[0], ref.pos[:, 1] - ref.wall.pos[1])
    top = (ref.pos[rho < 0.004 + ref.radius, 2] + ref.radius[rho < 0.004 + ref.radius]).max()
    assert pe.pos[2] - pe.height == pytest.approx(top + 20e-6, abs=1e-12)
    nbc.lib().nbc_free_mesh(C.byref(pe))
    r2 = subprocess.run([exe, "--kind", "1", "--clumps", "150", "--out", str(d)], capture_output=True, text=True)
    assert r2.returncode == 1 and "refusing to overwrite" in r2.stderr


def test_resume_from_binary_outputs(tmp_path):
    """nbc_resume_from_bin turns the binary outputs of a step (sorted-order
    ph with file indices, ch with the penetrator record first) back into a
    run directory in the text formats: rows in file order, bit-exact values,
    StartStep = step, the penetrator at its recorded state; it refuses to
    overwrite and rejects a truncated output."""
    import subprocess

    exe = os.path.join(os.path.dirname(nbc.LIB_PATH), "nbc_generate")
    run = tmp_path / "run"
    r = subprocess.run([exe, "--kind", "1", "--clumps", "120", "--seed", "4", "--cone", "0.004", "0.0069", "12",
                        "0.1", "--out", str(run)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    bed = nbc.generate_bed(1, 120, seed=4)
    n, nc = bed.n, bed.nc
    rng = np.random.default_rng(1)
    order = rng.permutation(n).astype(np.int32)  # sorted position k holds file sphere order[k]
    pos = bed.pos + rng.normal(0, 1e-4, bed.pos.shape)
    cvel = rng.normal(0, 0.1, (nc, 3))
    cangv = rng.normal(0, 10.0, (nc, 3))
    cpos = bed.cpos + 1e-4
    ppos, pvel = np.array([0.001, -0.002, 0.03]), np.array([0.0, 0.0, -0.1])
    pquat, pangv = np.array([0.0, 0.0, 0.0, 1.0]), np.zeros(3)
    step = 40
    sid = np.arange(n, dtype=np.int32)
    cid = np.ascontiguousarray(bed.clump_id[order])
    args = [np.ascontiguousarray(a) for a in (pos[order], bed.local[order], bed.mass[order], bed.radius[order])]
    assert nbc.lib().nbc_write_particles_bin(str(run / "OutPart").encode(), step, n, nbc.ptr(sid), nbc.ptr(order),
                                             nbc.ptr(cid), *[nbc.ptr(a) for a in args]) == 0
    assert nbc.lib().nbc_write_clumps_bin(str(run / "OutClump").encode(), step, nc,
                                          *[nbc.ptr(np.ascontiguousarray(a)) for a in
                                            (cpos, cvel, bed.cquat, cangv, bed.cprop, ppos, pvel, pquat, pangv)]) == 0
    out = tmp_path / "resume"
    assert nbc.lib().nbc_resume_from_bin(str(run).encode(), step, str(out).encode()) == 0
    p = nbc.Params()
    ini = C.create_string_buffer(256)
    assert nbc.lib().nbc_load_world_par(str(out / "world.par").encode(), C.byref(p), ini, 256) == 0
    q = nbc.Params()
    assert nbc.lib().nbc_load_world_par(str(run / "world.par").encode(), C.byref(q), ini, 256) == 0
    assert p.StartStep == step and p.ParticleNum == n and p.StepSize == q.StepSize and p.Thresh == q.Thresh
    c2 = np.zeros(n, np.int32)
    x2, l2, m2, r2 = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    assert nbc.lib().nbc_load_phase(str(out / ini.value.decode()).encode(), n,
                                    *[nbc.ptr(a) for a in (c2, x2, l2, m2, r2)]) == 0
    assert np.array_equal(c2, bed.clump_id) and np.array_equal(x2, pos) and np.array_equal(l2, bed.local)
    assert np.array_equal(m2, bed.mass) and np.array_equal(r2, bed.radius)
    cp, cv, cq, cw, cr = np.zeros((nc, 3)), np.zeros((nc, 3)), np.zeros((nc, 4)), np.zeros((nc, 3)), np.zeros((nc, 4))
    assert nbc.lib().nbc_load_clumps(str(out / "ch.00000000.bt").encode(), nc,
                                     *[nbc.ptr(a) for a in (cp, cv, cq, cw, cr)]) == 0
    assert np.array_equal(cp, cpos) and np.array_equal(cv, cvel) and np.array_equal(cw, cangv)
    assert np.array_equal(cq, bed.cquat) and np.array_equal(cr, bed.cprop)
    pe, pe0 = nbc.Penetrator(), nbc.Penetrator()
    assert nbc.lib().nbc_load_penetrator(str(out / "Rocks" / "Penetrator.bt").encode(), C.byref(pe)) == 0
    assert nbc.lib().nbc_load_penetrator(str(run / "Rocks" / "Penetrator.bt").encode(), C.byref(pe0)) == 0
    assert list(pe.pos) == ppos.tolist() and list(pe.vel) == pvel.tolist()
    assert (pe.shape, pe.radius, pe.height, pe.ntri, pe.motion) == (pe0.shape, pe0.radius, pe0.height, pe0.ntri,
                                                                     pe0.motion)
    assert list(pe.prop) == list(pe0.prop)
    nbc.lib().nbc_free_mesh(C.byref(pe))
    nbc.lib().nbc_free_mesh(C.byref(pe0))
    assert (out / "Rocks" / "CylindWall.bt").read_bytes() == (run / "Rocks" / "CylindWall.bt").read_bytes()
    assert (out / "nbc.par").read_bytes() == (run / "nbc.par").read_bytes()
    # refuses to overwrite; a truncated output is an I/O error
    assert nbc.lib().nbc_resume_from_bin(str(run).encode(), step, str(out).encode()) == nbc.NBC_E_IO
    f = run / "OutPart" / "ph.00000040.bin"
    f.write_bytes(f.read_bytes()[:-8])
    assert nbc.lib().nbc_resume_from_bin(str(run).encode(), step, str(tmp_path / "r2").encode()) == nbc.NBC_E_IO
    assert "truncated" in nbc.lib().nbc_last_error(None).decode()

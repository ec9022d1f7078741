"""The drop-in entry (nbc_run_simulation_clump, the RunSimulationClump
replacement) on the reference's own shipped inputs, on the GPU: output files,
their layout, and the state they hold against the FP64 oracle run on the same
inputs."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

import nbc
import orc

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")


def prepare(tmp_path, end_step, interval, extra=None):
    d = tmp_path / "run"
    shutil.copytree(FIX, d)
    txt = open(d / "world.par").read()
    txt = txt.replace("EndStep\t\t= 200000", f"EndStep\t\t= {end_step}")
    txt = txt.replace("OutputInterval\t= 10000", f"OutputInterval\t= {interval}")
    assert f"= {end_step}" in txt and f"= {interval}" in txt
    (d / "world.par").write_text(txt)
    if extra:
        (d / "nbc.par").write_text(extra)
    return d


def read_ph(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    ints = np.frombuffer(raw[4:4 + 12 * n], np.int32).reshape(3, n)
    dbl = np.frombuffer(raw[4 + 12 * n:], np.float64)
    pos = dbl[:3 * n].reshape(n, 3)
    return n, ints[1], ints[2], pos


def oracle_for_fixture(steps, d=FIX):
    d = str(d)
    p = nbc.Params()
    ini = C.create_string_buffer(256)
    assert nbc.lib().nbc_load_world_par(os.path.join(d, "world.par").encode(), C.byref(p), ini, 256) == 0
    n, nc = p.ParticleNum, p.ClumpNum
    cid = np.zeros(n, np.int32)
    pos, loc, m, r = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    assert nbc.lib().nbc_load_phase(os.path.join(d, ini.value.decode()).encode(), n,
                                    *[nbc.ptr(a) for a in (cid, pos, loc, m, r)]) == 0
    cp, cv, cq, cw, pr = np.zeros((nc, 3)), np.zeros((nc, 3)), np.zeros((nc, 4)), np.zeros((nc, 3)), np.zeros((nc, 4))
    assert nbc.lib().nbc_load_clumps(os.path.join(d, "ch.00000000.bt").encode(), nc,
                                     *[nbc.ptr(a) for a in (cp, cv, cq, cw, pr)]) == 0
    wall = nbc.Wall()
    assert nbc.lib().nbc_load_cube_wall(os.path.join(d, "Rocks", "CubeWall.bt").encode(), C.byref(wall)) == 0
    pene = nbc.Penetrator()
    assert nbc.lib().nbc_load_penetrator(os.path.join(d, "Rocks", "Penetrator.bt").encode(), C.byref(pene)) == 0
    bed = nbc.Bed(p, cid, pos, loc, m, r, cp, cv, cq, cw, pr, wall)
    o = orc.Oracle(orc.finalize_params(bed))
    o.upload(bed, wall, pene)
    o.step(steps)
    return o, bed


def test_run_simulation_clump_on_reference_inputs(tmp_path):
    print("[case] drop-in run on the shipped inputs, 300 steps", flush=True)
    d = prepare(tmp_path, 300, 150)
    rc = nbc.lib().nbc_run_simulation_clump(str(d).encode(), 0)
    assert rc == 0, nbc.lib().nbc_last_error(None)
    for f in ("OutPart/ph.00000150.bin", "OutPart/ph.00000300.bin", "OutClump/ch.00000300.bin",
              "OutForce/fc.00000300.bin", "OutForce/yi.00000300.bin", "OutCnt/Nei.00000300.bin",
              "OutCnt/Fcn.00000300.bin", "OutCnt/Cpt.00000300.bin", "OutClump/PenetratorShape.bt"):
        assert (d / f).exists(), f
    assert open(d / "OutClump/PenetratorShape.bt").read().split()[0] == "sphere"
    n, fid, cid, pos = read_ph(d / "OutPart/ph.00000300.bin")
    assert n == 576
    o, bed = oracle_for_fixture(300)
    so = o.spheres()
    assert np.array_equal(fid, so["fid"]) and np.array_equal(cid, so["cid"])
    assert np.abs(pos - so["pos"]).max() <= 1e-12
    # Nei rows: front neighbours first, back ones at the end of each row
    raw = open(d / "OutCnt/Nei.00000300.bin", "rb").read()
    N, B = np.frombuffer(raw[:8], np.int32)
    nei = np.frombuffer(raw[8:], np.int32).reshape(N, B)
    pr = o.pairs()
    for i in range(0, N, 29):
        ng = pr["foff"][i + 1] - pr["foff"][i]
        assert np.array_equal(nei[i, :ng], pr["fj"][pr["foff"][i]:pr["foff"][i + 1]])
        nb = pr["boff"][i + 1] - pr["boff"][i]
        if nb:
            assert np.array_equal(nei[i, B - nb:], pr["bi"][pr["boff"][i]:pr["boff"][i + 1]])
        assert np.all(nei[i, ng:B - nb] == -1)
    # Fcn / Cpt rows (src/dem_force_particle.cu:238-268): front entries +CacheFcn,
    # back entries -CacheFcn of the owner's slot, contact points as stored
    fcn, cpt = o.pair_exports()
    rows = {}
    for name in ("Fcn", "Cpt"):
        raw = open(d / f"OutCnt/{name}.00000300.bin", "rb").read()
        assert tuple(np.frombuffer(raw[:8], np.int32)) == (N, B)
        rows[name] = np.frombuffer(raw[8:], np.float64).reshape(N, B, 3)
    ef, ec = np.zeros((N, B, 3)), np.zeros((N, B, 3))
    for i in range(N):
        a, b = pr["foff"][i], pr["foff"][i + 1]
        ef[i, :b - a], ec[i, :b - a] = fcn[a:b], cpt[a:b]
        a, b = pr["boff"][i], pr["boff"][i + 1]
        if b > a:
            sl = pr["bslot"][a:b]
            assert np.all(sl >= 0)
            ef[i, B - (b - a):], ec[i, B - (b - a):] = -fcn[sl], cpt[sl]
    scale = np.abs(fcn).max()
    assert scale > 0, "no touching pair at the output step"
    assert np.abs(rows["Fcn"] - ef).max() <= 1e-12 * scale
    assert np.abs(rows["Cpt"] - ec).max() <= 1e-12


def test_mixed_precision_run_and_overflow_error(tmp_path):
    print("[case] drop-in mixed run + tiny BranchLmtLen overflow", flush=True)
    d = prepare(tmp_path, 100, 100, extra="Precision = mixed\n")
    assert nbc.lib().nbc_run_simulation_clump(str(d).encode(), 0) == 0
    d2 = prepare(tmp_path / "b", 10, 10)
    txt = open(d2 / "world.par").read().replace("BranchLmtLen         = 256", "BranchLmtLen         = 2")
    (d2 / "world.par").write_text(txt)
    rc = nbc.lib().nbc_run_simulation_clump(str(d2).encode(), 0)
    assert rc == nbc.NBC_E_NEIGHBOR_OVERFLOW
    assert "Neighbor list overflow (BranchLmtLen=2)" in nbc.lib().nbc_last_error(None).decode()


def test_resume_from_binary_outputs(tmp_path):
    """A run continued from its own binary outputs: the shipped inputs run 150
    steps, nbc_resume_from_bin turns the step-150 outputs into a run
    directory (StartStep 150), which runs to step 300; its step-300 output
    equals the FP64 oracle started from the same converted files (contact
    histories start from zero there, as the outputs do not hold them)."""
    print("[case] resume from the step-150 outputs of the shipped inputs", flush=True)
    d = prepare(tmp_path, 150, 150)
    assert nbc.lib().nbc_run_simulation_clump(str(d).encode(), 0) == 0, nbc.lib().nbc_last_error(None)
    r = tmp_path / "resumed"
    assert nbc.lib().nbc_resume_from_bin(str(d).encode(), 150, str(r).encode()) == 0, nbc.lib().nbc_last_error(None)
    txt = (r / "world.par").read_text()
    assert "StartStep = 150" in txt
    (r / "world.par").write_text(txt.replace("EndStep\t\t= 150", "EndStep\t\t= 300"))
    assert nbc.lib().nbc_run_simulation_clump(str(r).encode(), 0) == 0, nbc.lib().nbc_last_error(None)
    assert not (r / "OutPart/ph.00000150.bin").exists() and (r / "OutPart/ph.00000300.bin").exists()
    # the resumed run starts where the first one stopped
    n0, fid0, _, pos0 = read_ph(d / "OutPart/ph.00000150.bin")
    o0, bed0 = oracle_for_fixture(0, r)
    x0 = np.zeros_like(pos0)
    x0[fid0] = pos0
    assert np.array_equal(bed0.pos, x0)
    o0.close()
    n, fid, cid, pos = read_ph(r / "OutPart/ph.00000300.bin")
    o, bed = oracle_for_fixture(150, r)
    so = o.spheres()
    assert np.array_equal(fid, so["fid"]) and np.array_equal(cid, so["cid"])
    assert np.abs(pos - so["pos"]).max() <= 1e-12

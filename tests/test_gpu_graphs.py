"""HIP graph replays of steady chunks: a context with graphs on must give
bit-identical FP64 results to one with graphs off, call after call (several
replays per call, eager rebuilds, near-list refreshes and synchronisations
between them), on a single bed and on a batch of beds."""
import numpy as np
import pytest

import nbc

pytestmark = pytest.mark.gpu


def shaken(kind, nc, seed, speed, dt, **kw):
    b = nbc.generate_bed(kind, nc, seed=seed, **kw)
    rng = np.random.default_rng(seed)
    b.cvel[:] = rng.uniform(-speed, speed, b.cvel.shape)
    b.params.StepSize = dt
    return b


@pytest.mark.parametrize("prec", [nbc.NBC_FP64, nbc.NBC_MIXED])
def test_graph_replays_match_eager_steps(prec):
    print(f"[case] graph replays vs eager, 300-clump C1 bed, precision {prec}", flush=True)
    bed = shaken(1, 300, 3, 0.4, 5e-6)
    ctxs = []
    for g in (-1, 1):
        opt = nbc.make_options(precision=prec, wall_type=nbc.NBC_WALL_CYLINDER, graphs=g)
        c = nbc.Context(bed.finalized_params(opt), opt)
        c.upload(bed, bed.wall)
        ctxs.append(c)
    seen = 0
    for call in range(8):
        for c in ctxs:
            c.step(37)
        a, b = ctxs[0], ctxs[1]
        sa, sb = a.stats(), b.stats()
        seen = max(seen, sb.touching_ss + sb.touching_wall)
        assert (sa.rebuilds, sa.touching_ss, sa.touching_wall, sa.near_refreshes) == \
            (sb.rebuilds, sb.touching_ss, sb.touching_wall, sb.near_refreshes), call
        assert np.array_equal(a.spheres()["pos"], b.spheres()["pos"]), call
        ca, cb = a.clumps(), b.clumps()
        assert np.array_equal(ca["vel"], cb["vel"]) and np.array_equal(ca["angv"], cb["angv"]), call
    sa, sb = ctxs[0].stats(), ctxs[1].stats()
    assert sa.graph_replays == 0 and sb.graph_replays > 5
    assert sb.rebuilds >= 2 and seen > 0
    for c in ctxs:
        c.close()


def test_graph_replays_in_a_batch():
    print("[case] graph replays vs eager, batch of 3 beds", flush=True)
    beds = [shaken(1, nc, s, 0.5, 1e-5, container=(R, 0, 0, 0))
            for nc, s, R in ((150, 41, 0.009), (220, 42, 0.011), (90, 43, 0.007))]
    ctxs = []
    for g in (-1, 1):
        opt = nbc.make_options(precision=nbc.NBC_FP64, wall_type=nbc.NBC_WALL_CYLINDER, graphs=g)
        c = nbc.BatchContext([b.finalized_params(opt) for b in beds], opt)
        c.upload(beds)
        ctxs.append(c)
    for call in range(5):
        for c in ctxs:
            c.step(41)
        assert np.array_equal(ctxs[0].spheres()["pos"], ctxs[1].spheres()["pos"]), call
    assert ctxs[1].stats().graph_replays > 3 and ctxs[1].stats().rebuilds >= 2
    for c in ctxs:
        c.close()

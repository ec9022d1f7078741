"""Helpers to build small hand-made beds for tests (host-side data only)."""
import numpy as np

import nbc


def make_params(n, nc, **kw):
    p = nbc.Params()
    p.ParticleNum, p.ClumpNum = n, nc
    p.BranchLmtLen = kw.get("BranchLmtLen", 64)
    p.StartStep, p.EndStep = 0, kw.get("EndStep", 100)
    p.OutputInterval = kw.get("OutputInterval", 100)
    p.StepSize = kw.get("StepSize", 1e-6)
    p.Ymod = kw.get("Ymod", 1e8)
    p.Prat = kw.get("Prat", 0.3)
    p.Frcoe = kw.get("Frcoe", 0.5)
    p.NormRes = kw.get("NormRes", 0.5)
    p.Frvel = 5e-4
    p.Beta = kw.get("Beta", 0.5)
    p.MU_T = p.MU_R = 0.5
    p.Cohesion = kw.get("Cohesion", 0.0)
    p.Gravity = kw.get("Gravity", 0.0)
    p.Thresh = kw.get("Thresh", 1e-4)
    p.GridLevels = 1
    p.Update = 0
    lo = kw.get("lo", (-0.05, -0.05, -0.05))
    size = kw.get("size", (0.1, 0.1, 0.1))
    for a in range(3):
        p.WorldOrigin[a] = lo[a]
        p.WorldSize[a] = size[a]
        p.BoxLowLmt[a] = lo[a]
        p.BoxUpLmt[a] = lo[a] + size[a]
    p.isIntruder = kw.get("isIntruder", 0)
    p.isWallForce = kw.get("isWallForce", 0)
    p.isSelfGravity = 0
    p.isQuaternion = 1
    return p


def single_sphere_clumps(centres, radii, masses=None, vel=None, angv=None, quat=None, inertia=None, **kw):
    """One sphere per clump, local offset zero (the per-sphere case)."""
    centres = np.asarray(centres, float).reshape(-1, 3)
    n = len(centres)
    radii = np.broadcast_to(np.asarray(radii, float), (n,)).copy()
    masses = np.broadcast_to(np.asarray(2650.0 * 4.0 / 3.0 * np.pi * radii ** 3 if masses is None else masses, float),
                             (n,)).copy()
    vel = np.zeros((n, 3)) if vel is None else np.asarray(vel, float).reshape(n, 3)
    angv = np.zeros((n, 3)) if angv is None else np.asarray(angv, float).reshape(n, 3)
    quat = np.tile([0.0, 0.0, 0.0, 1.0], (n, 1)) if quat is None else np.asarray(quat, float).reshape(n, 4)
    if inertia is None:
        inertia = 0.4 * masses * radii ** 2
    prop = np.column_stack([inertia, inertia, inertia, masses])
    p = make_params(n, n, **kw)
    wall = kw.get("wall")
    return nbc.Bed(p, np.arange(n, dtype=np.int32), centres, np.zeros((n, 3)), masses, radii, centres.copy(), vel,
                   quat, angv, prop, wall)


def cube_wall(half, centre):
    w = nbc.Wall()
    w.type = nbc.NBC_WALL_CUBE
    w.side[:] = half
    w.pos[:] = centre
    w.quat[:] = (0, 0, 0, 1)
    return w


def cylinder_wall(radius, half_height, centre, axis=(0, 0, 1)):
    w = nbc.Wall()
    w.type = nbc.NBC_WALL_CYLINDER
    w.radius, w.height = radius, half_height
    w.dir[:] = axis
    w.prop[:] = (1, 1, 1, 1)
    w.pos[:] = centre
    w.quat[:] = (0, 0, 0, 1)
    return w


def quat_to_matrix(q):
    """Textbook body->world rotation of the unit quaternion (x, y, z, w)."""
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def rodrigues(axis, angle):
    k = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def check_mixed_history(ctx, ref, mu):
    """FP32-mixed pair histories against the oracle (needs a recorded step).

    The tangential history is discontinuous where the Coulomb cap switches
    (src/dem_force_clump.cu:100-110): a slipping pair stores the capped force
    tf, a sticking one tf - Cs vt. A pair whose |tf| lies within FP32 rounding
    of mu |Fn| can take the other branch, so pairs at the cap on either side
    (|h_t| = mu |Fn| to 1e-5) are counted apart; every other pair holds 1e-4
    of the largest history. The rolling and twisting histories are continuous
    at their caps."""
    ha, hb = ctx.pair_history(), ref.pair_history()
    fa, _ = ctx.pair_exports()
    fb, _ = ref.pair_exports()
    na, nb = np.linalg.norm(fa, axis=1), np.linalg.norm(fb, axis=1)
    ta, tb = np.linalg.norm(ha[:, :3], axis=1), np.linalg.norm(hb[:, :3], axis=1)
    at_cap = ((nb > 0) & (np.abs(tb - mu * nb) <= 1e-5 * mu * nb)) | ((na > 0) & (np.abs(ta - mu * na) <= 1e-5 * mu * na))
    scale = max(np.abs(hb).max(), 1e-300)
    err = np.abs(ha - hb).max(axis=1) / scale
    flips = int(np.count_nonzero(at_cap & (err > 1e-4)))
    e_rest = err[~at_cap].max() if np.any(~at_cap) else 0.0
    print(f"  mixed history: {int(at_cap.sum())} pairs at the Coulomb cap, {flips} took the other branch; "
          f"other pairs max rel err {e_rest:.2e}", flush=True)
    assert e_rest <= 1e-4
    assert flips <= max(2, at_cap.sum() // 100)

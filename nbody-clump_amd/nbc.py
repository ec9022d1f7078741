"""ctypes binding of libnbc.so (include/nbc.h) for tests, bench and scripts.

The library is the product: this module only marshals numpy arrays across
the C ABI. It raises if libnbc.so is missing — there is no fallback path.
Structures mirror include/nbc.h field for field.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbc.so")

NBC_OK, NBC_E_BADINPUT, NBC_E_NEIGHBOR_OVERFLOW, NBC_E_NONFINITE, NBC_E_HIP, NBC_E_NODEVICE, NBC_E_IO, NBC_E_PARAM = range(8)
NBC_WALL_NONE, NBC_WALL_CUBE, NBC_WALL_CYLINDER = 0, 1, 2
NBC_MAX_LEVELS = 8
NBC_FP64, NBC_MIXED = 0, 1
NBC_PENE_CYLINDER, NBC_PENE_SPHERE, NBC_PENE_CONE = 0, 1, 2
NBC_PENE_CONSTANT_VELOCITY, NBC_PENE_GRAVITY = 0, 1
NBC_DRIVER_CLUMP, NBC_DRIVER_CUBE, NBC_DRIVER_CYLINDER, NBC_DRIVER_SPLIT, NBC_DRIVER_HGRID = range(5)

D3 = C.c_double * 3
D4 = C.c_double * 4


class Params(C.Structure):
    _fields_ = [
        ("ParticleNum", C.c_int32), ("ClumpNum", C.c_int32), ("BranchLmtLen", C.c_int32),
        ("StartStep", C.c_int32), ("EndStep", C.c_int32), ("OutputInterval", C.c_int32),
        ("StepSize", C.c_double),
        ("Ymod", C.c_double), ("Prat", C.c_double), ("Frcoe", C.c_double), ("NormRes", C.c_double),
        ("Frvel", C.c_double), ("Beta", C.c_double), ("MU_T", C.c_double), ("MU_R", C.c_double),
        ("Cohesion", C.c_double), ("Gravity", C.c_double),
        ("Thresh", C.c_double), ("GridLevels", C.c_double), ("Update", C.c_double),
        ("WorldSize", D3), ("WorldOrigin", D3), ("BoxUpLmt", D3), ("BoxLowLmt", D3),
        ("isIntruder", C.c_int32), ("isSelfGravity", C.c_int32), ("isWallForce", C.c_int32),
        ("isQuaternion", C.c_int32),
        ("Damping", C.c_double), ("CellSize", D3), ("i_CellSize", D3), ("GridDim", C.c_uint32 * 3),
        ("CellNum", C.c_int32), ("finalized", C.c_int32),
        ("Levels", C.c_int32), ("LevelOff", C.c_int32 * NBC_MAX_LEVELS),
        ("LevelDim", (C.c_int32 * 3) * NBC_MAX_LEVELS), ("LevelCell", C.c_double * NBC_MAX_LEVELS),
        ("LevelRmin", C.c_double * NBC_MAX_LEVELS),
        ("LevelNested", C.c_int32), ("LevelBlock", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("wall_type", C.c_int32), ("block_size", C.c_int32),
        ("radix_bits", C.c_int32), ("pair_capacity", C.c_int32), ("batch_steps", C.c_int32),
        ("hgrid_levels", C.c_int32), ("graphs", C.c_int32),
        ("pene_frcoe", C.c_double), ("cell_size", C.c_double), ("near_skin", C.c_double),
        ("driver", C.c_int32), ("level_from_type", C.c_int32), ("hgrid_nested", C.c_int32),
        ("single_pass", C.c_int32), ("near_horizon", C.c_double),
    ]


class TuneEntry(C.Structure):
    _fields_ = [("opt", Options), ("ms_per_step", C.c_double)]


NBC_TUNE_CELL, NBC_TUNE_HGRID = 1, 2


class Wall(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("reserved", C.c_int32), ("side", D3), ("radius", C.c_double),
        ("height", C.c_double), ("dir", D3), ("prop", D4), ("pos", D3), ("vel", D3), ("quat", D4),
        ("angv", D3),
    ]


class Penetrator(C.Structure):
    _fields_ = [
        ("shape", C.c_int32), ("motion", C.c_int32), ("wait_for_contact", C.c_int32),
        ("contact_state", C.c_int32), ("radius", C.c_double), ("height", C.c_double), ("dir", D3),
        ("prop", D4), ("pos", D3), ("vel", D3), ("quat", D4), ("angv", D3), ("ntri", C.c_int32),
        ("reserved", C.c_int32), ("tri", C.POINTER(C.c_double)),
    ]


class SpheresIn(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("reserved", C.c_int32), ("clump_id", C.POINTER(C.c_int32)),
        ("pos", C.POINTER(C.c_double)), ("local", C.POINTER(C.c_double)),
        ("mass", C.POINTER(C.c_double)), ("radius", C.POINTER(C.c_double)),
    ]


class ClumpsIn(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("reserved", C.c_int32), ("pos", C.POINTER(C.c_double)),
        ("vel", C.POINTER(C.c_double)), ("quat", C.POINTER(C.c_double)),
        ("angv", C.POINTER(C.c_double)), ("prop", C.POINTER(C.c_double)),
    ]


class ParticlesIn(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("reserved", C.c_int32), ("type", C.POINTER(C.c_int32)),
        ("pos", C.POINTER(C.c_double)), ("vel", C.POINTER(C.c_double)), ("quat", C.POINTER(C.c_double)),
        ("angv", C.POINTER(C.c_double)), ("prop", C.POINTER(C.c_double)), ("radius", C.POINTER(C.c_double)),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("rebuilds", C.c_int64), ("front_pairs", C.c_int64),
        ("back_entries", C.c_int64), ("touching_ss", C.c_int64), ("touching_wall", C.c_int64),
        ("kinetic_energy", C.c_double), ("pene_force", D3), ("pene_torque", D3),
        ("max_cum_disp", C.c_double), ("near_refreshes", C.c_int64), ("near_pairs", C.c_int64),
        ("graph_replays", C.c_int64), ("single_pass_rebuilds", C.c_int64),
    ]


class BedSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nclumps", C.c_int32), ("seed", C.c_uint64), ("container", D4)]


class BedOut(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("nc", C.c_int32), ("clump_id", C.POINTER(C.c_int32)),
        ("pos", C.POINTER(C.c_double)), ("local", C.POINTER(C.c_double)),
        ("mass", C.POINTER(C.c_double)), ("radius", C.POINTER(C.c_double)),
        ("cpos", C.POINTER(C.c_double)), ("cvel", C.POINTER(C.c_double)),
        ("cquat", C.POINTER(C.c_double)), ("cangv", C.POINTER(C.c_double)),
        ("cprop", C.POINTER(C.c_double)), ("wall", Wall), ("params", Params),
    ]


# Every exported symbol of include/nbc.h, with its argtypes.
P = C.c_void_p
I32 = C.c_int32
_SIGS = {
    "nbc_version": ([], C.c_char_p),
    "nbc_last_error": ([P], C.c_char_p),
    "nbc_device_count": ([C.POINTER(I32)], C.c_int),
    "nbc_load_world_par": ([C.c_char_p, C.POINTER(Params), C.c_char_p, C.c_size_t], C.c_int),
    "nbc_finalize_params": ([C.POINTER(Params), P, I32, C.POINTER(Options)], C.c_int),
    "nbc_load_options": ([C.c_char_p, C.POINTER(Options)], C.c_int),
    "nbc_load_grid_levels": ([C.c_char_p, C.POINTER(Params)], C.c_int),
    "nbc_load_phase": ([C.c_char_p, I32, P, P, P, P, P], C.c_int),
    "nbc_load_clumps": ([C.c_char_p, I32, P, P, P, P, P], C.c_int),
    "nbc_load_particles": ([C.c_char_p, I32, P, P, P, P, P, P, P], C.c_int),
    "nbc_load_cube_wall": ([C.c_char_p, C.POINTER(Wall)], C.c_int),
    "nbc_load_cylinder_wall": ([C.c_char_p, C.POINTER(Wall)], C.c_int),
    "nbc_load_penetrator": ([C.c_char_p, C.POINTER(Penetrator)], C.c_int),
    "nbc_resume_from_bin": ([C.c_char_p, C.c_int32, C.c_char_p], C.c_int),
    "nbc_free_mesh": ([C.POINTER(Penetrator)], None),
    "nbc_make_cone_mesh": ([C.c_double, C.c_double, I32, P], C.c_int),
    "nbc_generate_bed": ([C.POINTER(BedSpec), C.POINTER(BedOut)], C.c_int),
    "nbc_free_bed": ([C.POINTER(BedOut)], None),
    "nbc_save_bed_text": ([C.c_char_p, C.POINTER(BedOut), C.POINTER(Penetrator)], C.c_int),
    "nbc_create": ([C.POINTER(P), C.POINTER(Params), C.POINTER(Options), I32], C.c_int),
    "nbc_create_batch": ([C.POINTER(P), I32, C.POINTER(Params), C.POINTER(Options), I32], C.c_int),
    "nbc_upload_batch": ([P, C.POINTER(SpheresIn), C.POINTER(ClumpsIn), C.POINTER(Wall)], C.c_int),
    "nbc_get_batch_ranges": ([P, P, P], C.c_int),
    "nbc_destroy": ([P], C.c_int),
    "nbc_upload": ([P, C.POINTER(SpheresIn), C.POINTER(ClumpsIn), C.POINTER(Wall), C.POINTER(Penetrator)], C.c_int),
    "nbc_upload_particles": ([P, C.POINTER(ParticlesIn), C.POINTER(Wall), C.POINTER(Penetrator)], C.c_int),
    "nbc_get_particles": ([P, P, P, P, P, P, P, P, P, P], C.c_int),
    "nbc_step": ([P, I32], C.c_int),
    "nbc_rebuild": ([P], C.c_int),
    "nbc_set_record": ([P, I32], C.c_int),
    "nbc_synchronize": ([P], C.c_int),
    "nbc_get_stats": ([P, C.POINTER(Stats)], C.c_int),
    "nbc_get_spheres": ([P, P, P, P, P, P, P, P], C.c_int),
    "nbc_get_grid": ([P, P, P, P, P], C.c_int),
    "nbc_get_pair_counts": ([P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "nbc_get_pairs": ([P, P, P, P, P, P, P], C.c_int),
    "nbc_get_pair_history": ([P, P], C.c_int),
    "nbc_get_sphere_forces": ([P, I32, P, P], C.c_int),
    "nbc_get_clumps": ([P, P, P, P, P, P, P], C.c_int),
    "nbc_get_penetrator": ([P, P, P, P, P, P, P, P], C.c_int),
    "nbc_get_pene_exports": ([P, P, P, P, P], C.c_int),
    "nbc_get_pair_exports": ([P, P, P], C.c_int),
    "nbc_grid_build": ([C.POINTER(Params), P, I32, P, P, P, P, P], C.c_int),
    "nbc_grid_build_noalloc": ([C.POINTER(Params), P, I32, P, P, P, P, P], C.c_int),
    "nbc_profile_enable": ([P, I32], C.c_int),
    "nbc_profile_get": ([P, I32, C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int),
    "nbc_write_particles_bin": ([C.c_char_p, I32, I32, P, P, P, P, P, P, P], C.c_int),
    "nbc_write_spheres_bin": ([C.c_char_p, I32, I32, P, P, P, P, P, P, P, P, P], C.c_int),
    "nbc_write_clumps_bin": ([C.c_char_p, I32, I32, P, P, P, P, P, P, P, P, P], C.c_int),
    "nbc_write_intruder_bin": ([C.c_char_p, I32, P, P, P, P], C.c_int),
    "nbc_write_force_history_bin": ([C.c_char_p, I32, I32, P, P], C.c_int),
    "nbc_write_yi_bin": ([C.c_char_p, I32, I32, P, P, P, P, P], C.c_int),
    "nbc_write_contacts_bin": ([C.c_char_p, I32, I32, I32, P, P, P], C.c_int),
    "nbc_run_simulation_clump": ([C.c_char_p, I32], C.c_int),
    "nbc_run_simulation": ([C.c_char_p, I32, I32], C.c_int),
    "nbc_autotune_batch": ([I32, C.POINTER(Params), C.POINTER(Options), C.POINTER(SpheresIn), C.POINTER(ClumpsIn),
                            C.POINTER(Wall), I32, I32, I32, C.POINTER(Options), C.POINTER(TuneEntry), I32,
                            C.POINTER(I32)], C.c_int),
    "nbc_autotune": ([C.POINTER(Params), C.POINTER(Options), C.POINTER(SpheresIn), C.POINTER(ClumpsIn),
                      C.POINTER(Wall), C.POINTER(Penetrator), I32, I32, I32, C.POINTER(Options),
                      C.POINTER(TuneEntry), I32, C.POINTER(I32)], C.c_int),
}

_lib = None


def lib():
    """Load libnbc.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libnbc.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class NbcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nbc error {code}: {msg}")
        self.code = code


def check(code, ctx=None):
    if code != 0:
        m = lib().nbc_last_error(ctx)
        raise NbcError(code, m.decode() if m else "")
    return code


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count():
    n = I32(0)
    rc = lib().nbc_device_count(C.byref(n))
    return n.value if rc == 0 else 0


# ---------------------------------------------------------------------------
class Bed:
    """Host state of one bed in file order (numpy copies)."""

    def __init__(self, params, clump_id, pos, local, mass, radius, cpos, cvel, cquat, cangv, cprop, wall):
        self.params = params
        self.clump_id = np.ascontiguousarray(clump_id, dtype=np.int32)
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.local = np.ascontiguousarray(local, dtype=np.float64).reshape(-1, 3)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.cpos = np.ascontiguousarray(cpos, dtype=np.float64).reshape(-1, 3)
        self.cvel = np.ascontiguousarray(cvel, dtype=np.float64).reshape(-1, 3)
        self.cquat = np.ascontiguousarray(cquat, dtype=np.float64).reshape(-1, 4)
        self.cangv = np.ascontiguousarray(cangv, dtype=np.float64).reshape(-1, 3)
        self.cprop = np.ascontiguousarray(cprop, dtype=np.float64).reshape(-1, 4)
        self.wall = wall

    @property
    def n(self):
        return len(self.radius)

    @property
    def nc(self):
        return len(self.cprop)

    def spheres_in(self):
        return SpheresIn(self.n, 0, self.clump_id.ctypes.data_as(C.POINTER(C.c_int32)),
                         self.pos.ctypes.data_as(C.POINTER(C.c_double)),
                         self.local.ctypes.data_as(C.POINTER(C.c_double)),
                         self.mass.ctypes.data_as(C.POINTER(C.c_double)),
                         self.radius.ctypes.data_as(C.POINTER(C.c_double)))

    def clumps_in(self):
        return ClumpsIn(self.nc, 0, self.cpos.ctypes.data_as(C.POINTER(C.c_double)),
                        self.cvel.ctypes.data_as(C.POINTER(C.c_double)),
                        self.cquat.ctypes.data_as(C.POINTER(C.c_double)),
                        self.cangv.ctypes.data_as(C.POINTER(C.c_double)),
                        self.cprop.ctypes.data_as(C.POINTER(C.c_double)))

    def finalized_params(self, opt=None):
        p = Params()
        C.memmove(C.byref(p), C.byref(self.params), C.sizeof(Params))
        check(lib().nbc_finalize_params(C.byref(p), ptr(self.radius), self.n, C.byref(opt) if opt else None))
        return p

    _ARRAYS = ("clump_id", "pos", "local", "mass", "radius", "cpos", "cvel", "cquat", "cangv", "cprop")

    def save(self, path):
        """Write the bed (arrays, parameter block, wall) to an .npz file."""
        np.savez(path, params=np.frombuffer(bytes(self.params), np.uint8),
                 wall=np.frombuffer(bytes(self.wall), np.uint8), **{k: getattr(self, k) for k in self._ARRAYS})

    @classmethod
    def load(cls, path):
        with np.load(path) as z:
            if z["params"].size != C.sizeof(Params) or z["wall"].size != C.sizeof(Wall):
                raise ValueError(f"{path}: parameter block layout does not match this library")
            p = Params.from_buffer_copy(z["params"].tobytes())
            w = Wall.from_buffer_copy(z["wall"].tobytes())
            return cls(p, *[z[k] for k in cls._ARRAYS], w)


class Particles:
    """Host state of a DEMmain single-sphere bed in file order (numpy copies):
    the columns of the 21-column ph file (type, pos, vel, quat, angv, prop =
    Ix Iy Iz mass, radius)."""

    _ARRAYS = ("type", "pos", "vel", "quat", "angv", "prop", "radius")

    def __init__(self, params, type_, pos, vel, quat, angv, prop, radius, wall):
        self.params = params
        self.type = np.ascontiguousarray(type_, dtype=np.int32)
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        self.quat = np.ascontiguousarray(quat, dtype=np.float64).reshape(-1, 4)
        self.angv = np.ascontiguousarray(angv, dtype=np.float64).reshape(-1, 3)
        self.prop = np.ascontiguousarray(prop, dtype=np.float64).reshape(-1, 4)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.wall = wall
        self.params.ClumpNum = self.params.ParticleNum = len(self.radius)

    @property
    def n(self):
        return len(self.radius)

    @classmethod
    def from_bed(cls, bed, wall=None):
        """A bed of one-sphere clumps (generator kind 3) as DEMmain particles."""
        if bed.n != bed.nc or np.any(bed.local != 0.0) or np.any(bed.clump_id != np.arange(bed.n)):
            raise ValueError("from_bed: every clump must be one sphere at its centre")
        p = Params()
        C.memmove(C.byref(p), C.byref(bed.params), C.sizeof(Params))
        return cls(p, np.zeros(bed.n, np.int32), bed.cpos.copy(), bed.cvel.copy(), bed.cquat.copy(),
                   bed.cangv.copy(), bed.cprop.copy(), bed.radius.copy(), wall if wall is not None else bed.wall)

    def particles_in(self):
        def dp(a):
            return a.ctypes.data_as(C.POINTER(C.c_double))
        return ParticlesIn(self.n, 0, self.type.ctypes.data_as(C.POINTER(C.c_int32)), dp(self.pos), dp(self.vel),
                           dp(self.quat), dp(self.angv), dp(self.prop), dp(self.radius))

    def finalized_params(self, opt=None):
        p = Params()
        C.memmove(C.byref(p), C.byref(self.params), C.sizeof(Params))
        check(lib().nbc_finalize_params(C.byref(p), ptr(self.radius), self.n, C.byref(opt) if opt else None))
        return p

    def save(self, path):
        np.savez(path, params=np.frombuffer(bytes(self.params), np.uint8),
                 wall=np.frombuffer(bytes(self.wall), np.uint8), **{k: getattr(self, k) for k in self._ARRAYS})

    @classmethod
    def load(cls, path):
        with np.load(path) as z:
            if z["params"].size != C.sizeof(Params) or z["wall"].size != C.sizeof(Wall):
                raise ValueError(f"{path}: parameter block layout does not match this library")
            p = Params.from_buffer_copy(z["params"].tobytes())
            w = Wall.from_buffer_copy(z["wall"].tobytes())
            return cls(p, *[z[k] for k in cls._ARRAYS], w)


def generate_bed(kind, nclumps=0, seed=1, container=(0, 0, 0, 0)):
    spec = BedSpec(kind, nclumps, seed, D4(*container))
    out = BedOut()
    check(lib().nbc_generate_bed(C.byref(spec), C.byref(out)))
    try:
        n, nc = out.n, out.nc

        def arr(p, m):
            return np.ctypeslib.as_array(p, shape=(m,)).copy()

        params = Params()
        C.memmove(C.byref(params), C.byref(out.params), C.sizeof(Params))
        wall = Wall()
        C.memmove(C.byref(wall), C.byref(out.wall), C.sizeof(Wall))
        bed = Bed(params, arr(out.clump_id, n), arr(out.pos, 3 * n), arr(out.local, 3 * n), arr(out.mass, n),
                  arr(out.radius, n), arr(out.cpos, 3 * nc), arr(out.cvel, 3 * nc), arr(out.cquat, 4 * nc),
                  arr(out.cangv, 3 * nc), arr(out.cprop, 4 * nc), wall)
    finally:
        lib().nbc_free_bed(C.byref(out))
    return bed


def make_options(precision=NBC_FP64, wall_type=0, block_size=0, radix_bits=0, pair_capacity=0, pene_frcoe=0.0,
                 cell_size=0.0, batch_steps=0, hgrid_levels=0, near_skin=0.0, graphs=0, driver=0, level_from_type=0,
                 hgrid_nested=0, single_pass=0, near_horizon=0.0):
    o = Options()
    o.hgrid_nested = hgrid_nested
    o.single_pass = single_pass
    o.driver = driver
    o.level_from_type = level_from_type
    o.precision, o.wall_type, o.block_size, o.radix_bits = precision, wall_type, block_size, radix_bits
    o.pair_capacity, o.pene_frcoe, o.cell_size, o.batch_steps = pair_capacity, pene_frcoe, cell_size, batch_steps
    o.hgrid_levels = hgrid_levels
    o.near_skin = near_skin
    o.near_horizon = near_horizon
    o.graphs = graphs
    return o


def sphere_penetrator(radius, pos, vel, mass=1.0, motion=NBC_PENE_CONSTANT_VELOCITY):
    p = Penetrator()
    p.shape, p.motion, p.radius, p.height = NBC_PENE_SPHERE, motion, radius, radius
    p.dir[:] = (0.0, 0.0, 1.0)
    p.prop[:] = (1.0, 1.0, 1.0, mass)
    p.pos[:] = pos
    p.vel[:] = vel
    p.quat[:] = (0.0, 0.0, 0.0, 1.0)
    return p


def cylinder_penetrator(radius, half_height, pos, vel, axis=(0.0, 0.0, 1.0), mass=1.0,
                        motion=NBC_PENE_CONSTANT_VELOCITY):
    """The reference's cylinder penetrator (PenetratorShape 0): radius, half
    height along the unit axis dir (src/dem_contact_penetrator.cu:42-83)."""
    p = Penetrator()
    p.shape, p.motion, p.radius, p.height = NBC_PENE_CYLINDER, motion, radius, half_height
    p.dir[:] = axis
    p.prop[:] = (1.0, 1.0, 1.0, mass)
    p.pos[:] = pos
    p.vel[:] = vel
    p.quat[:] = (0.0, 0.0, 0.0, 1.0)
    return p


def cone_penetrator(radius, height, pos, vel, nseg=64, mass=1.0):
    p = Penetrator()
    p.shape, p.motion, p.radius, p.height = NBC_PENE_CONE, NBC_PENE_CONSTANT_VELOCITY, radius, height
    p.dir[:] = (0.0, 0.0, 1.0)
    p.prop[:] = (1.0, 1.0, 1.0, mass)
    p.pos[:] = pos
    p.vel[:] = vel
    p.quat[:] = (0.0, 0.0, 0.0, 1.0)
    nt = lib().nbc_make_cone_mesh(radius, height, nseg, None)
    tri = np.zeros(9 * nt)
    lib().nbc_make_cone_mesh(radius, height, nseg, ptr(tri))
    p._tri = tri  # keep alive
    p.ntri = nt
    p.tri = tri.ctypes.data_as(C.POINTER(C.c_double))
    return p


class Context:
    """Device context: one bed resident on one GPU."""

    def __init__(self, params, opt=None, device=0):
        self.params = params
        self.h = C.c_void_p()
        check(lib().nbc_create(C.byref(self.h), C.byref(params), C.byref(opt) if opt else None, device))
        self.N = params.ParticleNum
        self.Nc = params.ClumpNum

    def close(self):
        if self.h:
            lib().nbc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return check(rc, self.h)

    def upload(self, bed, wall=None, pene=None):
        s, c = bed.spheres_in(), bed.clumps_in()
        self._bed = bed
        self._pene = pene
        self._c(lib().nbc_upload(self.h, C.byref(s), C.byref(c), C.byref(wall) if wall else None,
                                 C.byref(pene) if pene else None))

    def upload_particles(self, part, wall, pene=None):
        """DEMmain drivers: one sphere per particle (nbc_upload_particles)."""
        q = part.particles_in()
        self._part, self._pene = part, pene
        self._c(lib().nbc_upload_particles(self.h, C.byref(q), C.byref(wall), C.byref(pene) if pene else None))

    def particles(self):
        N = self.N
        d = dict(sid=np.zeros(N, np.int32), fid=np.zeros(N, np.int32), type=np.zeros(N, np.int32),
                 pos=np.zeros((N, 3)), vel=np.zeros((N, 3)), quat=np.zeros((N, 4)), angv=np.zeros((N, 3)),
                 prop=np.zeros((N, 4)), radius=np.zeros(N))
        self._c(lib().nbc_get_particles(self.h, *[ptr(d[k]) for k in ("sid", "fid", "type", "pos", "vel", "quat",
                                                                      "angv", "prop", "radius")]))
        return d

    def step(self, n=1):
        self._c(lib().nbc_step(self.h, n))

    def rebuild(self):
        self._c(lib().nbc_rebuild(self.h))

    def set_record(self, on=True):
        self._c(lib().nbc_set_record(self.h, int(on)))

    def sync(self):
        self._c(lib().nbc_synchronize(self.h))

    def stats(self):
        s = Stats()
        self._c(lib().nbc_get_stats(self.h, C.byref(s)))
        return s

    def spheres(self):
        N = self.N
        d = dict(fid=np.zeros(N, np.int32), cid=np.zeros(N, np.int32), pos=np.zeros((N, 3)),
                 cum=np.zeros((N, 3)), local=np.zeros((N, 3)), mass=np.zeros(N), radius=np.zeros(N))
        self._c(lib().nbc_get_spheres(self.h, ptr(d["fid"]), ptr(d["cid"]), ptr(d["pos"]), ptr(d["cum"]),
                                      ptr(d["local"]), ptr(d["mass"]), ptr(d["radius"])))
        return d

    def grid(self):
        N, CN = self.N, self.params.CellNum
        d = dict(cell=np.zeros(N, np.int32), perm=np.zeros(N, np.int32), beg=np.zeros(CN, np.int32),
                 end=np.zeros(CN, np.int32))
        self._c(lib().nbc_get_grid(self.h, ptr(d["cell"]), ptr(d["perm"]), ptr(d["beg"]), ptr(d["end"])))
        return d

    def pair_counts(self):
        a, b = C.c_int64(), C.c_int64()
        self._c(lib().nbc_get_pair_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def pairs(self):
        nf, nb = self.pair_counts()
        N = self.N
        d = dict(foff=np.zeros(N + 1, np.int32), fj=np.zeros(max(nf, 1), np.int32), fflag=np.zeros(max(nf, 1), np.uint8),
                 boff=np.zeros(N + 1, np.int32), bi=np.zeros(max(nb, 1), np.int32), bslot=np.zeros(max(nb, 1), np.int32))
        self._c(lib().nbc_get_pairs(self.h, ptr(d["foff"]), ptr(d["fj"]), ptr(d["fflag"]), ptr(d["boff"]),
                                    ptr(d["bi"]), ptr(d["bslot"])))
        d["fj"], d["fflag"], d["bi"], d["bslot"] = d["fj"][:nf], d["fflag"][:nf], d["bi"][:nb], d["bslot"][:nb]
        return d

    def pair_history(self):
        nf, _ = self.pair_counts()
        h = np.zeros((max(nf, 1), 9))
        self._c(lib().nbc_get_pair_history(self.h, ptr(h)))
        return h[:nf]

    def sphere_forces(self, kind=0):
        f, t = np.zeros((self.N, 3)), np.zeros((self.N, 3))
        self._c(lib().nbc_get_sphere_forces(self.h, kind, ptr(f), ptr(t)))
        return f, t

    def clumps(self):
        Nc = self.Nc
        d = dict(pos=np.zeros((Nc, 3)), vel=np.zeros((Nc, 3)), quat=np.zeros((Nc, 4)), angv=np.zeros((Nc, 3)),
                 F=np.zeros((Nc, 3)), T=np.zeros((Nc, 3)))
        self._c(lib().nbc_get_clumps(self.h, ptr(d["pos"]), ptr(d["vel"]), ptr(d["quat"]), ptr(d["angv"]),
                                     ptr(d["F"]), ptr(d["T"])))
        return d

    def penetrator(self):
        oi = max(self.params.OutputInterval, 1)
        d = dict(pos=np.zeros(3), vel=np.zeros(3), quat=np.zeros(4), angv=np.zeros(3), cache_f=np.zeros((oi, 3)),
                 cache_t=np.zeros((oi, 3)))
        st = I32()
        self._c(lib().nbc_get_penetrator(self.h, ptr(d["pos"]), ptr(d["vel"]), ptr(d["quat"]), ptr(d["angv"]),
                                         ptr(d["cache_f"]), ptr(d["cache_t"]), C.byref(st)))
        d["state"] = st.value
        return d

    def pene_exports(self):
        a = [np.zeros((self.N, 3)) for _ in range(4)]
        self._c(lib().nbc_get_pene_exports(self.h, *[ptr(x) for x in a]))
        return a

    def pair_exports(self):
        nf, _ = self.pair_counts()
        a, b = np.zeros((max(nf, 1), 3)), np.zeros((max(nf, 1), 3))
        self._c(lib().nbc_get_pair_exports(self.h, ptr(a), ptr(b)))
        return a[:nf], b[:nf]

    def profile(self, on=True):
        self._c(lib().nbc_profile_enable(self.h, int(on)))

    def profile_get(self, cls):
        ms, n = C.c_double(), C.c_int64()
        self._c(lib().nbc_profile_get(self.h, cls, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class BatchContext(Context):
    """Device context holding several independent beds, advanced together.

    Getters return all beds concatenated; `ranges()` gives each bed's sphere
    and clump index ranges (file indices and clump ids are offset by bed)."""

    def __init__(self, params_list, opt=None, device=0):
        n = len(params_list)
        arr = (Params * n)()
        for b, q in enumerate(params_list):
            C.memmove(C.byref(arr[b]), C.byref(q), C.sizeof(Params))
        self.h = C.c_void_p()
        check(lib().nbc_create_batch(C.byref(self.h), n, arr, C.byref(opt) if opt else None, device))
        self.nbeds = n
        self.bed_params = arr
        p = Params()
        C.memmove(C.byref(p), C.byref(arr[0]), C.sizeof(Params))
        p.ParticleNum = sum(q.ParticleNum for q in arr)
        p.ClumpNum = sum(q.ClumpNum for q in arr)
        p.CellNum = sum(q.CellNum for q in arr)
        self.params = p
        self.N = p.ParticleNum
        self.Nc = p.ClumpNum

    def upload(self, beds, walls=True):
        assert len(beds) == self.nbeds
        n = self.nbeds
        s_arr, c_arr, w_arr = (SpheresIn * n)(), (ClumpsIn * n)(), (Wall * n)()
        for b, bed in enumerate(beds):
            s_arr[b], c_arr[b] = bed.spheres_in(), bed.clumps_in()
            if walls and bed.wall is not None:
                C.memmove(C.byref(w_arr[b]), C.byref(bed.wall), C.sizeof(Wall))
        self._beds = list(beds)  # the structs point into the beds' arrays
        self._c(lib().nbc_upload_batch(self.h, s_arr, c_arr, w_arr if walls else None))

    def ranges(self):
        so, co = np.zeros(self.nbeds + 1, np.int32), np.zeros(self.nbeds + 1, np.int32)
        self._c(lib().nbc_get_batch_ranges(self.h, ptr(so), ptr(co)))
        return so, co


def autotune(bed, base=None, wall=None, pene=None, steps=300, flags=0, device=0):
    """Fastest options for this bed (nbc_autotune): returns (options, [(options, ms_per_step), ...])."""
    s, c = bed.spheres_in(), bed.clumps_in()
    best = Options()
    cap = 64
    ent = (TuneEntry * cap)()
    n = I32()
    check(lib().nbc_autotune(C.byref(bed.params), C.byref(base) if base else None, C.byref(s), C.byref(c),
                             C.byref(wall) if wall else None, C.byref(pene) if pene else None, device, steps, flags,
                             C.byref(best), ent, cap, C.byref(n)))
    return best, [(ent[k].opt, ent[k].ms_per_step) for k in range(min(n.value, cap))]


def autotune_batch(beds, base=None, steps=100, flags=0, device=0):
    """Fastest options for a batch of beds (nbc_autotune_batch): returns
    (options, [(options, ms_per_step), ...]). The beds' parameters must not be
    finalized."""
    n = len(beds)
    p_arr, s_arr, c_arr, w_arr = (Params * n)(), (SpheresIn * n)(), (ClumpsIn * n)(), (Wall * n)()
    for b, bed in enumerate(beds):
        C.memmove(C.byref(p_arr[b]), C.byref(bed.params), C.sizeof(Params))
        s_arr[b], c_arr[b] = bed.spheres_in(), bed.clumps_in()
        if bed.wall is not None:
            C.memmove(C.byref(w_arr[b]), C.byref(bed.wall), C.sizeof(Wall))
    best = Options()
    cap = 64
    ent = (TuneEntry * cap)()
    cnt = I32()
    check(lib().nbc_autotune_batch(n, p_arr, C.byref(base) if base else None, s_arr, c_arr, w_arr, device, steps,
                                   flags, C.byref(best), ent, cap, C.byref(cnt)))
    return best, [(ent[k].opt, ent[k].ms_per_step) for k in range(min(cnt.value, cap))]


def download_bed(ctx, template):
    """Current device state as a Bed in file order (sphere file ids restored).

    `template` supplies the parameters and wall of the returned bed."""
    sp = ctx.spheres()
    cl = ctx.clumps()
    order = np.argsort(sp["fid"], kind="stable")
    p = Params()
    C.memmove(C.byref(p), C.byref(template.params), C.sizeof(Params))
    prop = template.cprop.copy()
    return Bed(p, sp["cid"][order], sp["pos"][order], sp["local"][order], sp["mass"][order], sp["radius"][order],
               cl["pos"], cl["vel"], cl["quat"], cl["angv"], prop, template.wall)

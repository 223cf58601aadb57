"""ctypes binding of libmpm_hip.so (the C ABI declared in include/mpm_c_abi.h).

The HIP library is the only compute path: importing this module never falls back to anything else.
If the library has not been built, `load()` raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_hip.so")

MPM_OK = 0
KERNEL_QUADRATIC, KERNEL_CUBIC = 2, 3
SNOW, SAND = 0, 1
FRICTION_REFERENCE, FRICTION_COULOMB, FRICTION_STICKY = 0, 1, 2
LS_PLANE, LS_WALL2GROUND, LS_BOX, LS_SPHERE = 0, 1, 2, 3
STAGE_SORT, STAGE_P2G, STAGE_GRID, STAGE_G2P = 0, 1, 2, 3
NUM_STAGES = 4
STATUS_NONFINITE, STATUS_OUT_OF_DOMAIN, STATUS_CFL = 1, 2, 4
# return codes of every mpm_* call (include/mpm_c_abi.h MPM_OK, MPM_ERR_*)
OK, ERR_INVALID, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_UNSUPPORTED = 0, 1, 2, 3, 4, 5


class Params(ctypes.Structure):
    _fields_ = [
        ("kernel_order", ctypes.c_int32),
        ("material", ctypes.c_int32),
        ("stress_at_predicted_F", ctypes.c_int32),
        ("friction_model", ctypes.c_int32),
        ("gravity", ctypes.c_float * 3),
        ("youngs_modulus", ctypes.c_float),
        ("poisson_ratio", ctypes.c_float),
        ("critical_compression", ctypes.c_float),
        ("critical_stretch", ctypes.c_float),
        ("hardening", ctypes.c_float),
        ("dp_h", ctypes.c_float * 4),
        ("per_particle_phi", ctypes.c_int32),
        ("friction", ctypes.c_float),
        ("apic_damping", ctypes.c_float),
        ("fp64_svd_ratio", ctypes.c_float),
        ("fp64_apex_radius", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 4),
    ]


class LevelSetPrim(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("a", ctypes.c_float * 4), ("b", ctypes.c_float * 4)]


MAX_LEVELSETS = 16


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("min_bound", ctypes.c_double * 3),
        ("max_bound", ctypes.c_double * 3),
        ("res", ctypes.c_int32 * 3),
        ("n_particles", ctypes.c_int32),
        ("params", Params),
        ("n_levelsets", ctypes.c_int32),
        ("levelsets", LevelSetPrim * MAX_LEVELSETS),
    ]


SORT_COUNTING, SORT_RADIX = 0, 1
SHAPE_BALL, SHAPE_BLOCK, SHAPE_CYLINDER = 0, 1, 2
P2G_ROWS, P2G_LAYERS = 0, 1


class Tuning(ctypes.Structure):
    _fields_ = [("p2g_ranks", ctypes.c_int32), ("p2g_blocks", ctypes.c_int32), ("grid_blocks", ctypes.c_int32),
                ("p2g_ms", ctypes.c_float), ("grid_ms", ctypes.c_float), ("sort_mode", ctypes.c_int32),
                ("sort_ms", ctypes.c_float), ("p2g_pass", ctypes.c_int32)]


class MeshParams(ctypes.Structure):
    _fields_ = [("youngs_modulus", ctypes.c_float), ("poisson_ratio", ctypes.c_float),
                ("shear_stiffness", ctypes.c_float), ("stiffness", ctypes.c_float),
                ("friction_angle", ctypes.c_float), ("density", ctypes.c_float), ("thickness", ctypes.c_float),
                ("reference_normal_row", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("substeps", ctypes.c_int64),
        ("status", ctypes.c_int32),
        ("n_particles", ctypes.c_int32),
        ("n_active_tiles", ctypes.c_int32),
        ("n_active_blocks", ctypes.c_int32),
        ("n_active_nodes", ctypes.c_int64),
        ("fp64_svd_count", ctypes.c_int64),
        ("fp64_apex_count", ctypes.c_int64),
        ("surrogate_candidates", ctypes.c_int64),
        ("surrogate_accepted", ctypes.c_int64),
        ("pairs_logged", ctypes.c_int64),
    ]


_P = ctypes.POINTER
_f = _P(ctypes.c_float)
_d = _P(ctypes.c_double)
_i32 = _P(ctypes.c_int32)
_u32 = _P(ctypes.c_uint32)
_ctx = ctypes.c_void_p

# (name, restype, argtypes) for every exported entry point of include/mpm_c_abi.h
SIGNATURES = [
    ("mpm_params_default", None, [_P(Params)]),
    ("mpm_create", ctypes.c_int, [_P(ctypes.c_void_p), ctypes.c_int32]),
    ("mpm_destroy", ctypes.c_int, [_ctx]),
    ("mpm_last_error", ctypes.c_char_p, [_ctx]),
    ("mpm_set_grid", ctypes.c_int, [_ctx, _d, _d, _i32]),
    ("mpm_set_params", ctypes.c_int, [_ctx, _P(Params)]),
    ("mpm_set_levelset", ctypes.c_int, [_ctx, _P(LevelSetPrim), ctypes.c_int32]),
    ("mpm_set_scenes", ctypes.c_int, [_ctx, ctypes.c_int32, _P(SceneDesc)]),
    ("mpm_set_particles", ctypes.c_int, [_ctx, ctypes.c_int32, _f, _f, _f, _f, _f, _f, _f, _f]),
    ("mpm_get_particles", ctypes.c_int, [_ctx, _f, _f, _f, _f, _f, _f]),
    ("mpm_set_mesh", ctypes.c_int, [_ctx, ctypes.c_int32, _f, _f, ctypes.c_int32, _i32, _P(ctypes.c_uint8),
                                    _P(MeshParams)]),
    ("mpm_get_mesh", ctypes.c_int, [_ctx, _f, _f, _f]),
    ("mpm_get_stress", ctypes.c_int, [_ctx, _f]),
    ("mpm_set_stress", ctypes.c_int, [_ctx, _f]),
    ("mpm_estimate_volumes", ctypes.c_int, [_ctx]),
    ("mpm_substep", ctypes.c_int, [_ctx, ctypes.c_float]),
    ("mpm_run", ctypes.c_int, [_ctx, ctypes.c_float, ctypes.c_int32]),
    ("mpm_solve", ctypes.c_int, [_ctx, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    ("mpm_autotune", ctypes.c_int, [_ctx, ctypes.c_float, ctypes.c_int32, _P(Tuning)]),
    ("mpm_get_tuning", ctypes.c_int, [_ctx, _P(Tuning)]),
    ("mpm_set_tuning", ctypes.c_int, [_ctx, _P(Tuning)]),
    ("mpm_get_clock", ctypes.c_int,
     [_ctx, _d, _d, _i32, _P(ctypes.c_int64)]),
    ("mpm_stage", ctypes.c_int, [_ctx, ctypes.c_int32, ctypes.c_float]),
    ("mpm_get_sorted", ctypes.c_int, [_ctx, _u32, _i32, _i32]),
    ("mpm_get_grid", ctypes.c_int, [_ctx, _f, _f, _f, _f]),
    ("mpm_get_stats", ctypes.c_int, [_ctx, _P(Stats)]),
    ("mpm_clear_status", ctypes.c_int, [_ctx]),
    ("mpm_synchronize", ctypes.c_int, [_ctx]),
    ("mpm_time_substeps", ctypes.c_int, [_ctx, ctypes.c_float, ctypes.c_int32, _f, _f]),
    ("mpm_time_kernels", ctypes.c_int, [_ctx, ctypes.c_float, ctypes.c_int32, _f, _f, _f]),
    ("mpm_time_surrogate", ctypes.c_int, [_ctx, ctypes.c_float, ctypes.c_int32, _P(ctypes.c_float), _P(ctypes.c_int64)]),
    ("mpm_surrogate_nparam", ctypes.c_int32, []),
    ("mpm_surrogate_init", ctypes.c_int, [_ctx, ctypes.c_uint64]),
    ("mpm_surrogate_set_weights", ctypes.c_int, [_ctx, _f, ctypes.c_int32]),
    ("mpm_surrogate_get_weights", ctypes.c_int, [_ctx, _f, ctypes.c_int32]),
    ("mpm_surrogate_predict", ctypes.c_int, [_ctx, _f, _f, ctypes.c_int32]),
    ("mpm_surrogate_gradient", ctypes.c_int, [_ctx, _f, _f, ctypes.c_int32, _f, _f]),
    ("mpm_surrogate_gradient_noalloc", ctypes.c_int, [_ctx, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                      ctypes.c_void_p, ctypes.c_void_p]),
    ("mpm_surrogate_train", ctypes.c_int,
     [_ctx, _f, _f, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64, _f]),
    ("mpm_set_pair_logging", ctypes.c_int, [_ctx, ctypes.c_int64]),
    ("mpm_get_pairs", ctypes.c_int, [_ctx, _f, ctypes.c_int64, _P(ctypes.c_int64)]),
    ("mpm_set_surrogate_log", ctypes.c_int, [_ctx, ctypes.c_int64]),
    ("mpm_get_surrogate_log", ctypes.c_int, [_ctx, _f, ctypes.c_int64, _P(ctypes.c_int64)]),
    ("mpm_set_surrogate", ctypes.c_int, [_ctx, ctypes.c_int32, ctypes.c_float]),
    ("mpm_bench_copy", ctypes.c_int, [_ctx, ctypes.c_int64, ctypes.c_int32, _f]),
    ("mpm_generate_block", ctypes.c_int,
     [ctypes.c_uint64, _d, _d, _i32, ctypes.c_int32, ctypes.c_int64, _f]),
    ("mpm_generate_shape", ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, _d, ctypes.c_int64, _f]),
]

# void (*)(mpm_ctx*, int32_t frame, double t, void* user)
FRAME_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p)

_lib = None


def load(path=None):
    """Load libmpm_hip.so and attach the C signatures. Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: the HIP extension has not been built "
            "(run `python -c 'import __graft_entry__ as g; g.build()'` from the repository root)")
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def default_params():
    lib = load()
    p = Params()
    lib.mpm_params_default(ctypes.byref(p))
    return p

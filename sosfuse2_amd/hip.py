"""ctypes binding of libsosfuse2_hip.so (include/sosfuse2_hip.h). Loading fails loudly: there is no CPU fallback."""

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / 'libsosfuse2_hip.so'

c_i32, c_i64, c_u32, c_f32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_float
P = ctypes.c_void_p

LANE_CONSTANTS = 12
PROFILE_SHAPES, PROFILE_EXP, PROFILE_GAUSS, PROFILE_TRIANGLE, PROFILE_TABLE, PROFILE_AIR_SQUARED, PROFILE_ZERO = range(7)
FIRST, BOUNDARY, RECORDS, UNIFORM_PHASE, GROUND_RECORDS, NODE_GEOMETRY, TENSOR4 = 1, 2, 8, 16, 32, 64, 128


class Bundle(ctypes.Structure):
    _fields_ = [
        ('groups', c_i64), ('rays', c_i64), ('group_off_dev', P), ('starts_dev', P), ('targets_dev', P),
        ('ray_group_dev', P), ('dirs_dev', P), ('dloc_dev', P), ('weights_dev', P), ('edge_off_dev', P),
        ('edges_dev', P), ('lower_dev', P), ('end_j_dev', P), ('end_f_dev', P), ('end_col_dev', P), ('hits_dev', P),
        ('node_off_dev', P), ('node_geom_dev', P), ('node_sun_dev', P),
    ]


class Grid(ctypes.Structure):
    _fields_ = [
        ('nh', c_i32), ('nc', c_i32), ('L', c_i32), ('lanes', c_i32), ('nq', c_i32), ('strato', c_i32),
        ('top_km', c_f32), ('air_H', c_f32), ('aer_H', c_f32), ('o3_c', c_f32), ('o3_s', c_f32), ('no2_c', c_f32),
        ('no2_s', c_f32), ('st_c', c_f32), ('st_s', c_f32),
        ('gl_u_dev', P), ('gl_w_dev', P), ('h_dev', P), ('chi_dev', P), ('hr_dev', P), ('lane_dev', P),
        ('sun_nh', c_i32), ('sun_nx', c_i32), ('sun_h_dev', P), ('sun_col_dev', P), ('goff_dev', P),
        ('gdir_dev', P), ('gval_dev', P), ('gm_dev', P), ('blocks_per_cu', c_i32), ('sh_tile', c_i32),
        ('pool_max', c_i32),
        ('profile_kind', c_i32 * 6), ('profile_a', c_f32 * 6), ('profile_b', c_f32 * 6), ('profile_table_dev', P),
        ('profile_table_n', c_i32),
    ]


class Columns(ctypes.Structure):
    _fields_ = [
        ('top_km', ctypes.c_double), ('nq', c_i32), ('gl_u_dev', P), ('gl_w_dev', P), ('profile_kind', c_i32 * 6),
        ('profile_a', ctypes.c_double * 6), ('profile_b', ctypes.c_double * 6), ('profile_table_dev', P),
        ('profile_table_n', c_i32),
    ]


class Incident(ctypes.Structure):
    _fields_ = [
        ('nh', c_i32), ('nc', c_i32), ('nmu', c_i32), ('nphi', c_i32), ('nlevels', c_i32), ('h_dev', P),
        ('chi_dev', P), ('levels', ctypes.c_double * 8), ('gl_u_dev', P), ('gl_w_dev', P),
    ]


_SIGNATURES = {
    'sos_incident_count': [ctypes.POINTER(Incident), P, P],
    'sos_incident_fill': [ctypes.POINTER(Incident), P, P, P, P, P],
    'sos_pack_count': [ctypes.POINTER(Columns), c_i32, P, P, c_i64, P, P, P, P, P],
    'sos_pack_fill': [ctypes.POINTER(Columns), c_i32, P, P, P, c_i64, P, P, P, P, P, P, P, P, P, P],
    'sos_sun_table': [ctypes.POINTER(Columns), c_i32, c_i32, P, P, P, P],
    'sos_ray_columns': [ctypes.POINTER(Columns), c_i64, P, P, P, P, P, P],
    'sos_transport': [ctypes.POINTER(Grid), ctypes.POINTER(Bundle), c_i32, c_i32, c_i32, P, P, P, P, P, P, P],
    'sos_plan_count': [ctypes.POINTER(Grid), ctypes.POINTER(Bundle), P, P],
    'sos_plan_nodes': [ctypes.POINTER(Grid), ctypes.POINTER(Bundle), P, P, P, P, P, P],
    'sos_plan_geometry': [ctypes.POINTER(Grid), ctypes.POINTER(Bundle), P, P, P],
    'sos_plan_sun': [ctypes.POINTER(Grid), ctypes.POINTER(Bundle), P, P, P],
    'sos_ground_records': [ctypes.POINTER(Grid), c_i32, ctypes.POINTER(c_f32), ctypes.POINTER(c_f32), c_i64, P, P, P, P,
                           P, P, P],
    'sos_plan_basis': [ctypes.POINTER(Grid), ctypes.POINTER(Bundle), P, P],
    'sos_ground_pools': [ctypes.POINTER(Grid), P, P, P, P, P, P, P],
    'sos_sh_records': [ctypes.POINTER(Grid), P, P, c_i32, c_i64, P, P, P, P, P, P, P, P, P],
    'sos_project': [ctypes.POINTER(Grid), ctypes.POINTER(Bundle), P, c_i32, P, P, P, P],
    'sos_compose': [ctypes.POINTER(Grid), c_i32, P, P, P, P, P],
    'sos_make_plane': [ctypes.POINTER(Grid), c_i32, P, P, P, c_i32, P, P, P],
    'sos_source_metrics': [c_i64, c_i32, P, P, c_u32, P, P],
    'sos_anderson_residual': [c_i64, c_i32, c_i32, P, P, P, c_i32, P, P, P],
    'sos_anderson_gram': [c_i64, c_i32, c_i32, c_i32, c_u32, P, P, P, P],
    'sos_anderson_gram_static': [c_i64, c_i32, c_i32, c_i32, c_u32, P, P, P],
    'sos_anderson_solve': [c_i32, c_i32, c_i32, P, P, c_u32, P, P],
    'sos_anderson_mix': [c_i64, c_i32, c_i32, c_i32, c_i32, P, P, P, P, P],
    'sos_mlp_gemm': [c_i64, c_i64, c_i64, P, c_i64, c_i64, P, c_i64, c_i64, P, c_i32, P, c_i64, P, c_i64, c_i32, P, P],
    'sos_mlp_gelu_backward': [c_i64, c_i64, P, c_i64, P, c_i64, P, c_i64, P],
    'sos_mlp_mse': [c_i64, c_i64, P, c_i64, P, c_i64, P, P, c_i64, P, P, P],
    'sos_mlp_bias_grad': [c_i64, c_i64, P, c_i64, P, P],
    'sos_mlp_adam': [c_i64, P, P, P, P, c_i32, c_f32, c_f32, c_f32, c_f32, P],
    'sos_device_info': [ctypes.POINTER(c_i32), ctypes.POINTER(c_i32), ctypes.c_char_p, c_i32],
}

_lib = None


def exported_symbols():
    """Names the header declares (every one must be exported by the library)."""
    return sorted(_SIGNATURES) + ['sos_last_error'] + sorted(_EXTRA)


_EXTRA = {}


def register(name, argtypes):
    """Declare a further entry point (used by modules that add kernels, e.g. the warm-start network)."""
    _EXTRA[name] = argtypes
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f'E_HIP_MISSING: {LIB_PATH} is not built; run __graft_entry__.build()')
        _lib = ctypes.CDLL(str(LIB_PATH))
        for name, argtypes in {**_SIGNATURES, **_EXTRA}.items():
            fn = getattr(_lib, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        _lib.sos_last_error.restype = ctypes.c_char_p
        _lib.sos_last_error.argtypes = []
    return _lib


_started = False


def mark_started():
    """Record that this process has started the GPU runtime (Kernel creation, any library call)."""
    global _started
    _started = True


def started():
    return _started


def call(name, *args):
    """Call an entry point; a nonzero status raises with the library's message."""
    mark_started()
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise RuntimeError(f'E_HIP_{name.upper()}: status {rc}: {lib().sos_last_error().decode()}')


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())

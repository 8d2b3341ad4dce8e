"""ctypes binding of libhx.so (the C ABI in include/hx.h).

The library must be built in tree (hippie_amd/libhx.so, see _build.py); loading
fails loudly when it is missing -- there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhx.so")

HX_OK = 0
HX_ERR_INVALID_ARG = 1
HX_ERR_HIP = 2
HX_ERR_OUT_OF_MEMORY = 3
HX_ERR_UNSUPPORTED = 4
HX_ERR_STACK_OVERFLOW = 5

HX_METHOD_LBVH = 0
HX_METHOD_HLBVH = 1
HX_METHOD_PLOC = 2
HX_METHOD_ATRBVH = 3
HX_METHOD_INSERTION = 4
HX_METHOD_TR = 5
HX_LAYOUT_BIN = 2
HX_LAYOUT_QUAD = 4
HX_LAYOUT_OCT = 8
HX_RAYKEY_TWOPOINT = 0
HX_RAYKEY_AILA = 1


class HxBuildParams(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("morton_bits", ctypes.c_int32),
        ("max_leaf_size", ctypes.c_int32),
        ("adaptive_leaf_size", ctypes.c_int32),
        ("ct", ctypes.c_float),
        ("ci", ctypes.c_float),
        ("presplitting", ctypes.c_int32),
        ("presplitter_beta", ctypes.c_float),
        ("ploc_radius", ctypes.c_int32),
        ("ploc_max_radius", ctypes.c_int32),
        ("ploc_adaptive", ctypes.c_int32),
        ("atr_iterations", ctypes.c_int32),
        ("atr_treelet_size", ctypes.c_int32),
        ("tr_iterations", ctypes.c_int32),
        ("hlbvh_sah_bits", ctypes.c_int32),
        ("insertion_mod", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("use_scene_box", ctypes.c_int32),
        ("scene_box", ctypes.c_float * 6),
        ("keep_debug", ctypes.c_int32),
    ]


class HxBvhView(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_void_p),
        ("woop", ctypes.c_void_p),
        ("tri_indices", ctypes.c_void_p),
        ("n_interior", ctypes.c_int64),
        ("n_leaves", ctypes.c_int64),
        ("n_references", ctypes.c_int64),
        ("scene_box", ctypes.c_float * 6),
        ("layout", ctypes.c_int32),
    ]


class HxBvhDebugView(ctypes.Structure):
    _fields_ = [
        ("sorted_keys", ctypes.c_void_p),
        ("keys_are_64bit", ctypes.c_int32),
        ("sorted_tri", ctypes.c_void_p),
        ("left", ctypes.c_void_p),
        ("right", ctypes.c_void_p),
        ("parent", ctypes.c_void_p),
        ("node_min", ctypes.c_void_p),
        ("node_max", ctypes.c_void_p),
        ("state", ctypes.c_void_p),
    ]


_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (all entry points return int status unless listed in _RESTYPE)
SIGNATURES = {
    "hx_status_string": [ctypes.c_int],
    "hx_last_error": [],
    "hx_context_create": [ctypes.c_int, _pp],
    "hx_context_destroy": [_p],
    "hx_context_set_stream": [_p, _p],
    "hx_context_synchronize": [_p],
    "hx_context_set_profiling": [_p, ctypes.c_int],
    "hx_context_profile_count": [_p, ctypes.POINTER(_i32)],
    "hx_context_profile_entry": [_p, _i32, ctypes.c_char_p, _i32, ctypes.POINTER(_i64),
                                 ctypes.POINTER(ctypes.c_double)],
    "hx_context_profile_reset": [_p],
    "hx_context_set_tuning": [_p, ctypes.c_char_p, _i32],
    "hx_context_get_tuning": [_p, ctypes.c_char_p, ctypes.POINTER(_i32)],
    "hx_build_params_default": [ctypes.POINTER(HxBuildParams)],
    "hx_build": [_p, ctypes.POINTER(HxBuildParams), _p, _i64, _p, _i64, _pp],
    "hx_rebuild": [_p, _p, _p, _i64, _p, _i64],
    "hx_refit": [_p, _p, _p, _i64, _p, _i64],
    "hx_bvh_destroy": [_p],
    "hx_bvh_get_view": [_p, ctypes.POINTER(HxBvhView)],
    "hx_bvh_debug": [_p, ctypes.POINTER(HxBvhDebugView)],
    "hx_bvh_cost": [_p, ctypes.POINTER(ctypes.c_double)],
    "hx_bvh_info": [_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)],
    "hx_scene_box": [_p, _p, _p, _i64, _p],
    "hx_normalize_vertices": [_p, _p, _i64, _p],
    "hx_bvh_leaf_stats": [_p, _p, ctypes.POINTER(_i64)],
    "hx_bvh_validate": [_p, ctypes.POINTER(_i32)],
    "hx_morton_codes": [_p, _p, _p, _i64, _p, _i32, _p],
    "hx_sort_pairs_u64": [_p, _p, _p, _i64, _i32],
    "hx_trace": [_p, _p, _p, _i64, _i32, _p, _p],
    "hx_trace_sorted": [_p, _p, _p, _i64, _i32, _i32, _i32, _f32, _p, _p, _p],
    "hx_trace_sorted_timed": [_p, _p, _p, _i64, _i32, _i32, _i32, _f32, _p, _p, _p, ctypes.POINTER(_f32),
                              ctypes.POINTER(_f32)],
    "hx_init_seeds": [_p, _i64, _i32, _p],
    "hx_primary_rays": [_p, _i32, _p, _p, _i32, _i32, _f32, _p, _p],
    "hx_shadow_rays": [_p, _i64, _i64, _i32, _p, _f32, _p, _p, _p, _p],
    "hx_ao_rays": [_p, _i64, _i64, _i32, _p, _f32, _p, _p, _p, _p, _p],
    "hx_path_rays": [_p, _i32, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "hx_reconstruct_occlusion": [_p, _i64, _i64, _i32, _i32, _p, _p, _p, _p, _p],
    "hx_interpolate_vertices": [_p, _i64, _i64, _f32, _p, _p, _p, _i32],
    "hx_trace_sorted_noalloc": [_p, _p, _p, _i64, _i32, _i32, _i32, _f32, _p, _p, _p],
    "hx_trace_view": [_p, _p, _i64, _p, _p, _i64, _p, _i64, _i32, _p, _p, _p],
    "hx_ray_keys": [_p, _p, _i64, _i32, _i32, _f32, _p],
    "hx_memcpy": [_p, _p, _p, _i64],
    "hx_backward": [_p, _p, _p, _p, _p, _i64, _p, _p, _p, _p],
    "hx_build_batch": [_p, _p, _p, _i64, _p, _p, _i32, _p],
    "hx_batch_refit": [_p, _p, _p, _i64, _p, _i64],
    "hx_batch_rebuild": [_p, _p, _p, _i64, _p, _i64],
    "hx_batch_destroy": [_p],
    "hx_batch_mesh": [_p, _i32, _p],
    "hx_batch_trace": [_p, _p, _p, _p, _i64, _i32, _p, _p],
    "hx_batch_backward": [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _p, _p, _p],
    "hx_batch_recompute_hits": [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _p, _p],
    "hx_recompute_hits": [_p, _p, _p, _p, _p, _i64, _p, _p, _p],
    "hx_trace_overflow_count": [_p, ctypes.POINTER(_i64)],
}
_RESTYPE = {"hx_status_string": ctypes.c_char_p, "hx_last_error": ctypes.c_char_p}

_lib = None


class HxError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__(f"{where}: {message} (status {status})")
        self.status = status


def lib():
    """Load libhx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python hippie_amd/_build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(status, where):
    if status != HX_OK:
        msg = lib().hx_last_error()
        msg = msg.decode() if msg else lib().hx_status_string(status).decode()
        raise HxError(status, where, msg)


def call(name, *args):
    check(getattr(lib(), name)(*args), name)

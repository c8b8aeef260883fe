"""ctypes declarations of the C-ABI (include/skelcl_rt.h, include/skelcl_skeletons.h).

Loading fails loudly when libskelcl_amd.so is missing: there is no fallback
path.  Build it with ``python -m skelcl_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libskelcl_amd.so")

# enums
SKL_OK = 0
SKL_ERR_INVALID, SKL_ERR_NO_DEVICE, SKL_ERR_HIP, SKL_ERR_COMPILE = 1, 2, 3, 4
SKL_ERR_PARSE, SKL_ERR_UNSUPPORTED, SKL_ERR_NOT_INIT = 5, 6, 7
(SKL_MAP, SKL_MAP_INDEX, SKL_MAP_INDEXPOINT, SKL_ZIP, SKL_REDUCE, SKL_SCAN,
 SKL_MAPOVERLAP, SKL_ALLPAIRS, SKL_ALLPAIRS_GENERAL) = range(9)
SKL_PADDING_NEUTRAL, SKL_PADDING_NEAREST = 0, 1
SKL_PREC_DEFAULT, SKL_PREC_FP32, SKL_PREC_BF16X3, SKL_PREC_FP8, SKL_PREC_FP64 = 0, 1, 2, 3, 4
SKL_ARG_VALUE, SKL_ARG_BUFFER, SKL_ARG_MATRIX, SKL_ARG_LOCAL = 0, 1, 2, 3


class skl_type(C.Structure):
    _fields_ = [("name", C.c_char_p), ("size", C.c_size_t)]


class skl_skeleton_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("source", C.c_char_p), ("func", C.c_char_p),
        ("source2", C.c_char_p), ("func2", C.c_char_p),
        ("identity", C.c_char_p),
        ("types", skl_type * 3), ("ntypes", C.c_int),
        ("overlap_range", C.c_int), ("padding", C.c_int), ("neutral", C.c_char_p),
        ("work_group_size", C.c_int), ("precision", C.c_int),
        ("deriv_source", C.c_char_p), ("deriv_func", C.c_char_p),
        ("deriv_func2", C.c_char_p),
        ("build_options", C.c_char_p),
    ]


class skl_arg(C.Structure):
    _fields_ = [("kind", C.c_int), ("value", C.c_void_p), ("size", C.c_size_t),
                ("cols", C.c_uint)]


class skl_device_info(C.Structure):
    _fields_ = [
        ("name", C.c_char * 256), ("arch", C.c_char * 64),
        ("compute_units", C.c_int), ("wavefront_size", C.c_int),
        ("max_work_group_size", C.c_int), ("lds_per_block", C.c_size_t),
        ("global_mem", C.c_size_t), ("clock_khz", C.c_int),
        ("memory_clock_khz", C.c_int), ("memory_bus_width", C.c_int),
        ("supports_double", C.c_int), ("l2_cache_bytes", C.c_int),
    ]


class skl_gemm_desc(C.Structure):
    _fields_ = [("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p),
                ("M", C.c_int), ("N", C.c_int), ("K", C.c_int), ("pad", C.c_int)]


_vp = C.c_void_p
_sz = C.c_size_t
_i = C.c_int
_argp = C.POINTER(skl_arg)

SIGNATURES = {
    # runtime
    "skl_last_error": (C.c_char_p, []),
    "skl_free_string": (None, [_vp]),
    "skl_init": (_i, [_i, _i]),
    "skl_terminate": (_i, []),
    "skl_is_initialized": (_i, []),
    "skl_device_count": (_i, [C.POINTER(_i)]),
    "skl_device_get_info": (_i, [_i, C.POINTER(skl_device_info)]),
    "skl_device_synchronize": (_i, []),
    "skl_stream_create": (_i, [C.POINTER(_vp)]),
    "skl_stream_destroy": (_i, [_vp]),
    "skl_stream_synchronize": (_i, [_vp]),
    "skl_stream_add_callback": (_i, [_vp, _vp, _vp]),
    "skl_malloc": (_i, [C.POINTER(_vp), _sz]),
    "skl_free": (_i, [_vp]),
    "skl_host_malloc": (_i, [C.POINTER(_vp), _sz]),
    "skl_host_free": (_i, [_vp]),
    "skl_memcpy_h2d": (_i, [_vp, _sz, _vp, _sz, _vp, _i]),
    "skl_memcpy_d2h": (_i, [_vp, _vp, _sz, _sz, _vp, _i]),
    "skl_memcpy_d2d": (_i, [_vp, _sz, _vp, _sz, _sz, _vp]),
    "skl_fill": (_i, [_vp, _vp, _sz, _sz, _vp]),
    "skl_event_create": (_i, [C.POINTER(_vp)]),
    "skl_event_destroy": (_i, [_vp]),
    "skl_event_record": (_i, [_vp, _vp]),
    "skl_event_synchronize": (_i, [_vp]),
    "skl_event_elapsed_ms": (_i, [_vp, _vp, C.POINTER(C.c_float)]),
    "skl_program_compile": (_i, [C.c_char_p, C.c_char_p, C.POINTER(_vp), C.POINTER(_sz),
                                 C.POINTER(_vp)]),
    "skl_program_build": (_i, [C.c_char_p, C.c_char_p, C.POINTER(_vp), C.POINTER(_vp)]),
    "skl_program_release": (_i, [_vp]),
    "skl_kernel_get": (_i, [_vp, C.c_char_p, C.POINTER(_vp)]),
    "skl_kernel_launch": (_i, [_vp, C.POINTER(C.c_uint), C.POINTER(C.c_uint), _sz,
                               C.POINTER(_vp), _vp]),
    "skl_sha1_hex": (_i, [C.c_char_p, _sz, C.c_char_p]),
    "skl_graph_begin_capture": (_i, [_vp]),
    "skl_graph_end_capture": (_i, [_vp, C.POINTER(_vp)]),
    "skl_graph_launch": (_i, [_vp, _vp]),
    "skl_graph_destroy": (_i, [_vp]),
    "skl_generate_f32_sym": (_i, [_vp, _sz, C.c_uint64, _vp]),
    "skl_generate_f32_unit": (_i, [_vp, _sz, C.c_uint64, C.c_float, _vp]),
    "skl_generate_i32_below": (_i, [_vp, _sz, C.c_uint64, C.c_uint32, _vp]),
    # skeletons
    "skl_common_definition_append": (_i, [C.c_char_p, _i]),
    "skl_common_definitions": (_i, [C.POINTER(_vp)]),
    "skl_skeleton_create": (_i, [C.POINTER(skl_skeleton_desc), C.POINTER(_vp)]),
    "skl_skeleton_release": (_i, [_vp]),
    "skl_skeleton_set_work_group_size": (_i, [_vp, _i]),
    "skl_skeleton_work_group_size": (_i, [_vp, C.POINTER(_i)]),
    "skl_skeleton_set_precision": (_i, [_vp, _i]),
    "skl_skeleton_source": (_i, [_vp, C.POINTER(_vp)]),
    "skl_skeleton_check": (_i, [C.POINTER(skl_skeleton_desc), C.POINTER(_vp),
                                C.POINTER(_vp)]),
    "skl_skeleton_extra_count": (_i, [_vp, C.POINTER(_i)]),
    "skl_describe_user_function": (_i, [C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "skl_skeleton_extra_info": (_i, [_vp, _i, C.POINTER(_i), C.POINTER(_sz), C.c_char_p,
                                     _sz]),
    "skl_map_run": (_i, [_vp, _vp, _vp, _sz, _argp, _i, _vp]),
    "skl_map_index_run": (_i, [_vp, _vp, _sz, _sz, _argp, _i, _vp]),
    "skl_map_indexpoint_run": (_i, [_vp, _vp, _sz, _sz, _sz, _argp, _i, _vp]),
    "skl_zip_run": (_i, [_vp, _vp, _vp, _vp, _sz, _argp, _i, _vp]),
    "skl_reduce_run": (_i, [_vp, _vp, _vp, _sz, _argp, _i, _vp]),
    "skl_reduce_run_noalloc": (_i, [_vp, _vp, _vp, _sz, _argp, _i, _vp]),
    "skl_scan_run": (_i, [_vp, _vp, _vp, _sz, _i, _argp, _i, _vp]),
    "skl_mapoverlap_run": (_i, [_vp, _vp, _vp, _sz, _sz, _argp, _i, _vp]),
    "skl_allpairs_run": (_i, [_vp, _vp, _vp, _vp, _sz, _sz, _sz, _argp, _i, _vp]),
    "skl_zip_reduce_run": (_i, [_vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "skl_reduce_segmented_run": (_i, [_vp, _vp, _vp, _sz, _vp, _i, _vp]),
    "skl_scan_segmented_run": (_i, [_vp, _vp, _vp, _sz, _vp, _i, _i, _vp]),
    "skl_map_backward": (_i, [_vp, _vp, _vp, _vp, _sz, _argp, _i, _vp]),
    "skl_zip_backward": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _sz, _argp, _i, _vp]),
    "skl_reduce_backward": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "skl_scan_backward": (_i, [_vp, _vp, _vp, _sz, _i, _vp]),
    "skl_mapoverlap_stencil_backward": (_i, [_vp, _vp, _vp, _vp, _sz, _sz, _vp]),
    "skl_allpairs_backward": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _sz, _sz, _sz, _vp]),
    "skl_matmul_f32": (_i, [_vp, _vp, _vp, _sz, _sz, _sz, _i, _i, _i, _vp]),
    "skl_matmul_batched_f32": (_i, [_vp, _i, _i, _vp]),
    "skl_batch_run": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _sz, _vp, _i, _vp, _i, _i, _i, _vp]),
    "skl_batch_check": (_i, [_vp, C.POINTER(_i)]),
    "skl_autotune": (_i, [_vp, _sz, _i]),
    "skl_autotune_args": (_i, [_vp, _sz, _argp, _i]),
    "skl_skeleton_launch_shape": (_i, [_vp, C.POINTER(_i)]),
    "skl_skeleton_set_launch_shape": (_i, [_vp, C.POINTER(_i)]),
    "skl_skeleton_set_mma_tile": (_i, [_vp, _i]),
    "skl_skeleton_mma_tile": (_i, [_vp, C.POINTER(_i)]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m skelcl_amd.build` "
                "(there is no fallback path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class SkelError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


def check(status: int) -> None:
    if status != SKL_OK:
        msg = lib().skl_last_error()
        raise SkelError(status, msg.decode(errors="replace") if msg else "")


def take_string(ptr: C.c_void_p) -> str:
    if not ptr.value:
        return ""
    s = C.cast(ptr, C.c_char_p).value.decode(errors="replace")
    lib().skl_free_string(ptr)
    return s

"""numpy front-end of the oracle (liboracle.so, built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the reference its result is compared with, and by bench.py's
``cpu_baseline`` leg.  The product
(``skelcl_amd``) never imports it.

Every function here is a thin typed wrapper; the algorithms and their
reference citations live in oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_sz = ctypes.c_size_t
_int = ctypes.c_int


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_splitmix64": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
            "orc_gen_u64": (None, [ctypes.c_uint64, _sz, _u64p]),
            "orc_gen_f32_sym": (None, [ctypes.c_uint64, _sz, _f32p]),
            "orc_gen_f32_unit": (None, [ctypes.c_uint64, _sz, ctypes.c_float, _f32p]),
            "orc_gen_i32_below": (None, [ctypes.c_uint64, _sz, ctypes.c_uint32, _i32p]),
            "orc_gen_loguniform": (None, [ctypes.c_uint64, _sz, ctypes.c_double,
                                          ctypes.c_double, _i64p]),
            "orc_saxpy_f32": (None, [ctypes.c_float, _f32p, _f32p, _f32p, _sz, _int]),
            "orc_saxpy_f32_fma": (None, [ctypes.c_float, _f32p, _f32p, _f32p, _sz]),
            "orc_scale_f32": (None, [ctypes.c_float, _f32p, _f32p, _sz]),
            "orc_zip_mul_f32": (None, [_f32p, _f32p, _f32p, _sz]),
            "orc_reduce_add_f32_reforder": (ctypes.c_float, [_f32p, _sz, _sz, _int]),
            "orc_reduce_add_u32_reforder": (ctypes.c_uint32, [_u32p, _sz, _sz]),
            "orc_sum_f64": (ctypes.c_double, [_f32p, _sz]),
            "orc_dot_f64": (ctypes.c_double, [_f32p, _f32p, _sz]),
            "orc_scan_excl_add_f32_reforder": (None, [_f32p, _f32p, _sz, _sz, _int]),
            "orc_scan_excl_add_u32_reforder": (None, [_u32p, _u32p, _sz, _sz, _int]),
            "orc_scan_add_u32": (None, [_u32p, _u32p, _sz, _int]),
            "orc_scan_add_f64": (None, [_f32p, _f64p, _sz, _int]),
            "orc_scan_add_adjoint_f64": (None, [_f32p, _f64p, _sz, _int]),
            "orc_mapoverlap_stencil_f32": (None, [_f32p, _f32p, _int, _int, _f32p, _int,
                                                  _int, ctypes.c_float, _int]),
            "orc_mapoverlap_stencil_f64": (None, [_f32p, _f64p, _int, _int, _f32p, _int,
                                                  _int, ctypes.c_float, _i64p, _int]),
            "orc_mapoverlap_stencil_f32_fma": (None, [_f32p, _f32p, _int, _int, _f32p, _int,
                                                      _int, ctypes.c_float, _i64p, _int]),
            "orc_mapoverlap_stencil_adjoint_f64": (None, [_f32p, _f64p, _int, _int, _f32p,
                                                          _int, _int, _i64p, _int]),
            "orc_getdata_i32": (ctypes.c_int32, [_i32p, _int, _int, _int, _int, _int,
                                                 _int, _int, ctypes.c_int32]),
            "orc_allpairs_mm_f32_seqk": (None, [_f32p, _f32p, _f32p, _int, _int, _int,
                                                ctypes.c_float, ctypes.c_void_p, _int]),
            "orc_allpairs_mm_f64": (None, [_f32p, _f32p, _f64p, _int, _int, _int,
                                           ctypes.c_void_p, _int]),
            "orc_allpairs_sqdist_f32": (None, [_f32p, _f32p, _f32p, _int, _int, _int,
                                               ctypes.c_void_p, _int]),
            "orc_allpairs_sqdist_f64": (None, [_f32p, _f32p, _f64p, _int, _int, _int,
                                               ctypes.c_void_p, _int]),
            "orc_segmented_sum_f64": (None, [_f32p, _i64p, _int, _f64p]),
            "orc_segmented_scan_f64": (None, [_f32p, _i64p, _int, _int, _f64p]),
            "orc_segmented_sum_u32": (None, [_u32p, _i64p, _int, _u32p]),
            "orc_segmented_scan_u32": (None, [_u32p, _i64p, _int, _int, _u32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---- generator -------------------------------------------------------------
def splitmix64(seed: int, i: int) -> int:
    return int(lib().orc_splitmix64(seed, i))


def gen_u64(seed, n):
    out = np.empty(n, np.uint64)
    lib().orc_gen_u64(seed, n, out)
    return out


def gen_f32_sym(seed, n):
    out = np.empty(n, np.float32)
    lib().orc_gen_f32_sym(seed, n, out)
    return out


def gen_f32_unit(seed, n, scale=1.0):
    out = np.empty(n, np.float32)
    lib().orc_gen_f32_unit(seed, n, scale, out)
    return out


def gen_i32_below(seed, n, m):
    out = np.empty(n, np.int32)
    lib().orc_gen_i32_below(seed, n, m, out)
    return out


def gen_loguniform(seed, n, e0, span):
    out = np.empty(n, np.int64)
    lib().orc_gen_loguniform(seed, n, e0, span, out)
    return out


# ---- Map / Zip -----------------------------------------------------------------
def saxpy_f32(a, x, y, nthreads=1):
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    out = np.empty_like(x)
    lib().orc_saxpy_f32(a, x, y, out, x.size, nthreads)
    return out


def saxpy_f32_fma(a, x, y):
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    out = np.empty_like(x)
    lib().orc_saxpy_f32_fma(a, x, y, out, x.size)
    return out


def scale_f32(a, x):
    x = _c(x, np.float32)
    out = np.empty_like(x)
    lib().orc_scale_f32(a, x, out, x.size)
    return out


def zip_mul_f32(x, y):
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    out = np.empty_like(x)
    lib().orc_zip_mul_f32(x, y, out, x.size)
    return out


# ---- Reduce --------------------------------------------------------------------
def reduce_add_f32_reforder(x, wg=256, nthreads=1):
    x = _c(x, np.float32)
    return np.float32(lib().orc_reduce_add_f32_reforder(x, x.size, wg, nthreads))


def reduce_add_u32_reforder(x, wg=256):
    x = _c(np.asarray(x).view(np.uint32) if np.asarray(x).dtype == np.int32 else x,
           np.uint32)
    return np.uint32(lib().orc_reduce_add_u32_reforder(x, x.size, wg))


def sum_f64(x):
    x = _c(x, np.float32)
    return float(lib().orc_sum_f64(x, x.size))


def dot_f64(x, y):
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    return float(lib().orc_dot_f64(x, y, x.size))


# ---- Scan ----------------------------------------------------------------------
def scan_excl_add_f32_reforder(x, wg=256, nthreads=1):
    x = _c(x, np.float32)
    out = np.empty_like(x)
    lib().orc_scan_excl_add_f32_reforder(x, out, x.size, wg, nthreads)
    return out


def scan_excl_add_u32_reforder(x, wg=256, nthreads=1):
    x = np.ascontiguousarray(x)
    if x.dtype == np.int32:
        x = x.view(np.uint32)
    x = _c(x, np.uint32)
    out = np.empty_like(x)
    lib().orc_scan_excl_add_u32_reforder(x, out, x.size, wg, nthreads)
    return out


def scan_add_u32(x, inclusive):
    x = np.ascontiguousarray(x)
    if x.dtype == np.int32:
        x = x.view(np.uint32)
    x = _c(x, np.uint32)
    out = np.empty_like(x)
    lib().orc_scan_add_u32(x, out, x.size, int(bool(inclusive)))
    return out


def scan_add_f64(x, inclusive):
    x = _c(x, np.float32)
    out = np.empty(x.size, np.float64)
    lib().orc_scan_add_f64(x, out, x.size, int(bool(inclusive)))
    return out


def scan_add_adjoint_f64(g, inclusive):
    g = _c(g, np.float32)
    out = np.empty(g.size, np.float64)
    lib().orc_scan_add_adjoint_f64(g, out, g.size, int(bool(inclusive)))
    return out


# ---- MapOverlap ----------------------------------------------------------------
NEUTRAL, NEAREST = 0, 1


def mapoverlap_stencil_f32(img, w, R, padding=NEUTRAL, neutral=0.0, nthreads=1):
    img = _c(img, np.float32)
    w = _c(w, np.float32)
    rows, cols = img.shape
    out = np.empty_like(img)
    lib().orc_mapoverlap_stencil_f32(img, out, rows, cols, w.ravel(), R, padding,
                                     neutral, nthreads)
    return out


def mapoverlap_stencil_f64(img, w, R, pix, padding=NEUTRAL, neutral=0.0):
    img = _c(img, np.float32)
    w = _c(w, np.float32)
    pix = _c(pix, np.int64)
    rows, cols = img.shape
    out = np.empty(pix.size, np.float64)
    lib().orc_mapoverlap_stencil_f64(img, out, rows, cols, w.ravel(), R, padding,
                                     neutral, pix, pix.size)
    return out


def mapoverlap_stencil_f32_fma(img, w, R, pix, padding=NEUTRAL, neutral=0.0):
    """The stencil as a contracting compiler evaluates `sum += getData * w`:
    one fmaf per term in dy-outer, dx-inner order, at the pixels `pix`."""
    img = _c(img, np.float32)
    w = _c(w, np.float32)
    pix = _c(pix, np.int64)
    rows, cols = img.shape
    out = np.empty(pix.size, np.float32)
    lib().orc_mapoverlap_stencil_f32_fma(img, out, rows, cols, w.ravel(), R, padding,
                                         neutral, pix, pix.size)
    return out


def mapoverlap_stencil_adjoint_f64(g, w, R, pix, padding=NEUTRAL):
    g = _c(g, np.float32)
    w = _c(w, np.float32)
    pix = _c(pix, np.int64)
    rows, cols = g.shape
    out = np.empty(pix.size, np.float64)
    lib().orc_mapoverlap_stencil_adjoint_f64(g, out, rows, cols, w.ravel(), R,
                                             padding, pix, pix.size)
    return out


def getdata_i32(img, r, c, dx, dy, padding=NEAREST, neutral=0):
    img = _c(img, np.int32)
    rows, cols = img.shape
    return int(lib().orc_getdata_i32(img, rows, cols, r, c, dx, dy, padding, neutral))


# ---- AllPairs ------------------------------------------------------------------
def _rows_arg(rows):
    if rows is None:
        return None, 0, None
    r = _c(rows, np.int32)
    return r.ctypes.data_as(ctypes.c_void_p), r.size, r


def allpairs_mm_f32_seqk(A, B, identity=0.0, rows=None):
    A = _c(A, np.float32)
    B = _c(B, np.float32)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    rp, nr, keep = _rows_arg(rows)
    C = np.empty((nr if rows is not None else M, N), np.float32)
    lib().orc_allpairs_mm_f32_seqk(A, B, C, M, N, K, identity, rp, nr)
    del keep
    return C


def allpairs_mm_f64(A, B, rows=None):
    A = _c(A, np.float32)
    B = _c(B, np.float32)
    M, K = A.shape
    _, N = B.shape
    rp, nr, keep = _rows_arg(rows)
    C = np.empty((nr if rows is not None else M, N), np.float64)
    lib().orc_allpairs_mm_f64(A, B, C, M, N, K, rp, nr)
    del keep
    return C


def allpairs_sqdist_f32(L, Rt, rows=None):
    L = _c(L, np.float32)
    Rt = _c(Rt, np.float32)
    nl, d = L.shape
    _, nr_ = Rt.shape
    rp, nr, keep = _rows_arg(rows)
    C = np.empty((nr if rows is not None else nl, nr_), np.float32)
    lib().orc_allpairs_sqdist_f32(L, Rt, C, nl, nr_, d, rp, nr)
    del keep
    return C


def allpairs_sqdist_f64(L, Rt, rows=None):
    L = _c(L, np.float32)
    Rt = _c(Rt, np.float32)
    nl, d = L.shape
    _, nr_ = Rt.shape
    rp, nr, keep = _rows_arg(rows)
    C = np.empty((nr if rows is not None else nl, nr_), np.float64)
    lib().orc_allpairs_sqdist_f64(L, Rt, C, nl, nr_, d, rp, nr)
    del keep
    return C


# ---- segmented -------------------------------------------------------------------
def _check_offsets(offsets, n):
    if offsets.size < 1 or offsets[0] < 0 or offsets[-1] > n or np.any(np.diff(offsets) < 0):
        raise ValueError("offsets must be non-decreasing and inside [0, n]")


def segmented_sum_f64(x, offsets):
    x = _c(x, np.float32)
    offsets = _c(offsets, np.int64)
    _check_offsets(offsets, x.size)
    out = np.empty(offsets.size - 1, np.float64)
    lib().orc_segmented_sum_f64(x, offsets, offsets.size - 1, out)
    return out


def segmented_scan_f64(x, offsets, inclusive):
    x = _c(x, np.float32)
    offsets = _c(offsets, np.int64)
    _check_offsets(offsets, x.size)
    out = np.empty(x.size, np.float64)
    lib().orc_segmented_scan_f64(x, offsets, offsets.size - 1, int(bool(inclusive)), out)
    return out


def segmented_sum_u32(x, offsets):
    """Per-segment sums of 32-bit integers modulo 2^32 (uint32 out)."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.int32:
        x = x.view(np.uint32)
    x = _c(x, np.uint32)
    offsets = _c(offsets, np.int64)
    _check_offsets(offsets, x.size)
    out = np.empty(offsets.size - 1, np.uint32)
    lib().orc_segmented_sum_u32(x, offsets, offsets.size - 1, out)
    return out


def segmented_scan_u32(x, offsets, inclusive, fill=0):
    """Per-segment scans of 32-bit integers modulo 2^32 (uint32 out);
    elements outside [offsets[0], offsets[-1]) hold `fill`."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.int32:
        x = x.view(np.uint32)
    x = _c(x, np.uint32)
    offsets = _c(offsets, np.int64)
    _check_offsets(offsets, x.size)
    out = np.full(x.size, fill, np.uint32)
    lib().orc_segmented_scan_u32(x, offsets, offsets.size - 1, int(bool(inclusive)), out)
    return out


def gaussian_weights(R=5, sigma=2.0):
    """(2R+1)^2 normalised Gaussian table, computed in FP64, rounded to f32:
    w(i,j) = exp(-(i^2+j^2)/(2 sigma^2)) / sum."""
    ax = np.arange(-R, R + 1, dtype=np.float64)
    g = np.exp(-(ax[:, None] ** 2 + ax[None, :] ** 2) / (2.0 * sigma * sigma))
    return (g / g.sum()).astype(np.float32)

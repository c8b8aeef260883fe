"""Parity at the benchmark configurations' full sizes, checked against the
oracle on the same seeded inputs:

* Reduce(+) f32 over 2^28 elements against the FP64 sum;
* saxpy (Zip a*x+y, a = 2.5) over 2^28 elements, bit-exact against fmaf;
* the 11 x 11 Gaussian MapOverlap over an 8192 x 8192 image, bit-exact
  against the oracle's fmaf chain at every pixel within R of a border and
  4096 interior ones, both paddings;
* the 8192^3 AllPairs(+,*) product in FP32, BF16X3 and FP8 on 64 sampled
  rows against FP64 (normwise), and squared-Euclidean AllPairs on 16384
  points of dim 256 on sampled rows;
* the batched configuration in full: 4096 segments (the committed lengths)
  for segmented Reduce and Scan, and the 512 products; and the same in one
  skl_batch_run launch.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle
from skelcl_amd import _lib

pytestmark = pytest.mark.gpu

N28 = 1 << 28
ADD_F = "float func(float x, float y){ return x+y; }"
MUL_F = "float func(float x, float y){ return x*y; }"


def test_reduce_f32_full_2e28(skl):
    x = oracle.gen_f32_sym(0x2002, N28)
    r = skl.Reduce("float(float)", ADD_F)
    got = float(r(skl.Vector(x))[0])
    ref = oracle.sum_f64(x)
    tol = 2 * np.spacing(np.float32(abs(ref))) + 1e-6 * np.abs(x).sum() * 2.0 ** -24
    print(f"reduce 2^28: gpu={got!r} fp64={ref!r} err={abs(got - ref):.3e} tol={tol:.3e}")
    assert abs(got - ref) <= tol


def test_saxpy_full_2e28_bit_exact(skl):
    x = oracle.gen_f32_sym(0x2003, N28)
    y = oracle.gen_f32_sym(0x2004, N28)
    z = skl.Zip("float(float, float)", "float func(float x, float y, float a){ return a*x+y; }")
    got = np.asarray(z(skl.Vector(x), skl.Vector(y), np.float32(2.5)))
    ref = oracle.saxpy_f32_fma(np.float32(2.5), x, y)
    bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, (bad[:10], got[bad[:10]], ref[bad[:10]])


GAUSS = """float func(input_matrix_t img, __global float* w) {
  float sum = 0.0f;
  for (int dy = -5; dy <= 5; ++dy)
    for (int dx = -5; dx <= 5; ++dx)
      sum += getData(img, dx, dy) * w[(dy + 5) * 11 + (dx + 5)];
  return sum;
}"""


def _border_and_interior(rows, cols, R, n_interior, seed):
    r, c = np.mgrid[0:rows, 0:cols]
    edge = (r < R) | (r >= rows - R) | (c < R) | (c >= cols - R)
    pix = np.flatnonzero(edge.ravel())
    rng = np.random.default_rng(seed)
    inner = rng.integers(0, rows * cols, 4 * n_interior)
    inner = inner[~edge.ravel()[inner]][:n_interior]
    return np.unique(np.concatenate([pix, inner])).astype(np.int64)


@pytest.mark.parametrize("padding", [0, 1])
def test_mapoverlap_gauss_8192_bit_exact(skl, padding):
    rows = cols = 8192
    img = oracle.gen_f32_unit(0x3001, rows * cols, 255.0).reshape(rows, cols)
    w = oracle.gaussian_weights(5, 2.0)
    m = skl.MapOverlap("float(float)", GAUSS, 5, padding, 0)
    out = np.asarray(m(skl.Matrix(img), skl.Vector(w.ravel()))).ravel()
    pix = _border_and_interior(rows, cols, 5, 4096, 0x3003)
    ref = oracle.mapoverlap_stencil_f32_fma(img, w, 5, pix, padding, 0.0)
    got = out[pix]
    bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
    print(f"8192^2 padding={padding}: {pix.size} pixels, {bad.size} differ")
    assert bad.size == 0, (pix[bad[:10]], got[bad[:10]], ref[bad[:10]])
    # the FP64 stencil on the same pixels, for the record
    r64 = oracle.mapoverlap_stencil_f64(img, w, 5, pix[:2000], padding)
    assert np.max(np.abs(got[:2000] - r64) / np.maximum(np.abs(r64), 1.0)) < 1e-5


def _matmul(A_dev, B_dev, M, N, K, prec):
    C_dev = _dev(4 * M * N)
    _lib.check(_lib.lib().skl_matmul_f32(C_dev.ptr, A_dev.ptr, B_dev.ptr, M, N, K, 0, 0, prec,
                                         None))
    return C_dev


def _dev(nbytes):
    import skelcl_amd as skl
    return skl.DeviceBuffer(nbytes)


def _rows_of(C_dev, rows, N):
    out = np.empty((len(rows), N), np.float32)
    for i, r in enumerate(rows):
        C_dev.download(out[i], offset=int(r) * N * 4)
    return out


def test_allpairs_matmul_8192_sampled_rows(skl):
    """C = A.B at 8192^3: FP32 within 5e-6 normwise of FP64, BF16X3 no worse
    than FP32, FP8 (per-128-tile scales) within 5e-2, on 64 sampled rows
    (the first, the last and 62 seeded ones)."""
    n = 8192
    A = oracle.gen_f32_sym(0x4001, n * n).reshape(n, n)
    B = oracle.gen_f32_sym(0x4002, n * n).reshape(n, n)
    rng = np.random.default_rng(0x4005)
    rows = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, 62)])).astype(np.int32)
    ref = oracle.allpairs_mm_f64(A, B, rows=rows)
    dA, dB = skl.DeviceBuffer.from_array(A), skl.DeviceBuffer.from_array(B)
    errs = {}
    for prec in (skl.PREC_FP32, skl.PREC_BF16X3, skl.PREC_FP8):
        Cd = _matmul(dA, dB, n, n, n, prec)
        got = _rows_of(Cd, rows, n).astype(np.float64)
        errs[prec] = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        del Cd
    print(f"8192^3 normwise errors: fp32 {errs[1]:.3e} bf16x3 {errs[2]:.3e} fp8 {errs[3]:.3e}")
    assert errs[skl.PREC_FP32] < 5e-6
    assert errs[skl.PREC_BF16X3] <= errs[skl.PREC_FP32]
    assert errs[skl.PREC_FP8] < 5e-2


SQDIST = ("float func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim) {"
          " float s = 0; for (unsigned k = 0; k < dim; ++k) {"
          " float t = getElementFromRow(row, k) - getElementFromColumn(col, k); s += t*t; }"
          " return s; }")


def test_allpairs_sqdist_16384x256_sampled_rows(skl):
    n, d = 16384, 256
    P = oracle.gen_f32_sym(0x4003, n * d).reshape(n, d)
    Pt = np.ascontiguousarray(P.T)
    ap = skl.AllPairs("float(float, float)", SQDIST)
    out = np.asarray(ap(skl.Matrix(P), skl.Matrix(Pt)))
    rng = np.random.default_rng(0x4006)
    rows = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, 62)])).astype(np.int32)
    ref = oracle.allpairs_sqdist_f64(P, Pt, rows=rows)
    got = out[rows].astype(np.float64)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.abs(ref).max()
    assert np.all(np.diag(out[rows][:, rows]) == 0.0)  # d(p, p) = 0 exactly


def _cfg5():
    with open(os.path.join(os.path.dirname(__file__), "golden", "batched_shapes.json")) as f:
        return json.load(f)


def test_batched_segments_full_config(skl):
    """All 4096 segments (about 5.8e8 elements): every segmented Reduce
    result, every segment's last inclusive-Scan value against its FP64 sum,
    and 48 whole segments (the longest, the shortest and seeded ones) of
    both Scans against FP64 prefixes."""
    lens = np.array(_cfg5()["segment_lengths"], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = oracle.gen_f32_sym(0x5002, int(off[-1]))
    X = skl.Vector(x)
    sums = oracle.segmented_sum_f64(x, off)
    r = skl.Reduce("float(float)", ADD_F)
    got = np.asarray(r.segmented(X, off)).astype(np.float64)
    tol = np.spacing(np.abs(sums).astype(np.float32)).astype(np.float64) + 1e-12
    assert np.all(np.abs(got - sums) <= tol)
    s = skl.Scan("float(float)", ADD_F, "0.0f")
    order = np.argsort(lens)
    rng = np.random.default_rng(0x5007)
    pick = np.unique(np.concatenate([order[:8], order[-8:], rng.integers(0, lens.size, 32)]))
    for incl in (False, True):
        out = np.asarray(s.segmented(X, off, inclusive=incl))
        if incl:
            last = out[off[1:] - 1].astype(np.float64)
            scale = np.array([np.abs(x[a:b]).astype(np.float64).sum() for a, b in
                              zip(off[:-1], off[1:])])
            assert np.all(np.abs(last - sums) <= np.spacing(np.abs(sums).astype(np.float32))
                          + 1e-7 * scale * 2.0 ** -10)
        for k in pick:
            a, b = int(off[k]), int(off[k + 1])
            seg = x[a:b].astype(np.float64)
            ref = np.cumsum(seg) if incl else np.concatenate([[0.0], np.cumsum(seg)[:-1]])
            t = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
            assert np.all(np.abs(out[a:b].astype(np.float64) - ref) <= t), (incl, k)


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_products_full_config(skl, prec):
    """All 512 products of the batched configuration in one call, each
    against numpy FP64 (normwise 1e-6)."""
    shapes = _cfg5()["gemm_shapes"]
    descs = (_lib.skl_gemm_desc * len(shapes))()
    keep, refs = [], []
    for i, (M, N, K) in enumerate(shapes):
        A = oracle.gen_f32_sym(0x5004 + 2 * i, M * K).reshape(M, K)
        B = oracle.gen_f32_sym(0x5005 + 2 * i, K * N).reshape(K, N)
        dA, dB = skl.DeviceBuffer.from_array(A), skl.DeviceBuffer.from_array(B)
        dC = skl.DeviceBuffer(4 * M * N)
        descs[i].A, descs[i].B, descs[i].C = dA.ptr, dB.ptr, dC.ptr
        descs[i].M, descs[i].N, descs[i].K = M, N, K
        keep += [dA, dB]
        refs.append((dC, M, N, A.astype(np.float64) @ B.astype(np.float64)))
    dd = skl.DeviceBuffer(C.sizeof(descs))
    _lib.check(_lib.lib().skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs),
                                         None, 1))
    _lib.check(_lib.lib().skl_matmul_batched_f32(dd.ptr, len(shapes), prec, None))
    worst = 0.0
    for dC, M, N, ref in refs:
        c = dC.to_numpy(np.float32, M * N).reshape(M, N)
        worst = max(worst, np.linalg.norm(c - ref) / np.linalg.norm(ref))
    print(f"512 products prec={prec}: worst normwise error {worst:.3e}")
    assert worst < 1e-6


def test_reduce_all_sizes_8192_to_8447(skl):  # ReduceTests.cpp:155-170, every size
    r = skl.Reduce("float(float)", ADD_F)
    buf = skl.DeviceBuffer.from_array(np.ones(8447, np.float32))
    out = skl.DeviceBuffer(16)
    res = np.empty(1, np.float32)
    for n in range(8192, 8448):
        _lib.check(_lib.lib().skl_reduce_run(r.handle, out.ptr, buf.ptr, n, None, 0, None))
        out.download(res)
        assert res[0] == n, n


def test_batch_one_launch_full_config(skl):
    """The batched configuration in one skl_batch_run launch: all 4096
    segments (sums and inclusive scan from one read of the 5.8e8 elements)
    and all 512 products at BF16X3, two product work-groups per CU.  Every
    sum against FP64, every segment's last scan value against its sum, 48
    whole segments of the scan against FP64 prefixes, no offsets flagged,
    and the products bit for bit against skl_matmul_batched_f32 (the same
    tile code)."""
    cfg = _cfg5()
    lens = np.array(cfg["segment_lengths"], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n, nseg = int(off[-1]), lens.size
    x = oracle.gen_f32_sym(0x5012, n)
    X, O = skl.DeviceBuffer.from_array(x), skl.DeviceBuffer.from_array(off)
    S, Y = skl.DeviceBuffer(4 * nseg), skl.DeviceBuffer(4 * n)
    shapes = cfg["gemm_shapes"]
    d1 = (_lib.skl_gemm_desc * len(shapes))()
    d2 = (_lib.skl_gemm_desc * len(shapes))()
    keep, outs = [], []
    for i, (M, N, K) in enumerate(shapes):
        dA = skl.DeviceBuffer.from_array(oracle.gen_f32_sym(0x5104 + 2 * i, M * K))
        dB = skl.DeviceBuffer.from_array(oracle.gen_f32_sym(0x5105 + 2 * i, K * N))
        c1, c2 = skl.DeviceBuffer(4 * M * N), skl.DeviceBuffer(4 * M * N)
        for d, c in ((d1, c1), (d2, c2)):
            d[i].A, d[i].B, d[i].C = dA.ptr, dB.ptr, c.ptr
            d[i].M, d[i].N, d[i].K = M, N, K
        keep += [dA, dB]
        outs.append((c1, c2, M * N))
    dd = []
    for d in (d1, d2):
        b = skl.DeviceBuffer(C.sizeof(d))
        _lib.check(_lib.lib().skl_memcpy_h2d(b.ptr, 0, C.cast(d, C.c_void_p), C.sizeof(d),
                                             None, 1))
        dd.append(b)
    r = skl.Reduce("float(float)", ADD_F)
    s = skl.Scan("float(float)", ADD_F, "0.0f")
    _lib.check(_lib.lib().skl_batch_run(r.handle, S.ptr, s.handle, Y.ptr, 1, X.ptr, n, O.ptr,
                                        nseg, dd[0].ptr, len(shapes), 2, 2, None))
    bad = C.c_int(-1)
    _lib.check(_lib.lib().skl_batch_check(s.handle, C.byref(bad)))
    assert bad.value == 0
    _lib.check(_lib.lib().skl_matmul_batched_f32(dd[1].ptr, len(shapes), 2, None))
    for c1, c2, mn in outs:
        assert np.array_equal(c1.to_numpy(np.float32, mn).view(np.uint32),
                              c2.to_numpy(np.float32, mn).view(np.uint32))
    sums = oracle.segmented_sum_f64(x, off)
    got = S.to_numpy(np.float32, nseg).astype(np.float64)
    tol = np.spacing(np.abs(sums).astype(np.float32)).astype(np.float64) + 1e-12
    assert np.all(np.abs(got - sums) <= tol)
    out = Y.to_numpy(np.float32, n)
    last = out[off[1:] - 1].astype(np.float64)
    scale = np.array([np.abs(x[a:b]).astype(np.float64).sum() for a, b in zip(off[:-1], off[1:])])
    assert np.all(np.abs(last - sums) <= np.spacing(np.abs(sums).astype(np.float32))
                  + 1e-7 * scale * 2.0 ** -10)
    order = np.argsort(lens)
    rng = np.random.default_rng(0x5013)
    pick = np.unique(np.concatenate([order[:8], order[-8:], rng.integers(0, nseg, 32)]))
    for k in pick:
        a, b = int(off[k]), int(off[k + 1])
        ref = np.cumsum(x[a:b].astype(np.float64))
        t = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
        assert np.all(np.abs(out[a:b].astype(np.float64) - ref) <= t), k

"""MapOverlap and AllPairs on the GPU: the reference's known-answer tests
(test/MapOverlapTests.cpp, test/AllPairsTests.cpp) restated, and parity with
the oracle / FP64 at stated tolerances."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GAUSS_SRC = """float func(input_matrix_t img, __global float* w) {
  float sum = 0.0f;
  for (int dy = -5; dy <= 5; ++dy)
    for (int dx = -5; dx <= 5; ++dx)
      sum += getData(img, dx, dy) * w[(dy + 5) * 11 + (dx + 5)];
  return sum;
}"""


def _shift(inp, dx, dy):
    rows, cols = inp.shape
    r = np.clip(np.arange(rows)[:, None] + dy, 0, rows - 1)
    c = np.clip(np.arange(cols)[None, :] + dx, 0, cols - 1)
    return inp[r, c]


@pytest.mark.parametrize("dx,dy,rng,size", [(-1, 0, 1, 1024), (-2, 0, 2, 100), (1, 0, 1, 100),
                                            (2, 0, 2, 100), (0, -1, 1, 100), (0, 1, 1, 100),
                                            (0, -2, 2, 100), (0, 2, 2, 100), (3, -4, 5, 257)])
def test_mapoverlap_shift_kats(skl, dx, dy, rng, size):  # MapOverlapTests.cpp:187-447
    m = skl.MapOverlap("int(int)", f"int func(input_matrix_t f){{ return getData(f, {dx}, {dy}); }}",
                       rng)
    i, j = np.mgrid[0:size, 0:size]
    inp = (i * j + i).astype(np.int32)
    out = np.asarray(m(skl.Matrix(inp)))
    assert np.array_equal(out, _shift(inp, dx, dy))


def test_mapoverlap_neutral_nonsquare(skl):
    m = skl.MapOverlap("int(int)", "int func(input_matrix_t f){ return getData(f, -1, -1) + "
                       "getData(f, 1, 1); }", 1, skl.NEUTRAL, -7)
    inp = np.arange(37 * 53, dtype=np.int32).reshape(37, 53)
    out = np.asarray(m(skl.Matrix(inp)))
    p = np.pad(inp, 1, constant_values=-7)
    assert np.array_equal(out, p[:-2, :-2] + p[2:, 2:])


@pytest.mark.parametrize("rows,cols,padding", [(64, 64, 0), (100, 37, 1), (257, 513, 0),
                                               (1024, 1000, 1)])
def test_mapoverlap_gauss_parity(skl, rows, cols, padding):
    """11x11 weighted stencil, every pixel: bit-exact vs the oracle's fmaf
    chain (the user loop as the default contraction evaluates it); within
    1e-5 relative of FP64."""
    w = oracle.gaussian_weights(5, 2.0)
    img = oracle.gen_f32_unit(0x3001, rows * cols, 255.0).reshape(rows, cols)
    m = skl.MapOverlap("float(float)", GAUSS_SRC, 5, padding, 0)
    out = np.asarray(m(skl.Matrix(img), skl.Vector(w.ravel()))).ravel()
    pix = np.arange(rows * cols, dtype=np.int64)
    fma = oracle.mapoverlap_stencil_f32_fma(img, w, 5, pix, padding, 0.0)
    assert np.array_equal(out.view(np.uint32), fma.view(np.uint32))
    sub = pix[::97]
    ref = oracle.mapoverlap_stencil_f64(img, w, 5, sub, padding)
    assert np.max(np.abs(out[sub] - ref) / np.maximum(1.0, np.abs(ref))) < 1e-5


@pytest.mark.parametrize("padding", [0, 1])
def test_mapoverlap_gauss_contraction_off_bit_exact(skl, padding):
    """With the parity setting (-ffp-contract=off) the stencil is bit-exact
    vs the oracle's two-rounding form (each product, then each sum, rounded
    as written)."""
    rows, cols = 300, 257
    w = oracle.gaussian_weights(5, 2.0)
    img = oracle.gen_f32_unit(0x3004, rows * cols, 255.0).reshape(rows, cols)
    m = skl.MapOverlap("float(float)", GAUSS_SRC, 5, padding, 0,
                       build_options="-ffp-contract=off")
    out = np.asarray(m(skl.Matrix(img), skl.Vector(w.ravel())))
    o32 = oracle.mapoverlap_stencil_f32(img, w, 5, padding)
    assert np.array_equal(out.view(np.uint32), o32.view(np.uint32))


# ------------------------------------------------------------------ AllPairs
def _ints(seed, shape):
    return (oracle.gen_i32_below(seed, int(np.prod(shape)), 10).reshape(shape)).astype(np.float32)


@pytest.mark.parametrize("M,K,N", [(64, 64, 64), (32, 128, 32), (100, 60, 11), (1, 1, 1),
                                   (129, 257, 65)])
def test_allpairs_matmul_kats(skl, M, K, N):  # AllPairsTests.cpp:77-133
    zip_ = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    red = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    ap = skl.AllPairs("float(float, float)", reduce=red, zip=zip_)
    A, B = _ints(1, (M, K)), _ints(2, (K, N))
    out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B)))
    assert np.array_equal(out, (A.astype(np.float64) @ B).astype(np.float32))


def test_allpairs_user_args_kat(skl):  # AllPairsTests.cpp:136-173
    zip_ = skl.Zip("float(float, float)", "float func(float x, float y, float a){ return x*y+a; }")
    red = skl.Reduce("float(float)", "float func(float x, float y, float b){ return x+y+b; }")
    ap = skl.AllPairs("float(float, float)", reduce=red, zip=zip_)
    A, B = _ints(3, (16, 8)), _ints(4, (8, 12))
    out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B), 2.0, 1.0))  # reduce's b, then zip's a
    exp = np.zeros((16, 12), np.float32)
    for k in range(8):
        exp = (exp + (A[:, k:k + 1] * B[k:k + 1, :] + np.float32(1))) + np.float32(2)
    assert np.array_equal(out, exp)


def test_allpairs_plain_fp32_bit_exact_vs_fma_fold(skl):
    """The FP32 (+,*) path folds k in order with fused multiply-adds from the
    identity: bit-identical to the oracle's sequential fold in FP64-then-round
    form for the integer KAT and within 1e-6 of FP64 on random data."""
    zip_ = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    red = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    ap = skl.AllPairs("float(float, float)", reduce=red, zip=zip_)
    A = oracle.gen_f32_sym(0x4001, 200 * 300).reshape(200, 300)
    B = oracle.gen_f32_sym(0x4002, 300 * 150).reshape(300, 150)
    out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B))).astype(np.float64)
    ref = oracle.allpairs_mm_f64(A, B)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-6


@pytest.mark.parametrize("n,d", [(50, 20), (100, 256), (33, 300), (70, 256), (130, 64)])
def test_allpairs_general_sqdist(skl, n, d):
    """d <= 256: 64 columns staged in LDS, rows read as wave-uniform scalar
    loads (16-byte staging when n % 4 == 0, element staging otherwise, partial
    64 x 64 tiles); d = 300: the form reading both operands from memory."""
    src = ("float func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim) {"
           " float s = 0; for (unsigned k = 0; k < dim; ++k) {"
           " float t = getElementFromRow(row, k) - getElementFromColumn(col, k); s += t*t; }"
           " return s; }")
    ap = skl.AllPairs("float(float, float)", src)
    P = oracle.gen_f32_sym(0x4003, n * d).reshape(n, d)
    Pt = np.ascontiguousarray(P.T)
    out = np.asarray(ap(skl.Matrix(P), skl.Matrix(Pt))).astype(np.float64)
    ref = oracle.allpairs_sqdist_f64(P, Pt)
    assert np.max(np.abs(out - ref)) <= 1e-5 * max(1.0, np.abs(ref).max())


def test_allpairs_general_mixed_width_types(skl):
    """float rows against double columns at d = 150: the columns no longer fit
    the 64-column staging (64 x 150 doubles > 64 KB) but 32 rows and 32
    columns do, which selects the form staging both in LDS."""
    src = ("double func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim) {"
           " double s = 0; for (unsigned k = 0; k < dim; ++k) {"
           " double t = getElementFromRow(row, k) - getElementFromColumn(col, k); s += t*t; }"
           " return s; }")
    ap = skl.AllPairs("double(float, double)", src)
    n, d = 45, 150
    P = oracle.gen_f32_sym(0x4005, n * d).reshape(n, d)
    Q = oracle.gen_f32_sym(0x4006, d * 37).reshape(d, 37).astype(np.float64) * 0.5
    out = np.asarray(ap(skl.Matrix(P), skl.Matrix(Q))).astype(np.float64)
    ref = ((P.astype(np.float64)[:, :, None] - Q[None, :, :]) ** 2).sum(axis=1)
    assert np.max(np.abs(out - ref)) <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n_cols", [36, 70])
def test_allpairs_general_int_exact(skl, n_cols):
    """int rows and columns through the column-staged kernel: 16-byte staging
    copies the 4-byte elements as raw bits (n_cols % 4 == 0), element staging
    otherwise; the products are exact in int32."""
    src = ("int func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim) {"
           " int s = 0; for (unsigned k = 0; k < dim; ++k)"
           " s += getElementFromRow(row, k) * getElementFromColumn(col, k); return s; }")
    ap = skl.AllPairs("int(int, int)", src)
    A = oracle.gen_i32_below(0x4101, 67 * 64, 100).reshape(67, 64).astype(np.int32) - 50
    B = oracle.gen_i32_below(0x4102, 64 * n_cols, 100).reshape(64, n_cols).astype(np.int32) - 50
    out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B)))
    assert np.array_equal(out, (A.astype(np.int64) @ B.astype(np.int64)).astype(np.int32))


def test_allpairs_general_extra_arg_kat(skl):  # AllPairsTests.cpp:226-269
    src = ("float func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim, int start) {"
           "float res = 0; for (int i = start; i < dim; ++i) { res += getElementFromRow(row, i) *"
           " getElementFromColumn(col, i); } return res; }")
    ap = skl.AllPairs("float(float, float)", src)
    A, B = _ints(7, (20, 10)), _ints(8, (10, 30))
    out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B), 3))
    assert np.array_equal(out, (A[:, 3:].astype(np.float64) @ B[3:]).astype(np.float32))


# --------------------------------------------------------- precision forms
def _matmul(skl, A, B, prec, ta=0, tb=0):
    from skelcl_amd import _lib
    M = A.shape[1] if ta else A.shape[0]
    N = B.shape[0] if tb else B.shape[1]
    K = A.shape[0] if ta else A.shape[1]
    dA, dB = skl.DeviceBuffer.from_array(A), skl.DeviceBuffer.from_array(B)
    dC = skl.DeviceBuffer(4 * M * N)
    _lib.check(_lib.lib().skl_matmul_f32(dC.ptr, dA.ptr, dB.ptr, M, N, K, ta, tb, prec, None))
    return dC.to_numpy(np.float32, M * N).reshape(M, N)


@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (128, 128, 128), (200, 136, 300),
                                   (1000, 1000, 1000)])
def test_matmul_bf16x3_accuracy(skl, M, N, K):
    """bf16x3 (3-term split, 6 MFMA products) must be no less accurate than
    plain FP32 (normwise, vs FP64)."""
    A = oracle.gen_f32_sym(0x4001, M * K).reshape(M, K)
    B = oracle.gen_f32_sym(0x4002, K * N).reshape(K, N)
    ref = oracle.allpairs_mm_f64(A, B)
    c3 = _matmul(skl, A, B, skl.PREC_BF16X3).astype(np.float64)
    c32 = _matmul(skl, A, B, skl.PREC_FP32).astype(np.float64)
    e3 = np.linalg.norm(c3 - ref) / np.linalg.norm(ref)
    e32 = np.linalg.norm(c32 - ref) / np.linalg.norm(ref)
    print(f"{M}x{N}x{K}: bf16x3 {e3:.3e}  fp32 {e32:.3e}")
    assert e32 < 1e-6
    assert e3 <= max(e32, 2e-7)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_matmul_bf16x3_large_tiles_ragged(skl, ta, tb):
    """Shapes with at least one 256 x 128 tile per CU take the 8-wave kernel:
    ragged M, N, K (A padded to a pair of 128-row blocks), both transposes."""
    M, N, K = 4100, 4000, 200
    A = oracle.gen_f32_sym(0x4101, M * K).reshape((K, M) if ta else (M, K))
    B = oracle.gen_f32_sym(0x4102, K * N).reshape((N, K) if tb else (K, N))
    Am = np.ascontiguousarray(A.T if ta else A).astype(np.float64)
    Bm = np.ascontiguousarray(B.T if tb else B).astype(np.float64)
    ref = Am @ Bm
    c3 = _matmul(skl, A, B, skl.PREC_BF16X3, ta, tb).astype(np.float64)
    c32 = _matmul(skl, A, B, skl.PREC_FP32, ta, tb).astype(np.float64)
    e3 = np.linalg.norm(c3 - ref) / np.linalg.norm(ref)
    e32 = np.linalg.norm(c32 - ref) / np.linalg.norm(ref)
    print(f"{M}x{N}x{K} ta={ta} tb={tb}: bf16x3 {e3:.3e}  fp32 {e32:.3e}")
    assert e3 <= max(e32, 2e-7)
    # every row and column is covered (no tile left out at the ragged edges)
    assert np.abs(c3 - ref).max() <= 1e-5 * np.abs(ref).max()


def test_matmul_bf16x3_integers_exact(skl):
    A, B = _ints(11, (96, 160)), _ints(12, (160, 80))
    c = _matmul(skl, A, B, skl.PREC_BF16X3)
    assert np.array_equal(c, (A.astype(np.float64) @ B).astype(np.float32))


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("prec", [1, 2, 3])
def test_matmul_transposes(skl, ta, tb, prec):
    M, N, K = 70, 90, 130
    A = oracle.gen_f32_sym(21, M * K).reshape((K, M) if ta else (M, K))
    B = oracle.gen_f32_sym(22, K * N).reshape((N, K) if tb else (K, N))
    Am = A.T if ta else A
    Bm = B.T if tb else B
    ref = np.ascontiguousarray(Am).astype(np.float64) @ np.ascontiguousarray(Bm).astype(np.float64)
    c = _matmul(skl, np.ascontiguousarray(A), np.ascontiguousarray(B), prec, ta, tb)
    err = np.linalg.norm(c - ref) / np.linalg.norm(ref)
    assert err < (5e-2 if prec == 3 else 1e-6), err


@pytest.mark.parametrize("M,N,K", [(128, 128, 128), (300, 200, 500)])
def test_matmul_fp8_tile_scaled(skl, M, N, K):
    """FP8 e4m3 with one scale per 128x128 tile: normwise error of a few
    percent at most (3 mantissa bits)."""
    A = oracle.gen_f32_sym(0x4001, M * K).reshape(M, K)
    B = oracle.gen_f32_sym(0x4002, K * N).reshape(K, N)
    ref = oracle.allpairs_mm_f64(A, B)
    c = _matmul(skl, A, B, skl.PREC_FP8).astype(np.float64)
    err = np.linalg.norm(c - ref) / np.linalg.norm(ref)
    print(f"fp8 {M}x{N}x{K}: {err:.3e}")
    assert err < 5e-2


def test_allpairs_skeleton_precision_dispatch(skl):
    zip_ = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    red = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    ap = skl.AllPairs("float(float, float)", reduce=red, zip=zip_)
    A = oracle.gen_f32_sym(31, 256 * 512).reshape(256, 512)
    B = oracle.gen_f32_sym(32, 512 * 384).reshape(512, 384)
    ref = oracle.allpairs_mm_f64(A, B)
    for p, tol in ((skl.PREC_FP32, 1e-6), (skl.PREC_BF16X3, 1e-6), (skl.PREC_FP8, 5e-2)):
        ap.set_precision(p)
        out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B))).astype(np.float64)
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol, p


def test_allpairs_addmul_int_inputs_take_the_fold(skl):
    """AllPairs<float(int, int)> with (+, *): the inputs are ints, so the
    float products must not run (they would read int bits as floats); the
    user functions' own fold gives the exact integer sums."""
    zip_ = skl.Zip("float(int, int)", "float func(int x, int y){ return x*y; }")
    red = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    ap = skl.AllPairs("float(int, int)", reduce=red, zip=zip_)
    A = oracle.gen_i32_below(0x4201, 70 * 90, 30).reshape(70, 90).astype(np.int32) - 15
    B = oracle.gen_i32_below(0x4202, 90 * 50, 30).reshape(90, 50).astype(np.int32) - 15
    out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B)))
    assert np.array_equal(out, (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32))


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_allpairs_nonzero_identity_starts_the_fold(skl, prec):
    """A Reduce identity of 1 starts every fold (AllPairsKernel.cl:63), at
    every precision setting: integer-valued inputs make the in-order fold
    exact, so the result is 1 + A.B exactly."""
    zip_ = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    red = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }", "1")
    ap = skl.AllPairs("float(float, float)", reduce=red, zip=zip_)
    ap.set_precision(prec)
    A, B = _ints(41, (96, 130)), _ints(42, (130, 70))
    out = np.asarray(ap(skl.Matrix(A), skl.Matrix(B)))
    assert np.array_equal(out, (1.0 + A.astype(np.float64) @ B).astype(np.float32))


def test_matmul_fp8_tiny_tile_stays_finite(skl):
    """A 128x128 tile of values near 1e-38 (below 448 * 2^-128) gets the
    smallest E8M0 scale; its products flush towards zero, never NaN/inf."""
    M = N = K = 256
    A = oracle.gen_f32_sym(0x4301, M * K).reshape(M, K)
    B = oracle.gen_f32_sym(0x4302, K * N).reshape(K, N)
    A[:128, :128] *= np.float32(1e-38)
    c = _matmul(skl, A, B, skl.PREC_FP8)
    assert np.isfinite(c).all()
    ref = oracle.allpairs_mm_f64(A, B)
    assert np.linalg.norm(c - ref) / np.linalg.norm(ref) < 5e-2

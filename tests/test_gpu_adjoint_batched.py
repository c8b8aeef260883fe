"""Adjoints, batched segmented Reduce/Scan, batched small products, fused
Zip->Reduce under graph capture (extensions beyond the reference, no reference
counterpart).  Adjoints are checked against FP64 central finite differences
and FP64 reference gradients."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_map_backward_vs_fd(skl):
    src = "float func(float x, float a){ return a*x*x + sinf(x); }"
    dsrc = "float dfunc(float x, float a){ return 2.0f*a*x + cosf(x); }"
    m = skl.Map("float(float)", src, deriv_source=dsrc, deriv_func="dfunc")
    x = oracle.gen_f32_sym(1, 4097)
    g = oracle.gen_f32_sym(2, 4097)
    d = np.asarray(m.backward(skl.Vector(x), skl.Vector(g), np.float32(1.5))).astype(np.float64)
    xd = x.astype(np.float64)
    ref = g * (3.0 * xd + np.cos(xd))
    assert np.max(np.abs(d - ref)) < 1e-5
    # FP64 central finite differences of loss = sum(g * f(x)) at a few points
    for i in (0, 100, 4096):
        h = 1e-4
        f = lambda v: 1.5 * v * v + np.sin(v)
        fd = g[i] * (f(xd[i] + h) - f(xd[i] - h)) / (2 * h)
        assert abs(fd - d[i]) < 1e-4


def test_zip_backward(skl):
    src = "float func(float x, float y){ return x*y + x; }"
    dsrc = ("float dl(float x, float y){ return y + 1.0f; }\n"
            "float dr(float x, float y){ return x; }")
    z = skl.Zip("float(float, float)", src, deriv_source=dsrc, deriv_func="dl", deriv_func2="dr")
    x, y, g = (oracle.gen_f32_sym(s, 1000) for s in (3, 4, 5))
    dl, dr = z.backward(skl.Vector(x), skl.Vector(y), skl.Vector(g))
    assert np.allclose(np.asarray(dl), g * (y + 1), atol=1e-6)
    assert np.allclose(np.asarray(dr), g * x, atol=1e-6)


def test_reduce_backward_broadcast(skl):
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    d = np.asarray(r.backward(skl.Vector(np.array([0.75], np.float32)), 12345))
    assert d.shape == (12345,) and np.all(d == np.float32(0.75))


@pytest.mark.parametrize("n", [1, 1000, 4097, 300001, 1 << 20, 65540, 3 * (1 << 20) + 2,
                               5 * (1 << 20) + 3])
@pytest.mark.parametrize("inclusive", [False, True])
def test_scan_backward_reversed_scan(skl, n, inclusive):
    """Reversed scan; lengths with n % 4 != 0 read and write their full tiles
    as dword-aligned packs."""
    s = skl.Scan("float(float)", "float func(float x, float y){ return x+y; }", "0.0f")
    g = oracle.gen_f32_sym(0x5005, n)
    d = np.asarray(s.backward(skl.Vector(g), inclusive=inclusive)).astype(np.float64)
    ref = oracle.scan_add_adjoint_f64(g, inclusive)
    tol = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
    assert np.all(np.abs(d - ref) <= tol)


GAUSS = """float func(input_matrix_t img, __global float* w) {
  float sum = 0.0f;
  for (int dy = -5; dy <= 5; ++dy)
    for (int dx = -5; dx <= 5; ++dx)
      sum += getData(img, dx, dy) * w[(dy + 5) * 11 + (dx + 5)];
  return sum;
}"""


@pytest.mark.parametrize("padding", [0, 1])
def test_mapoverlap_stencil_backward(skl, padding):
    """loss = sum(W * out): d loss / d img is the transposed stencil of W;
    checked against the FP64 adjoint and FP64 central FD on 64 pixels."""
    rows, cols = 200, 333
    w = oracle.gaussian_weights(5, 2.0)
    img = oracle.gen_f32_unit(0x3001, rows * cols, 255.0).reshape(rows, cols)
    W = oracle.gen_f32_sym(0x3002, rows * cols).reshape(rows, cols)
    m = skl.MapOverlap("float(float)", GAUSS, 5, padding, 0)
    wv = skl.Vector(w.ravel())
    d = np.asarray(m.stencil_backward(skl.Matrix(W), wv)).astype(np.float64)
    rng = np.random.default_rng(0x3003)
    pix = np.unique(np.concatenate([rng.integers(0, rows * cols, 60),
                                    [0, cols - 1, (rows - 1) * cols, rows * cols - 1]])
                    ).astype(np.int64)
    ref = oracle.mapoverlap_stencil_adjoint_f64(W, w, 5, pix, padding)
    assert np.max(np.abs(d.ravel()[pix] - ref)) < 1e-5
    # central finite differences of the FP64 loss (linear: exact up to rounding)
    for p in pix[:8]:
        r, c = divmod(int(p), cols)
        win = np.arange(max(0, r - 6) * cols, min(rows, r + 7) * cols, dtype=np.int64)
        h = 0.5
        ip, im = img.astype(np.float64).copy(), img.astype(np.float64).copy()
        ip[r, c] += h
        im[r, c] -= h
        lp = np.dot(W.ravel()[win], oracle.mapoverlap_stencil_f64(ip.astype(np.float32), w, 5, win,
                                                                  padding))
        lm = np.dot(W.ravel()[win], oracle.mapoverlap_stencil_f64(im.astype(np.float32), w, 5, win,
                                                                  padding))
        fd = (lp - lm) / (2 * h)
        assert abs(fd - d[r, c]) < 1e-3 * max(1.0, abs(fd))


def _stencil_src(R):
    W = 2 * R + 1
    return ("float func(input_matrix_t img, __global float* w) { float sum = 0.0f;"
            f" for (int dy = -{R}; dy <= {R}; ++dy) for (int dx = -{R}; dx <= {R}; ++dx)"
            f" sum += getData(img, dx, dy) * w[(dy + {R}) * {W} + (dx + {R})]; return sum; }}")


@pytest.mark.parametrize("R", [1, 3, 5, 6])
@pytest.mark.parametrize("shape", [(37, 70), (5, 9), (130, 64)])
@pytest.mark.parametrize("padding", [0, 1])
def test_mapoverlap_stencil_backward_radii(skl, R, shape, padding):
    """Every pixel of the transposed stencil against the FP64 adjoint: radii
    on the register-tiled form (1..5) and the generic one (6), NEUTRAL and
    NEAREST, matrices smaller than 2R (overlapping border bands) and ones
    spanning several 64 x 64 tiles."""
    rows, cols = shape
    w = oracle.gaussian_weights(R, 1.5)
    W = oracle.gen_f32_sym(0x3010 + R, rows * cols).reshape(rows, cols)
    m = skl.MapOverlap("float(float)", _stencil_src(R), R, padding, 0)
    d = np.asarray(m.stencil_backward(skl.Matrix(W), skl.Vector(w.ravel()))).astype(np.float64)
    pix = np.arange(rows * cols, dtype=np.int64)
    ref = oracle.mapoverlap_stencil_adjoint_f64(W, w, R, pix, padding)
    assert np.max(np.abs(d.ravel() - ref)) < 1e-5 * max(1.0, np.abs(ref).max())


def test_allpairs_backward(skl):
    zip_ = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    red = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    ap = skl.AllPairs("float(float, float)", reduce=red, zip=zip_)
    M, K, N = 96, 130, 70
    A = oracle.gen_f32_sym(0x4001, M * K).reshape(M, K)
    B = oracle.gen_f32_sym(0x4002, K * N).reshape(K, N)
    dC = oracle.gen_f32_sym(0x4004, M * N).reshape(M, N)
    for prec in (skl.PREC_FP32, skl.PREC_BF16X3):
        ap.set_precision(prec)
        dA, dB = ap.backward(skl.Matrix(A), skl.Matrix(B), skl.Matrix(dC))
        rA = dC.astype(np.float64) @ B.T.astype(np.float64)
        rB = A.T.astype(np.float64) @ dC.astype(np.float64)
        assert np.linalg.norm(np.asarray(dA) - rA) / np.linalg.norm(rA) < 1e-6
        assert np.linalg.norm(np.asarray(dB) - rB) / np.linalg.norm(rB) < 1e-6


# ------------------------------------------------------------- batched
def _segments(n_seg, seed, lo=3.0, span=3.0):
    lens = oracle.gen_loguniform(seed, n_seg, lo, span)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    return lens, off


def test_segmented_reduce_small(skl):
    lens = np.array([0, 1, 5, 4096, 4097, 70000, 0, 3], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = oracle.gen_f32_sym(0x5002, int(off[-1]))
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    got = np.asarray(r.segmented(skl.Vector(x), off)).astype(np.float64)
    ref = oracle.segmented_sum_f64(x, off)
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref).astype(np.float32)) + 1e-12)


@pytest.mark.parametrize("prec", ["DEFAULT", "FP64"])
def test_segmented_reduce_scan_cfg5_subset(skl, prec):
    """The batched configuration's segment lengths (log-uniform [1e3, 1e6),
    committed fixture), first 256 segments; compensated FP32 inside chunks
    and tiles (DEFAULT) and FP64 throughout (FP64)."""
    with open(os.path.join(os.path.dirname(__file__), "golden", "batched_shapes.json")) as f:
        fx = json.load(f)
    lens = np.array(fx["segment_lengths"][:256], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = oracle.gen_f32_sym(0x5002, int(off[-1]))
    X = skl.Vector(x)
    P = getattr(skl, "PREC_" + prec)
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }", precision=P)
    got = np.asarray(r.segmented(X, off)).astype(np.float64)
    ref = oracle.segmented_sum_f64(x, off)
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref).astype(np.float32)) + 1e-12)
    s = skl.Scan("float(float)", "float func(float x, float y){ return x+y; }", "0.0f",
                 precision=P)
    for incl in (False, True):
        got = np.asarray(s.segmented(X, off, inclusive=incl)).astype(np.float64)
        ref = oracle.segmented_scan_f64(x, off, incl)
        tol = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
        assert np.all(np.abs(got - ref) <= tol), incl


def test_segmented_long_segment(skl):
    """A segment of 5e6 elements (1221 scan chunks, 306 reduce chunks) between
    short and empty ones: multi-window look-back inside one segment."""
    lens = np.array([3, 5_000_000, 0, 4096, 1], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = oracle.gen_f32_sym(0x5006, int(off[-1]))
    X = skl.Vector(x)
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    got = np.asarray(r.segmented(X, off)).astype(np.float64)
    ref = oracle.segmented_sum_f64(x, off)
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref).astype(np.float32)) + 1e-12)
    s = skl.Scan("float(float)", "float func(float x, float y){ return x+y; }", "0.0f")
    for incl in (False, True):
        got = np.asarray(s.segmented(X, off, inclusive=incl)).astype(np.float64)
        ref = oracle.segmented_scan_f64(x, off, incl)
        tol = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
        assert np.all(np.abs(got - ref) <= tol), incl


def test_segmented_invalid_offsets_flagged_on_device(skl):
    """Offsets checked by the prologue kernel (the host check is bypassed by
    passing device offsets): Reduce writes NaN, Scan writes nothing."""
    from skelcl_amd import _lib
    n = 64
    x = oracle.gen_f32_sym(0x5007, n)
    X = skl.DeviceBuffer.from_array(x)
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    s = skl.Scan("float(float)", "float func(float x, float y){ return x+y; }", "0.0f")
    for bad in ([0, 40, 20, 64], [0, 10, 65], [-1, 10, 20]):
        off = skl.DeviceBuffer.from_array(np.array(bad, np.int64))
        nseg = len(bad) - 1
        o = skl.DeviceBuffer.from_array(np.full(nseg, 7.0, np.float32))
        _lib.check(_lib.lib().skl_reduce_segmented_run(r.handle, o.ptr, X.ptr, n, off.ptr, nseg,
                                                       None))
        assert np.all(np.isnan(o.to_numpy(np.float32, nseg))), bad
        so = skl.DeviceBuffer.from_array(np.full(n, 7.0, np.float32))
        _lib.check(_lib.lib().skl_scan_segmented_run(s.handle, so.ptr, X.ptr, n, off.ptr, nseg,
                                                     1, None))
        assert np.all(so.to_numpy(np.float32, n) == 7.0), bad
    # a valid call on the same objects afterwards is unaffected
    off = np.array([0, 30, 64], np.int64)
    got = np.asarray(r.segmented(skl.Vector(x), off)).astype(np.float64)
    ref = oracle.segmented_sum_f64(x, off)
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref).astype(np.float32)) + 1e-12)
    for incl in (False, True):
        got = np.asarray(s.segmented(skl.Vector(x), off, inclusive=incl)).astype(np.float64)
        ref = oracle.segmented_scan_f64(x, off, incl)
        assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref).astype(np.float32)) + 1e-9)


@pytest.mark.parametrize("case", ["wrap_long", "dense_tiny", "inner_range"])
def test_segmented_int_exact(skl, case):
    """int Reduce(+) / Scan(+) segmented (exact modulo 2^32) against the
    oracle, bit for bit: values whose sums wrap over segments of several
    Reduce chunks (16384) and Scan tiles (32768), runs of empty and
    one-element segments, and an uncovered head and tail (left as they
    are); invalid offsets make Reduce write INT_MIN."""
    from skelcl_amd import _lib
    rng = np.random.default_rng({"wrap_long": 21, "dense_tiny": 22, "inner_range": 23}[case])
    if case == "wrap_long":
        off = np.array([0, 3, 3, 300000, 300001, 1_000_000], np.int64)
        n = 1_000_000
        x = oracle.gen_u64(0x5201, n).astype(np.uint32).view(np.int32)
    elif case == "dense_tiny":
        lens = rng.integers(0, 4, 30000)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        n = int(off[-1]) + 11
        x = oracle.gen_i32_below(0x5202, n, 1000)
    else:
        off = np.array([1000, 1000, 20000, 20001, 60000, 100005], np.int64)
        n = 100100
        x = oracle.gen_i32_below(0x5203, n, 100000)
    nseg = len(off) - 1
    r = skl.Reduce("int(int)", "int func(int x, int y){ return x+y; }")
    s = skl.Scan("int(int)", "int func(int x, int y){ return x+y; }", "0")
    X = skl.DeviceBuffer.from_array(x)
    O = skl.DeviceBuffer.from_array(off)
    o = skl.DeviceBuffer.from_array(np.full(nseg, 0x5EED, np.int32))
    _lib.check(_lib.lib().skl_reduce_segmented_run(r.handle, o.ptr, X.ptr, n, O.ptr, nseg, None))
    assert np.array_equal(o.to_numpy(np.uint32, nseg), oracle.segmented_sum_u32(x, off))
    for incl in (0, 1):
        so = skl.DeviceBuffer.from_array(np.full(n, 0x5EED, np.int32))
        _lib.check(_lib.lib().skl_scan_segmented_run(s.handle, so.ptr, X.ptr, n, O.ptr, nseg,
                                                     incl, None))
        assert np.array_equal(so.to_numpy(np.uint32, n),
                              oracle.segmented_scan_u32(x, off, bool(incl), fill=0x5EED)), incl
    # the Python mirror over the same skeletons
    got = np.asarray(r.segmented(skl.Vector(x), off)).astype(np.int32).view(np.uint32)
    assert np.array_equal(got, oracle.segmented_sum_u32(x, off))
    bad = skl.DeviceBuffer.from_array(np.array([0, 40, 20, n], np.int64))
    o = skl.DeviceBuffer.from_array(np.full(3, 7, np.int32))
    _lib.check(_lib.lib().skl_reduce_segmented_run(r.handle, o.ptr, X.ptr, n, bad.ptr, 3, None))
    assert np.all(o.to_numpy(np.int32, 3) == np.iinfo(np.int32).min)


def _seg_scan_raw(skl, x, off, n, inclusive, fill=7.0):
    from skelcl_amd import _lib
    X = skl.DeviceBuffer.from_array(x)
    o = skl.DeviceBuffer.from_array(np.array(off, np.int64))
    y = skl.DeviceBuffer.from_array(np.full(n, fill, np.float32))
    s = skl.Scan("float(float)", "float func(float x, float y){ return x+y; }", "0.0f")
    _lib.check(_lib.lib().skl_scan_segmented_run(s.handle, y.ptr, X.ptr, n, o.ptr, len(off) - 1,
                                                 int(inclusive), None))
    return y.to_numpy(np.float32, n)


SEG_TILE = 32768  # elements per tile of the segmented Scan


@pytest.mark.parametrize("case", ["inner_range", "dense_tiny", "tile_edges", "ragged_tail"])
@pytest.mark.parametrize("inclusive", [False, True])
def test_segmented_scan_layouts(skl, case, inclusive):
    """Segment layouts against tile boundaries: segments that leave the ends
    of the array uncovered (untouched outputs), runs of empty and one-element
    segments (many heads per tile), heads exactly on, before and after tile
    starts, and a last tile cut short at a length that is not a multiple of 4."""
    rng = np.random.default_rng({"inner_range": 1, "dense_tiny": 2, "tile_edges": 3,
                                 "ragged_tail": 4}[case])
    if case == "inner_range":
        n = 5 * SEG_TILE + 123
        off = [1000, 1000, 20000, 20001, 60000, 4 * SEG_TILE + 5]
    elif case == "dense_tiny":
        lens = rng.integers(0, 4, 30000)
        off = np.concatenate([[0], np.cumsum(lens)]).tolist()
        n = off[-1] + 50
    elif case == "tile_edges":
        T = SEG_TILE
        off = [0, T - 1, T, T + 1, 2 * T, 3 * T, 3 * T, 5 * T - 1, 7 * T, 7 * T + 3]
        n = off[-1]
    else:
        n = 9 * SEG_TILE + 7
        lens = rng.integers(1, 40000, 12)
        off = np.concatenate([[0], np.minimum(np.cumsum(lens), n)]).tolist()
        off[-1] = n
    x = oracle.gen_f32_sym(0x5010, n)
    got = _seg_scan_raw(skl, x, off, n, inclusive)
    ref = oracle.segmented_scan_f64(x, np.array(off, np.int64), inclusive)
    inside = np.zeros(n, bool)
    inside[off[0]:off[-1]] = True
    tol = np.spacing(np.abs(ref[inside]).astype(np.float32)).astype(np.float64) + 1e-9
    assert np.all(np.abs(got[inside].astype(np.float64) - ref[inside]) <= tol)
    assert np.all(got[~inside] == 7.0)


SUM_CHUNK = 16384  # elements per chunk of the segmented Reduce


@pytest.mark.parametrize("case", ["dense_tiny", "chunk_multiples"])
def test_segmented_reduce_layouts(skl, case):
    """Runs of empty and one-element segments, and segments whose lengths are
    exact multiples of the chunk (full chunks only) or one off, starting at
    offsets that are not multiples of 4 (the full-chunk path reads
    dword-aligned packs)."""
    rng = np.random.default_rng(7 if case == "dense_tiny" else 8)
    if case == "dense_tiny":
        lens = rng.integers(0, 4, 30000)
    else:
        C = SUM_CHUNK
        lens = np.array([1, C, 2 * C, C - 1, C + 1, 0, 3 * C, 5, 7 * C + 2], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = oracle.gen_f32_sym(0x5040, int(off[-1]))
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    got = np.asarray(r.segmented(skl.Vector(x), off)).astype(np.float64)
    ref = oracle.segmented_sum_f64(x, off)
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref).astype(np.float32)) + 1e-12)


def test_segmented_scan_repeated_and_resized(skl):
    """The tile ticket and status words are reused across calls: alternate
    sizes on one skeleton, every result checked."""
    from skelcl_amd import _lib
    s = skl.Scan("float(float)", "float func(float x, float y){ return x+y; }", "0.0f")
    for it, n in enumerate([300000, 40000, 300000, 17, 1 << 20]):
        lens = oracle.gen_loguniform(0x5020 + it, 40, 1.0, 4.0)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        off = np.minimum(off, n)
        off[-1] = n
        x = oracle.gen_f32_sym(0x5030 + it, n)
        X = skl.DeviceBuffer.from_array(x)
        O = skl.DeviceBuffer.from_array(off)
        y = skl.DeviceBuffer.from_array(np.zeros(n, np.float32))
        _lib.check(_lib.lib().skl_scan_segmented_run(s.handle, y.ptr, X.ptr, n, O.ptr,
                                                     len(off) - 1, 1, None))
        got = y.to_numpy(np.float32, n).astype(np.float64)
        ref = oracle.segmented_scan_f64(x, off, True)
        tol = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
        assert np.all(np.abs(got - ref) <= tol), n


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_matmul(skl, prec):
    from skelcl_amd import _lib
    with open(os.path.join(os.path.dirname(__file__), "golden", "batched_shapes.json")) as f:
        shapes = json.load(f)["gemm_shapes"][:40]
    descs = (_lib.skl_gemm_desc * len(shapes))()
    keep, refs = [], []
    for i, (M, N, K) in enumerate(shapes):
        A = oracle.gen_f32_sym(0x5004 + 3 * i, M * K).reshape(M, K)
        B = oracle.gen_f32_sym(0x5004 + 3 * i + 1, K * N).reshape(K, N)
        dA, dB = skl.DeviceBuffer.from_array(A), skl.DeviceBuffer.from_array(B)
        dC = skl.DeviceBuffer(4 * M * N)
        keep += [dA, dB, dC]
        descs[i].A, descs[i].B, descs[i].C = dA.value, dB.value, dC.value
        descs[i].M, descs[i].N, descs[i].K = M, N, K
        refs.append((dC, M, N, A.astype(np.float64) @ B.astype(np.float64)))
    dd = skl.DeviceBuffer(C.sizeof(descs))
    _lib.check(_lib.lib().skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs),
                                         None, 1))
    _lib.check(_lib.lib().skl_matmul_batched_f32(dd.ptr, len(shapes), prec, None))
    for dC, M, N, ref in refs:
        c = dC.to_numpy(np.float32, M * N).reshape(M, N)
        assert np.linalg.norm(c - ref) / np.linalg.norm(ref) < 1e-6


def test_batched_pipeline_graph_replays(skl):
    """The batched pipeline (segmented Reduce, segmented Scan, batched BF16X3
    products: nine launches) captured in one graph after an eager warm-up:
    every replay, with the outputs poisoned before it and eager work of the
    same skeletons on other data between replays, reproduces the eager
    results bit for bit (the tile tickets and look-back epochs re-arm inside
    the kernels; no memset node is recorded)."""
    from skelcl_amd import _lib
    L = _lib.lib()
    with open(os.path.join(os.path.dirname(__file__), "golden", "batched_shapes.json")) as f:
        fx = json.load(f)
    lens = np.array(fx["segment_lengths"][:24], np.int64) // 16
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n, nseg = int(off[-1]), len(lens)
    x = oracle.gen_f32_sym(0x5050, n)
    X, O = skl.DeviceBuffer.from_array(x), skl.DeviceBuffer.from_array(off)
    x2 = oracle.gen_f32_sym(0x5051, n)
    X2 = skl.DeviceBuffer.from_array(x2)
    sums, ys = skl.DeviceBuffer(4 * nseg), skl.DeviceBuffer(4 * n)
    sums2, ys2 = skl.DeviceBuffer(4 * nseg), skl.DeviceBuffer(4 * n)
    shapes = fx["gemm_shapes"][:16]
    descs = (_lib.skl_gemm_desc * len(shapes))()
    keep, outs = [], []
    for i, (M, N, K) in enumerate(shapes):
        dA = skl.DeviceBuffer.from_array(oracle.gen_f32_sym(0x5060 + 2 * i, M * K))
        dB = skl.DeviceBuffer.from_array(oracle.gen_f32_sym(0x5061 + 2 * i, K * N))
        dC = skl.DeviceBuffer(4 * M * N)
        keep += [dA, dB, dC]
        descs[i].A, descs[i].B, descs[i].C = dA.value, dB.value, dC.value
        descs[i].M, descs[i].N, descs[i].K = M, N, K
        outs.append((dC, M * N))
    dd = skl.DeviceBuffer(C.sizeof(descs))
    _lib.check(L.skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs), None, 1))
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }", "0")
    s = skl.Scan("float(float)", "float func(float x, float y){ return x+y; }", "0.0f")
    st = skl.Stream()

    def pipeline(xin, so, yo):
        _lib.check(L.skl_reduce_segmented_run(r.handle, so.ptr, xin.ptr, n, O.ptr, nseg,
                                              st.handle))
        _lib.check(L.skl_scan_segmented_run(s.handle, yo.ptr, xin.ptr, n, O.ptr, nseg, 1,
                                            st.handle))
        _lib.check(L.skl_matmul_batched_f32(dd.ptr, len(shapes), 2, st.handle))

    pipeline(X, sums, ys)  # eager: allocates the scratch the graph then uses
    st.synchronize()
    ref_sums = sums.to_numpy(np.float32, nseg)
    ref_ys = ys.to_numpy(np.float32, n)
    ref_cs = [c.to_numpy(np.float32, m) for c, m in outs]
    g = C.c_void_p()
    _lib.check(L.skl_graph_begin_capture(st.handle))
    pipeline(X, sums, ys)
    _lib.check(L.skl_graph_end_capture(st.handle, C.byref(g)))
    for k in range(4):
        print(f"replay {k}", flush=True)
        sums.fill(b"\x00\x00\xc0\x7f", 4, nseg, st)
        ys.fill(b"\x00\x00\xc0\x7f", 4, n, st)
        for c, m in outs:
            c.fill(b"\x00\x00\xc0\x7f", 4, m, st)
        _lib.check(L.skl_graph_launch(g, st.handle))
        st.synchronize()
        assert np.array_equal(sums.to_numpy(np.float32, nseg).view(np.uint32),
                              ref_sums.view(np.uint32)), k
        assert np.array_equal(ys.to_numpy(np.float32, n).view(np.uint32),
                              ref_ys.view(np.uint32)), k
        for (c, m), want in zip(outs, ref_cs):
            assert np.array_equal(c.to_numpy(np.float32, m).view(np.uint32),
                                  want.view(np.uint32)), k
        pipeline(X2, sums2, ys2)  # eager work of the same skeletons between replays
        st.synchronize()
    _lib.check(L.skl_graph_destroy(g))
    got2 = sums2.to_numpy(np.float32, nseg).astype(np.float64)
    ref2 = oracle.segmented_sum_f64(x2, off)
    assert np.all(np.abs(got2 - ref2) <= np.spacing(np.abs(ref2).astype(np.float32)) + 1e-12)


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_matmul_ragged_shapes(skl, prec):
    """Partial 128x128 tiles, K not a multiple of the 32-deep slice, and
    degenerate 1x1x1 products in one batch."""
    from skelcl_amd import _lib
    shapes = [(1, 1, 1), (3, 130, 7), (200, 1, 65), (129, 257, 33), (64, 64, 1), (17, 300, 95)]
    descs = (_lib.skl_gemm_desc * len(shapes))()
    keep, refs = [], []
    for i, (M, N, K) in enumerate(shapes):
        A = oracle.gen_f32_sym(0x5100 + 2 * i, M * K).reshape(M, K)
        B = oracle.gen_f32_sym(0x5101 + 2 * i, K * N).reshape(K, N)
        dA, dB = skl.DeviceBuffer.from_array(A), skl.DeviceBuffer.from_array(B)
        dC = skl.DeviceBuffer(4 * M * N)
        keep += [dA, dB, dC]
        descs[i].A, descs[i].B, descs[i].C = dA.value, dB.value, dC.value
        descs[i].M, descs[i].N, descs[i].K = M, N, K
        refs.append((dC, M, N, A.astype(np.float64) @ B.astype(np.float64)))
    dd = skl.DeviceBuffer(C.sizeof(descs))
    _lib.check(_lib.lib().skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs),
                                         None, 1))
    _lib.check(_lib.lib().skl_matmul_batched_f32(dd.ptr, len(shapes), prec, None))
    for dC, M, N, ref in refs:
        c = dC.to_numpy(np.float32, M * N).reshape(M, N)
        assert np.linalg.norm(c - ref) / max(np.linalg.norm(ref), 1e-30) < 1e-6, (M, N)


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_matmul_mixed_classes(skl, prec):
    """One batch holding products with K and N multiples of 4 (16-byte
    staging, partial 128x128 tiles included) and products without (element
    staging): each class is numbered and run by its own kernel."""
    from skelcl_amd import _lib
    shapes = [(129, 260, 36), (3, 130, 7), (1, 4, 4), (200, 8, 64), (64, 64, 1),
              (130, 132, 100), (0, 8, 8), (257, 12, 4)]
    descs = (_lib.skl_gemm_desc * len(shapes))()
    keep, refs = [], []
    for i, (M, N, K) in enumerate(shapes):
        A = oracle.gen_f32_sym(0x5200 + 2 * i, max(M * K, 1)).reshape(-1)[:M * K].reshape(M, K)
        B = oracle.gen_f32_sym(0x5201 + 2 * i, K * N).reshape(K, N)
        dA = skl.DeviceBuffer.from_array(A) if M * K else skl.DeviceBuffer(16)
        dB = skl.DeviceBuffer.from_array(B)
        dC = skl.DeviceBuffer(max(4 * M * N, 16))
        keep += [dA, dB, dC]
        descs[i].A, descs[i].B, descs[i].C = dA.value, dB.value, dC.value
        descs[i].M, descs[i].N, descs[i].K = M, N, K
        if M:
            refs.append((dC, M, N, A.astype(np.float64) @ B.astype(np.float64)))
    dd = skl.DeviceBuffer(C.sizeof(descs))
    _lib.check(_lib.lib().skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs),
                                         None, 1))
    _lib.check(_lib.lib().skl_matmul_batched_f32(dd.ptr, len(shapes), prec, None))
    for dC, M, N, ref in refs:
        c = dC.to_numpy(np.float32, M * N).reshape(M, N)
        assert np.linalg.norm(c - ref) / max(np.linalg.norm(ref), 1e-30) < 1e-6, (M, N)


def test_graph_capture_zip_reduce(skl):
    """Zip->Reduce fused, captured into a graph and replayed: every replay
    equals the eager result (the reduce counter re-arms itself, no memset
    nodes are recorded)."""
    from skelcl_amd import _lib
    L = _lib.lib()
    n = (1 << 20) + 3
    x, y = oracle.gen_f32_sym(0x1001, n), oracle.gen_f32_sym(0x1002, n)
    z = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    X, Y = skl.Vector(x), skl.Vector(y)
    st = skl.Stream()
    out = skl.DeviceBuffer(4)
    # eager (also allocates the reduce scratch outside the capture)
    _lib.check(L.skl_zip_reduce_run(z.handle, r.handle, out.ptr, X.device_ptr(), Y.device_ptr(),
                                    n, st.handle))
    st.synchronize()
    eager = out.to_numpy(np.float32, 1)[0]
    g = C.c_void_p()
    _lib.check(L.skl_graph_begin_capture(st.handle))
    _lib.check(L.skl_zip_reduce_run(z.handle, r.handle, out.ptr, X.device_ptr(), Y.device_ptr(),
                                    n, st.handle))
    _lib.check(L.skl_graph_end_capture(st.handle, C.byref(g)))
    for k in range(5):
        out.fill(b"\x00\x00\xc0\x7f", 4, 1, st)  # NaN between replays
        _lib.check(L.skl_graph_launch(g, st.handle))
        st.synchronize()
        assert out.to_numpy(np.float32, 1)[0] == eager, k
        # eager work between replays
        _lib.check(L.skl_zip_reduce_run(z.handle, r.handle, out.ptr, X.device_ptr(),
                                        Y.device_ptr(), n, st.handle))
        st.synchronize()
    _lib.check(L.skl_graph_destroy(g))
    ref = oracle.dot_f64(x, y)
    assert abs(float(eager) - ref) <= 2 * np.spacing(np.float32(abs(ref))) + 1e-12 * n


def test_autotune_keeps_results_and_caches(skl, tmp_path, monkeypatch):
    """The autotuner changes only launch shapes: results stay bit-identical,
    and the chosen shape is cached and reused."""
    import ctypes as C
    from skelcl_amd import _lib
    monkeypatch.setenv("SKELCL_CACHE_DIR", str(tmp_path))
    L = _lib.lib()
    s = skl.Scan("int(int)", "int func(int x, int y){ return x+y; }")
    x = oracle.gen_i32_below(77, 1 << 20, 1000)
    before = np.asarray(s(skl.Vector(x), inclusive=True))
    _lib.check(L.skl_autotune(s.handle, 1 << 20, 1))
    shape = (C.c_int * 4)()
    _lib.check(L.skl_skeleton_launch_shape(s.handle, shape))
    assert shape[2] in (2, 4, 8, 16)
    after = np.asarray(s(skl.Vector(x), inclusive=True))
    assert np.array_equal(before, after)
    m = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    _lib.check(L.skl_autotune(m.handle, 1 << 22, 1))
    a, b = oracle.gen_f32_sym(1, 12345), oracle.gen_f32_sym(2, 12345)
    assert np.array_equal(np.asarray(m(skl.Vector(a), skl.Vector(b))), a * b)
    _lib.check(L.skl_autotune(m.handle, 0, 0))  # back to defaults


@pytest.mark.parametrize("prec", ["BF16X3", "FP8"])
def test_mma_tiles_agree_and_autotune_picks_one(skl, prec, tmp_path, monkeypatch):
    """Every MFMA work-group tile of the AllPairs (+,*) product gives the
    same bits (each output tile runs the same MFMAs in the same k order), on
    ragged shapes where the size rule would not pick the larger tiles; the
    autotuner keeps the size rule or picks one of them, and the result stays
    the same."""
    import ctypes as C
    from skelcl_amd import _lib
    monkeypatch.setenv("SKELCL_CACHE_DIR", str(tmp_path))
    L = _lib.lib()
    P = getattr(skl, "PREC_" + prec)
    add = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }", "0")
    mul = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    ap = skl.AllPairs("float(float, float)", reduce=add, zip=mul, precision=P)
    for (M, N, K) in [(300, 517, 260), (512, 512, 512)]:
        A = skl.Matrix(oracle.gen_f32_sym(0x7101, M * K).reshape(M, K))
        B = skl.Matrix(oracle.gen_f32_sym(0x7102, K * N).reshape(K, N))
        outs = []
        for tile in (0, 1, 2, 3):
            _lib.check(L.skl_skeleton_set_mma_tile(ap.handle, tile))
            print(f"{prec} {M}x{N}x{K} tile {tile}", flush=True)
            outs.append(np.asarray(ap(A, B)).view(np.uint32).copy())
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])
    _lib.check(L.skl_skeleton_set_mma_tile(ap.handle, 0))
    _lib.check(L.skl_autotune(ap.handle, 512 * 512, 1))
    t = C.c_int(-1)
    _lib.check(L.skl_skeleton_mma_tile(ap.handle, C.byref(t)))
    # 0: the size rule stayed (no tile was clearly faster)
    assert t.value in ((0, 1, 2) if prec == "BF16X3" else (0, 1, 2, 3))
    assert np.array_equal(np.asarray(ap(A, B)).view(np.uint32), outs[0])


def test_fused_reduce_follows_a_new_zip(skl):
    """One Reduce fused with a Zip, the Zip released, then a Zip with another
    function (possibly at the released one's address): the fused program
    belongs to the Zip's function, so each result matches its own function."""
    import gc
    n = 100003
    x, y = oracle.gen_f32_sym(0x1001, n), oracle.gen_f32_sym(0x1002, n)
    X, Y = skl.Vector(x), skl.Vector(y)
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    exact = {"x*y": (x.astype(np.float64) * y).sum(), "x-y": (x.astype(np.float64) - y).sum(),
             "x+2*y": (x.astype(np.float64) + 2 * y.astype(np.float64)).sum()}
    for _ in range(2):
        for body, want in exact.items():
            z = skl.Zip("float(float, float)", f"float func(float x, float y){{ return {body}; }}")
            got = float(r.fused(z, X, Y)[0])
            del z
            gc.collect()
            assert abs(got - want) <= 1e-5 * max(1.0, np.abs(x).sum()), body


# ------------------------------------------- adjoint rigour (FD + guards)
def test_zip_backward_vs_fd(skl):
    """A non-linear Zip: the GPU adjoint against the FP64 reference gradient
    and FP64 central finite differences of loss = sum(g * f(x, y))."""
    src = "float func(float x, float y){ return x*y + sinf(x)*y*y; }"
    dsrc = ("float dl(float x, float y){ return y + cosf(x)*y*y; }\n"
            "float dr(float x, float y){ return x + 2.0f*sinf(x)*y; }")
    z = skl.Zip("float(float, float)", src, deriv_source=dsrc, deriv_func="dl", deriv_func2="dr")
    n = 5003
    x, y, g = (oracle.gen_f32_sym(s, n).astype(np.float32) for s in (0x6001, 0x6002, 0x6003))
    dl, dr = (np.asarray(v).astype(np.float64) for v in z.backward(skl.Vector(x), skl.Vector(y),
                                                                      skl.Vector(g)))
    xd, yd, gd = x.astype(np.float64), y.astype(np.float64), g.astype(np.float64)
    assert np.max(np.abs(dl - gd * (yd + np.cos(xd) * yd * yd))) < 1e-5
    assert np.max(np.abs(dr - gd * (xd + 2 * np.sin(xd) * yd))) < 1e-5
    f = lambda a, b: a * b + np.sin(a) * b * b
    h = 1e-5
    for i in (0, 17, 2500, n - 1):
        fdl = gd[i] * (f(xd[i] + h, yd[i]) - f(xd[i] - h, yd[i])) / (2 * h)
        fdr = gd[i] * (f(xd[i], yd[i] + h) - f(xd[i], yd[i] - h)) / (2 * h)
        assert abs(fdl - dl[i]) < 1e-5 * max(1.0, abs(fdl)), i
        assert abs(fdr - dr[i]) < 1e-5 * max(1.0, abs(fdr)), i


def test_reduce_backward_vs_fd_of_the_gpu_forward(skl):
    """Reduce(+) is linear, so central differences of the GPU forward itself
    with h = 1 give the gradient up to FP32 rounding of the sums: every
    sampled d loss / d x_i equals the broadcast upstream value."""
    n = 100003
    x = oracle.gen_f32_sym(0x6004, n)
    up = np.float32(0.75)
    r = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    d = np.asarray(r.backward(skl.Vector(np.array([up], np.float32)), n))
    loss = lambda v: float(up) * float(r(skl.Vector(v))[0])
    for i in (0, 1, 4099, n - 1):
        xp, xm = x.copy(), x.copy()
        xp[i] += np.float32(1.0)
        xm[i] -= np.float32(1.0)
        fd = (loss(xp) - loss(xm)) / 2.0
        assert abs(fd - float(d[i])) < 2e-5 * max(1.0, abs(fd)), (i, fd, d[i])


def test_allpairs_backward_vs_fd_of_the_gpu_forward(skl):
    """The (+,*) product is bilinear: central differences of the GPU forward
    of loss = sum(dC * (A.B)) with h = 1 at sampled entries of A and B match
    dA = dC.B^T and dB = A^T.dC up to FP32 rounding."""
    zip_ = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    red = skl.Reduce("float(float)", "float func(float x, float y){ return x+y; }")
    ap = skl.AllPairs("float(float, float)", reduce=red, zip=zip_)
    M, K, N = 70, 130, 90
    A = oracle.gen_f32_sym(0x6005, M * K).reshape(M, K)
    B = oracle.gen_f32_sym(0x6006, K * N).reshape(K, N)
    dC = oracle.gen_f32_sym(0x6007, M * N).reshape(M, N)
    dA, dB = (np.asarray(v).astype(np.float64)
              for v in ap.backward(skl.Matrix(A), skl.Matrix(B), skl.Matrix(dC)))
    dC64 = dC.astype(np.float64)
    loss = lambda a, b: float((dC64 * np.asarray(ap(skl.Matrix(a), skl.Matrix(b)))).sum())
    for (i, k) in ((0, 0), (69, 129), (33, 7)):
        ap_, am = A.copy(), A.copy()
        ap_[i, k] += 1.0
        am[i, k] -= 1.0
        fd = (loss(ap_, B) - loss(am, B)) / 2.0
        assert abs(fd - dA[i, k]) < 1e-4 * max(1.0, abs(fd)), (i, k, fd, dA[i, k])
    for (k, j) in ((0, 0), (129, 89), (64, 45)):
        bp, bm = B.copy(), B.copy()
        bp[k, j] += 1.0
        bm[k, j] -= 1.0
        fd = (loss(A, bp) - loss(A, bm)) / 2.0
        assert abs(fd - dB[k, j]) < 1e-4 * max(1.0, abs(fd)), (k, j, fd, dB[k, j])


@pytest.mark.parametrize("src,why", [
    # weights indexed transposed: not the stencil of w (w is asymmetric)
    ("float func(input_matrix_t img, __global float* w) { float s = 0.0f;"
     " for (int dy = -2; dy <= 2; ++dy) for (int dx = -2; dx <= 2; ++dx)"
     " s += getData(img, dx, dy) * w[(dx + 2) * 5 + (dy + 2)]; return s; }", "transposed"),
    # a non-linear term
    ("float func(input_matrix_t img, __global float* w) { float s = 0.0f;"
     " for (int dy = -2; dy <= 2; ++dy) for (int dx = -2; dx <= 2; ++dx)"
     " { float v = getData(img, dx, dy); s += v * v * w[(dy + 2) * 5 + (dx + 2)]; } return s; }",
     "square"),
    # a term left out
    ("float func(input_matrix_t img, __global float* w) { float s = 0.0f;"
     " for (int dy = -2; dy <= 2; ++dy) for (int dx = -2; dx < 2; ++dx)"
     " s += getData(img, dx, dy) * w[(dy + 2) * 5 + (dx + 2)]; return s; }", "dropped"),
    # another extra argument
    ("float func(input_matrix_t img, __global float* w, float k) {"
     " return k * getData(img, 0, 0) * w[12]; }", "extra"),
])
def test_stencil_backward_rejects_other_user_functions(skl, src, why):
    from skelcl_amd import _lib
    m = skl.MapOverlap("float(float)", src, 2, 0, 0)
    W = oracle.gen_f32_sym(0x6010, 30 * 40).reshape(30, 40)
    w = oracle.gen_f32_sym(0x6011, 25)
    with pytest.raises(_lib.SkelError) as e:
        m.stencil_backward(skl.Matrix(W), skl.Vector(w))
    assert e.value.status == _lib.SKL_ERR_UNSUPPORTED, why
    assert "not the weighted stencil" in str(e.value)


@pytest.mark.parametrize("padding", [0, 1])
def test_stencil_backward_accepts_the_weighted_stencil(skl, padding):
    """The same form with the weight index written another way is accepted
    (the probe checks values, not spelling)."""
    src = ("float func(input_matrix_t img, __global float* w) { float s = 0.0f;"
           " for (int i = 0; i < 25; ++i) s += w[i] * getData(img, i % 5 - 2, i / 5 - 2);"
           " return s; }")
    m = skl.MapOverlap("float(float)", src, 2, padding, 0)
    rows, cols = 30, 40
    W = oracle.gen_f32_sym(0x6012, rows * cols).reshape(rows, cols)
    w = oracle.gen_f32_sym(0x6013, 25).reshape(5, 5)
    d = np.asarray(m.stencil_backward(skl.Matrix(W), skl.Vector(w.ravel()))).astype(np.float64)
    pix = np.arange(rows * cols, dtype=np.int64)
    ref = oracle.mapoverlap_stencil_adjoint_f64(W, w, 2, pix, padding)
    assert np.max(np.abs(d.ravel() - ref)) < 1e-5 * max(1.0, np.abs(ref).max())

"""Seeded random shapes for the extensions whose layouts have the most
cases: the one-launch batch (segment layouts with empty and uncovered
parts, with and without products), the two-launch segmented Reduce/Scan
(float and int32), the native weighted-stencil forward and adjoint at
random sizes, radii, paddings and rows per thread, dense products in
the three precision forms at random ragged shapes and transposes, the
plain Reduce and Scan at random lengths, the generated MapOverlap kernel
on int32 with random getData combinations, AllPairs with a user
function at random ragged sizes, and Map and Zip at random lengths.  Integer results are
checked bit for bit against the oracle, the stencil forward bit for bit
against the oracle's fmaf chain, float sums and scans within one float ulp
of FP64, products normwise against FP64.  Every case prints its seed and
shape first."""
import ctypes as C

import numpy as np
import pytest

import oracle
from skelcl_amd import _lib

pytestmark = pytest.mark.gpu

ADD_F = "float func(float x, float y){ return x+y; }"
ADD_I = "int func(int x, int y){ return x+y; }"


def _layout(rng, n_max, nseg_max):
    """Offsets with runs of empty segments, short and long ones, and an
    uncovered head and tail; n >= the last offset."""
    nseg = int(rng.integers(1, nseg_max + 1))
    kind = rng.integers(0, 3, nseg)
    lens = np.where(kind == 0, 0, np.where(kind == 1, rng.integers(1, 40, nseg),
                                           rng.integers(1, max(2, n_max // max(nseg, 1)), nseg)))
    head = int(rng.integers(0, 3000)) if rng.random() < 0.5 else 0
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64) + head
    tail = int(rng.integers(0, 3000)) if rng.random() < 0.5 else 0
    return off, int(off[-1]) + tail


def _descs(skl, shapes, seed):
    descs = (_lib.skl_gemm_desc * len(shapes))()
    keep, refs = [], []
    for i, (M, N, K) in enumerate(shapes):
        A = oracle.gen_f32_sym(seed + 2 * i, M * K).reshape(M, K)
        B = oracle.gen_f32_sym(seed + 2 * i + 1, K * N).reshape(K, N)
        dA, dB = skl.DeviceBuffer.from_array(A), skl.DeviceBuffer.from_array(B)
        dC = skl.DeviceBuffer(4 * M * N)
        keep += [dA, dB, dC]
        descs[i].A, descs[i].B, descs[i].C = dA.ptr, dB.ptr, dC.ptr
        descs[i].M, descs[i].N, descs[i].K = M, N, K
        refs.append((dC, M, N, A.astype(np.float64) @ B.astype(np.float64)))
    dd = skl.DeviceBuffer(C.sizeof(descs))
    _lib.check(_lib.lib().skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs),
                                         None, 1))
    return dd, keep, refs


@pytest.mark.parametrize("case", range(16))
def test_random_batch_int_exact(skl, case):
    rng = np.random.default_rng(0xB000 + case)
    off, n = _layout(rng, 2_000_000, 3000)
    nseg = off.size - 1
    inclusive = bool(rng.integers(0, 2))
    with_products = case % 2 == 1
    print(f"case {case}: n={n} nseg={nseg} inclusive={inclusive} products={with_products}",
          flush=True)
    x = oracle.gen_u64(0xB100 + case, max(n, 1))[:n].astype(np.uint32).view(np.int32)
    X = skl.DeviceBuffer.from_array(x) if n else None
    O = skl.DeviceBuffer.from_array(off)
    S = skl.DeviceBuffer.from_array(np.full(nseg, 0x5EED, np.int32))
    Y = skl.DeviceBuffer.from_array(np.full(max(n, 1), 0x5EED, np.int32))
    dd, count, refs = None, 0, []
    if with_products:
        shapes = [tuple(int(v) for v in rng.integers(1, 300, 3)) for _ in range(5)]
        dd, keep, refs = _descs(skl, shapes, 0xB200 + 16 * case)
        count = len(shapes)
    red = skl.Reduce("int(int)", ADD_I)
    scan = skl.Scan("int(int)", ADD_I, "0")
    _lib.check(_lib.lib().skl_batch_run(red.handle, S.ptr, scan.handle, Y.ptr, int(inclusive),
                                        X.ptr if X else None, n, O.ptr, nseg,
                                        dd.ptr if dd else None, count, 2, 2, None))
    bad = C.c_int(-1)
    _lib.check(_lib.lib().skl_batch_check(scan.handle, C.byref(bad)))
    assert bad.value == 0
    assert np.array_equal(S.to_numpy(np.uint32, nseg), oracle.segmented_sum_u32(x, off))
    if n:
        assert np.array_equal(Y.to_numpy(np.uint32, n),
                              oracle.segmented_scan_u32(x, off, inclusive, fill=0x5EED))
    for dC, M, N, ref in refs:
        c = dC.to_numpy(np.float32, M * N).reshape(M, N)
        assert np.linalg.norm(c - ref) <= 1e-6 * max(np.linalg.norm(ref), 1e-30), (M, N)


@pytest.mark.parametrize("case", range(12))
def test_random_segmented_pipeline(skl, case):
    rng = np.random.default_rng(0xC000 + case)
    off, n = _layout(rng, 3_000_000, 2000)
    nseg = off.size - 1
    inclusive = bool(rng.integers(0, 2))
    form = ("int", "DEFAULT", "FP64")[case % 3]
    print(f"case {case}: n={n} nseg={nseg} inclusive={inclusive} form={form}", flush=True)
    O = skl.DeviceBuffer.from_array(off)
    if form == "int":
        x = oracle.gen_i32_below(0xC100 + case, max(n, 1), 1 << 30)[:n]
        r, s = skl.Reduce("int(int)", ADD_I), skl.Scan("int(int)", ADD_I, "0")
        fill = np.int32(0x5EED)
    else:
        x = oracle.gen_f32_sym(0xC100 + case, max(n, 1))[:n]
        P = getattr(skl, "PREC_" + form)
        r = skl.Reduce("float(float)", ADD_F, precision=P)
        s = skl.Scan("float(float)", ADD_F, "0.0f", precision=P)
        fill = np.float32(7.0)
    X = skl.DeviceBuffer.from_array(x) if n else skl.DeviceBuffer(16)
    S = skl.DeviceBuffer.from_array(np.full(nseg, fill))
    Y = skl.DeviceBuffer.from_array(np.full(max(n, 1), fill))
    _lib.check(_lib.lib().skl_reduce_segmented_run(r.handle, S.ptr, X.ptr, n, O.ptr, nseg, None))
    _lib.check(_lib.lib().skl_scan_segmented_run(s.handle, Y.ptr, X.ptr, n, O.ptr, nseg,
                                                 int(inclusive), None))
    if form == "int":
        assert np.array_equal(S.to_numpy(np.uint32, nseg), oracle.segmented_sum_u32(x, off))
        if n:
            assert np.array_equal(Y.to_numpy(np.uint32, n),
                                  oracle.segmented_scan_u32(x, off, inclusive,
                                                            fill=int(fill)))
        return
    sums = oracle.segmented_sum_f64(x, off)
    got = S.to_numpy(np.float32, nseg).astype(np.float64)
    assert np.all(np.abs(got - sums) <= np.spacing(np.abs(sums).astype(np.float32)) + 1e-12)
    if n:
        ref = oracle.segmented_scan_f64(x, off, inclusive)
        y = Y.to_numpy(np.float32, n)
        inside = np.zeros(n, bool)
        inside[off[0]:off[-1]] = True
        tol = np.spacing(np.abs(ref[inside]).astype(np.float32)).astype(np.float64) + 1e-9
        assert np.all(np.abs(y[inside].astype(np.float64) - ref[inside]) <= tol)
        assert np.all(y[~inside] == fill)


def _stencil_src(n):
    return f"""float func(input_matrix_t img, __global float* w) {{
  float sum = 0.0f;
  for (int dy = -{n}; dy <= {n}; ++dy)
    for (int dx = -{n}; dx <= {n}; ++dx)
      sum += getData(img, dx, dy) * w[(dy + {n}) * {2 * n + 1} + (dx + {n})];
  return sum;
}}"""


@pytest.mark.parametrize("case", range(16))
def test_random_stencil_forward_and_adjoint(skl, case):
    rng = np.random.default_rng(0xD000 + case)
    rows, cols = (int(v) for v in rng.integers(1, 700, 2))
    R = int(rng.integers(1, 6))
    padding = int(rng.integers(0, 2))
    rb = int(rng.choice([4, 6, 8]))
    print(f"case {case}: {rows}x{cols} R={R} padding={padding} rb={rb}", flush=True)
    w = oracle.gen_f32_sym(0xD100 + case, (2 * R + 1) ** 2).reshape(2 * R + 1, 2 * R + 1)
    img = oracle.gen_f32_sym(0xD200 + case, rows * cols).reshape(rows, cols)
    m = skl.MapOverlap("float(float)", _stencil_src(R), R, padding, 0)
    sh = (C.c_int * 4)()
    _lib.check(_lib.lib().skl_skeleton_launch_shape(m.handle, sh))
    sh[3] = rb
    _lib.check(_lib.lib().skl_skeleton_set_launch_shape(m.handle, sh))
    wv = skl.Vector(w.ravel())
    got = np.asarray(m(skl.Matrix(img), wv)).astype(np.float32)
    pix = np.arange(rows * cols, dtype=np.int64)
    ref = oracle.mapoverlap_stencil_f32_fma(img, w, R, pix, padding, 0.0).reshape(rows, cols)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    g = oracle.gen_f32_sym(0xD300 + case, rows * cols).reshape(rows, cols)
    d = np.asarray(m.stencil_backward(skl.Matrix(g), wv)).astype(np.float64).ravel()
    dref = oracle.mapoverlap_stencil_adjoint_f64(g, w, R, pix, padding)
    # a border pixel under NEAREST collects up to (R + 1)^2 clamped terms per weight
    mult = (R + 1) ** 2 if padding == oracle.NEAREST else 1
    scale = np.abs(w).sum() * np.abs(g).max() * mult
    assert np.max(np.abs(d - dref)) <= 4e-6 * max(scale, 1e-30)


@pytest.mark.parametrize("case", range(18))
def test_random_products_all_forms(skl, case):
    """Dense C = A.B at random ragged shapes and transposes in the three
    forms, against FP64 (normwise: FP32 and BF16X3 1e-6, FP8 5e-2), and the
    MFMA forms' work-group tiles against each other bit for bit."""
    rng = np.random.default_rng(0xE000 + case)
    M, N, K = (int(v) for v in rng.integers(1, 900, 3))
    ta, tb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
    prec = (skl.PREC_FP32, skl.PREC_BF16X3, skl.PREC_FP8)[case % 3]
    print(f"case {case}: M={M} N={N} K={K} ta={ta} tb={tb} prec={prec}", flush=True)
    A = oracle.gen_f32_sym(0xE100 + case, M * K).reshape((K, M) if ta else (M, K))
    B = oracle.gen_f32_sym(0xE200 + case, K * N).reshape((N, K) if tb else (K, N))
    ref = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)
    dA, dB = skl.DeviceBuffer.from_array(A), skl.DeviceBuffer.from_array(B)
    dC = skl.DeviceBuffer(4 * M * N)
    _lib.check(_lib.lib().skl_matmul_f32(dC.ptr, dA.ptr, dB.ptr, M, N, K, ta, tb, prec, None))
    c = dC.to_numpy(np.float32, M * N).reshape(M, N)
    err = np.linalg.norm(c - ref) / max(np.linalg.norm(ref), 1e-30)
    assert err < (5e-2 if prec == skl.PREC_FP8 else 1e-6), err
    if prec == skl.PREC_FP32:
        return
    if ta or tb:
        return  # the tile is chosen through AllPairs, which takes untransposed operands
    add = skl.Reduce("float(float)", ADD_F, "0")
    mul = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    ap = skl.AllPairs("float(float, float)", reduce=add, zip=mul, precision=prec)
    mA, mB = skl.Matrix(A), skl.Matrix(B)
    outs = []
    for t in (1, 2, 3):
        _lib.check(_lib.lib().skl_skeleton_set_mma_tile(ap.handle, t))
        outs.append(np.asarray(ap(mA, mB)).view(np.uint32).copy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert np.array_equal(outs[0], c.view(np.uint32))


@pytest.mark.parametrize("case", range(16))
def test_random_reduce_scan_sizes(skl, case):
    """Reduce(+) and Scan(+) at random lengths (tile remainders of every
    kind): int32 bit for bit against the oracle (modulo 2^32, values that
    wrap), float within one float ulp of FP64 (DEFAULT and FP64 forms)."""
    rng = np.random.default_rng(0xF000 + case)
    n = int(rng.integers(1, 5_000_000))
    inclusive = bool(rng.integers(0, 2))
    print(f"case {case}: n={n} inclusive={inclusive}", flush=True)
    xi = oracle.gen_u64(0xF100 + case, n).astype(np.uint32).view(np.int32)
    Xi = skl.Vector(xi)
    r = skl.Reduce("int(int)", ADD_I)
    s = skl.Scan("int(int)", ADD_I, "0")
    assert np.uint32(np.asarray(r(Xi))[0]) == np.uint32(xi.view(np.uint32).sum(dtype=np.uint64)
                                                        & 0xFFFFFFFF)
    got = np.asarray(s(Xi, inclusive=inclusive)).view(np.uint32)
    assert np.array_equal(got, oracle.scan_add_u32(xi, inclusive))
    xf = oracle.gen_f32_sym(0xF200 + case, n)
    Xf = skl.Vector(xf)
    ref_sum = oracle.sum_f64(xf)
    ref_scan = oracle.scan_add_f64(xf, inclusive)
    tol = np.spacing(np.abs(ref_scan).astype(np.float32)).astype(np.float64) + 1e-9
    for form in ("DEFAULT", "FP64"):
        P = getattr(skl, "PREC_" + form)
        rf = skl.Reduce("float(float)", ADD_F, precision=P)
        sf = skl.Scan("float(float)", ADD_F, "0.0f", precision=P)
        v = float(np.asarray(rf(Xf))[0])
        assert abs(v - ref_sum) <= np.spacing(np.float32(abs(ref_sum))) + 1e-12, form
        y = np.asarray(sf(Xf, inclusive=inclusive)).astype(np.float64)
        assert np.all(np.abs(y - ref_scan) <= tol), form


@pytest.mark.parametrize("case", range(12))
def test_random_mapoverlap_generated_int(skl, case):
    """The generated MapOverlap kernel (any user function) at random sizes,
    radii, rows per thread and both paddings, on int32 with a random
    combination of getData offsets: bit for bit against numpy over a padded
    copy, whose padding is checked against the oracle's getData first."""
    rng = np.random.default_rng(0xA000 + case)
    rows, cols = (int(v) for v in rng.integers(1, 400, 2))
    R = int(rng.integers(1, 7))
    padding = int(rng.integers(0, 2))
    neutral = int(rng.integers(-50, 50))
    rb = int(rng.choice([1, 2, 4, 6, 8]))
    terms = [(int(rng.integers(-R, R + 1)), int(rng.integers(-R, R + 1)),
              int(rng.integers(-9, 10))) for _ in range(5)]
    print(f"case {case}: {rows}x{cols} R={R} padding={padding} neutral={neutral} rb={rb} "
          f"terms={terms}", flush=True)
    body = " + ".join(f"({k}) * getData(img, {dx}, {dy})" for dx, dy, k in terms)
    src = f"int func(input_matrix_t img) {{ return {body}; }}"
    img = oracle.gen_i32_below(0xA100 + case, rows * cols, 1000).reshape(rows, cols)
    m = skl.MapOverlap("int(int)", src, R, padding, neutral)
    sh = (C.c_int * 4)()
    _lib.check(_lib.lib().skl_skeleton_launch_shape(m.handle, sh))
    sh[3] = rb
    _lib.check(_lib.lib().skl_skeleton_set_launch_shape(m.handle, sh))
    got = np.asarray(m(skl.Matrix(img))).astype(np.int64)
    if padding == oracle.NEAREST:
        pad = np.pad(img.astype(np.int64), R, mode="edge")
    else:
        pad = np.pad(img.astype(np.int64), R, mode="constant", constant_values=neutral)
    for _ in range(16):
        r, c = int(rng.integers(0, rows)), int(rng.integers(0, cols))
        dx, dy = int(rng.integers(-R, R + 1)), int(rng.integers(-R, R + 1))
        assert pad[R + r + dy, R + c + dx] == oracle.getdata_i32(img, r, c, dx, dy, padding,
                                                                 neutral)
    ref = np.zeros((rows, cols), np.int64)
    for dx, dy, k in terms:
        ref += k * pad[R + dy:R + dy + rows, R + dx:R + dx + cols]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("case", range(12))
def test_random_allpairs_general_int(skl, case):
    """AllPairs with a user function (the general kernels: column-staged for
    dim <= 256, the LDS and plain forms beyond) at random ragged sizes, int32
    with a weighted L1 distance and an extra argument: bit for bit against
    numpy."""
    rng = np.random.default_rng(0x9000 + case)
    M, N = (int(v) for v in rng.integers(1, 300, 2))
    D = int(rng.integers(1, 400))
    k = int(rng.integers(-5, 6))
    print(f"case {case}: M={M} N={N} D={D} k={k}", flush=True)
    src = ("int func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim, int k) {"
           " int s = 0; for (unsigned int i = 0; i < dim; ++i) {"
           " int d = getElementFromRow(row, i) - getElementFromColumn(col, i);"
           " s += (d < 0 ? -d : d) + k * (int)i; } return s; }")
    A = oracle.gen_i32_below(0x9100 + case, M * D, 2000).reshape(M, D) - 1000
    B = oracle.gen_i32_below(0x9200 + case, D * N, 2000).reshape(D, N) - 1000
    ap = skl.AllPairs("int(int, int)", src)
    got = np.asarray(ap(skl.Matrix(A.astype(np.int32)), skl.Matrix(B.astype(np.int32)), k))
    ref = (np.abs(A.astype(np.int64)[:, :, None] - B.astype(np.int64)[None, :, :]).sum(axis=1)
           + k * (D * (D - 1) // 2))
    assert np.array_equal(got.astype(np.int64), ref)


@pytest.mark.parametrize("case", range(10))
def test_random_map_zip_int(skl, case):
    """Map and Zip with an extra argument at random lengths (every tail of
    the 16-byte packs), int32 arithmetic that wraps: bit for bit against
    numpy modulo 2^32."""
    rng = np.random.default_rng(0x8000 + case)
    n = int(rng.integers(1, 3_000_000))
    a = int(rng.integers(-100000, 100000))
    print(f"case {case}: n={n} a={a}", flush=True)
    x = oracle.gen_u64(0x8100 + case, n).astype(np.uint32).view(np.int32)
    y = oracle.gen_u64(0x8200 + case, n).astype(np.uint32).view(np.int32)
    mp = skl.Map("int(int)", "int func(int x, int a){ return a * x + 3; }")
    zp = skl.Zip("int(int, int)", "int func(int x, int y, int a){ return x * a - (y ^ a); }")
    got_m = np.asarray(mp(skl.Vector(x), np.int32(a))).view(np.uint32)
    got_z = np.asarray(zp(skl.Vector(x), skl.Vector(y), np.int32(a))).view(np.uint32)
    m32 = 0xFFFFFFFF
    ref_m = ((a * x.astype(np.int64) + 3) & m32).astype(np.uint32)
    ref_z = ((x.astype(np.int64) * a - (y ^ np.int32(a)).astype(np.int64)) & m32).astype(np.uint32)
    assert np.array_equal(got_m, ref_m)
    assert np.array_equal(got_z, ref_z)

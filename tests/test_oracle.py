"""Pins the oracle (oracle/oracle.c) against things other than its own
formulas: published generator vectors, the reference's known-answer tests
(test/*.cpp), closed forms, brute force and textbook routines."""
import numpy as np
import pytest

import oracle


# ---------------------------------------------------------------- generator
def test_splitmix64_published_vectors():
    # SplitMix64 (Steele, Lea, Flood 2014), seed 1234567: the first outputs
    # published with the reference implementation (e.g. Rosetta Code).
    exp = [6457827717110365317, 3203168211198807973, 9817491932198370423,
           4593380528125082431, 16408922859458223821]
    assert [oracle.splitmix64(1234567, i) for i in range(5)] == exp


def test_generator_ranges_and_grids():
    x = oracle.gen_f32_sym(0x2002, 1 << 16)
    assert x.min() >= -1.0 and x.max() < 1.0
    assert np.all(x * 2 ** 23 == np.round(x * 2 ** 23))  # exact 2^-23 grid
    u = oracle.gen_f32_unit(0x3001, 1 << 16, 255.0)
    assert u.min() >= 0 and u.max() < 255.0
    k = oracle.gen_i32_below(0x2001, 1 << 16, 1000)
    assert k.min() >= 0 and k.max() < 1000 and len(np.unique(k)) == 1000
    # the float is 2*(z>>40)*2^-24 - 1 of the same z
    z = oracle.gen_u64(0x2002, 8)
    assert np.array_equal(x[:8], (2.0 * ((z >> np.uint64(40)).astype(np.float64) * 2.0 ** -24)
                                  - 1.0).astype(np.float32))
    # Lemire's bounded integer from the high 32 bits
    z = oracle.gen_u64(0x2001, 8)
    assert np.array_equal(k[:8], (((z >> np.uint64(32)) * np.uint64(1000)) >> np.uint64(32))
                          .astype(np.int32))


def test_generator_fixture():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "generator.json")) as f:
        fx = json.load(f)
    for seed_hex, vals in fx["u64"].items():
        assert oracle.gen_u64(int(seed_hex, 16), len(vals)).tolist() == [int(v) for v in vals]


# ---------------------------------------------------------------- Reduce
def test_reduce_kats():  # ReduceTests.cpp:77-136
    assert oracle.reduce_add_f32_reforder(np.arange(100, dtype=np.float32)) == 4950
    assert oracle.reduce_add_u32_reforder(np.arange(1587, dtype=np.int32)) == 1258491
    assert oracle.reduce_add_u32_reforder(np.ones(100000000, np.int32)) == 100000000


@pytest.mark.parametrize("n", list(range(1, 256)) + list(range(8192, 8448)))
def test_reduce_sizes_kat(n):  # ReduceTests.cpp:138-170
    assert oracle.reduce_add_f32_reforder(np.ones(n, np.float32)) == n


@pytest.mark.parametrize("n", [1, 7, 8191, 8192, 8193, 50000, 1 << 20])
def test_reduce_int_closed_form(n):
    x = oracle.gen_i32_below(11, n, 1000)
    assert oracle.reduce_add_u32_reforder(x) == np.uint32(x.astype(np.uint64).sum() % (1 << 32))


def test_reduce_int_wraps_mod_2_32():
    x = np.full(1 << 22, 1000, np.int32)
    assert oracle.reduce_add_u32_reforder(x) == np.uint32((1000 << 22) % (1 << 32))


@pytest.mark.parametrize("n,wg", [(300, 256), (20000, 256), (20000, 64), (100000, 1024)])
def test_reduce_float_order_brute_force(n, wg):
    """Exactly representable integer inputs: any order is exact, so the
    restatement must equal the plain sum; and a hand-written K9->K10 fold in
    Python (tiny sizes) equals it bit for bit on non-integer data."""
    x = oracle.gen_i32_below(3, n, 100).astype(np.float32)
    assert oracle.reduce_add_f32_reforder(x, wg) == np.float32(x.astype(np.float64).sum())
    y = oracle.gen_f32_sym(4, n)
    G = 8192
    D = min(n, G)
    p = [np.float32(y[g]) for g in range(D)]
    for g in range(D):
        for i in range(g + G, n, G):
            p[g] = np.float32(p[g] + y[i])
    L = min(D, wg)
    b = [p[t] for t in range(L)]
    for t in range(L):
        for i in range(t + L, D, L):
            b[t] = np.float32(b[t] + p[i])
    m = L
    while m > 1:
        mod, act = (m + 1) // 2, m // 2
        for i in range(act):
            b[i] = np.float32(b[i] + b[i + mod])
        m = (m + 1) // 2
    assert oracle.reduce_add_f32_reforder(y, wg) == b[0]


def test_sum_f64_is_accurate():
    x = oracle.gen_f32_sym(5, 1 << 20)
    from fractions import Fraction
    exact = float(sum(Fraction(float(v)) for v in x[:20000]))
    assert oracle.sum_f64(x[:20000]) == exact
    assert abs(oracle.sum_f64(x) - np.sum(x.astype(np.float64))) < 1e-9


def test_dot_f64():
    x = oracle.gen_f32_sym(1, 4096)
    y = oracle.gen_f32_sym(2, 4096)
    from fractions import Fraction
    exact = float(sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y)))
    assert oracle.dot_f64(x, y) == exact


# ---------------------------------------------------------------- Scan
@pytest.mark.parametrize("n", [10, 1024])
def test_scan_kats(n):  # ScanTests.cpp:71-107 (exclusive: ones -> i)
    assert np.array_equal(oracle.scan_excl_add_u32_reforder(np.ones(n, np.int32)),
                          np.arange(n, dtype=np.uint32))
    assert np.array_equal(oracle.scan_excl_add_f32_reforder(np.ones(n, np.float32)),
                          np.arange(n, dtype=np.float32))


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 65535, 65536, 65537, 300000,
                               256 ** 3 + 1])
def test_scan_int_reforder_equals_closed_form(n):
    """1..4 passes of the reference's multi-level scan: exact mod 2^32."""
    x = oracle.gen_i32_below(13, n, 1000)
    out = oracle.scan_excl_add_u32_reforder(x, 256, 4)
    cs = np.cumsum(x.astype(np.uint64)) % (1 << 32)
    exp = np.concatenate([[0], cs[:-1]]).astype(np.uint32)
    assert np.array_equal(out, exp)
    assert np.array_equal(oracle.scan_add_u32(x, False), exp)
    assert np.array_equal(oracle.scan_add_u32(x, True), cs.astype(np.uint32))


def test_scan_float_reforder_exact_on_integers():
    x = oracle.gen_i32_below(17, 100000, 50).astype(np.float32)
    out = oracle.scan_excl_add_f32_reforder(x)
    exp = np.concatenate([[0], np.cumsum(x.astype(np.float64))[:-1]]).astype(np.float32)
    assert np.array_equal(out, exp)


def test_scan_f64_and_adjoint():
    x = oracle.gen_f32_sym(21, 5000)
    inc = oracle.scan_add_f64(x, True)
    exc = oracle.scan_add_f64(x, False)
    assert np.allclose(inc, np.cumsum(x.astype(np.float64)), rtol=0, atol=1e-12)
    assert exc[0] == 0 and np.allclose(exc[1:], inc[:-1], rtol=0, atol=1e-12)
    # adjoint of y = scan(x): for loss = sum(g*y), dL/dx = reversed scan of g
    g = oracle.gen_f32_sym(22, 5000)
    for incl in (False, True):
        dx = oracle.scan_add_adjoint_f64(g, incl)
        # FP64 central finite differences at a few coordinates
        for i in (0, 17, 2500, 4999):
            h = 1e-3
            xp = x.astype(np.float64).copy()
            xm = xp.copy()
            xp[i] += h
            xm[i] -= h
            yp = np.cumsum(xp) if incl else np.concatenate([[0], np.cumsum(xp)[:-1]])
            ym = np.cumsum(xm) if incl else np.concatenate([[0], np.cumsum(xm)[:-1]])
            fd = (np.dot(g, yp) - np.dot(g, ym)) / (2 * h)
            assert abs(fd - dx[i]) < 1e-8 * max(1, abs(fd))


# ---------------------------------------------------------------- MapOverlap
def _shift_expect(inp, dx, dy):
    """NEAREST view: value at (r+dy, c+dx) clamped into the matrix."""
    rows, cols = inp.shape
    r = np.clip(np.arange(rows)[:, None] + dy, 0, rows - 1)
    c = np.clip(np.arange(cols)[None, :] + dx, 0, cols - 1)
    return inp[r, c]


@pytest.mark.parametrize("dx,dy", [(-1, 0), (-2, 0), (1, 0), (2, 0), (0, -1), (0, 1),
                                   (0, -2), (0, 2)])
def test_getdata_shift_kats(dx, dy):  # MapOverlapTests.cpp:187-447
    size = 100
    i, j = np.mgrid[0:size, 0:size]
    inp = (i * j + i).astype(np.int32)
    got = np.array([[oracle.getdata_i32(inp, r, c, dx, dy) for c in range(size)]
                    for r in range(size)], np.int32)
    assert np.array_equal(got, _shift_expect(inp, dx, dy))
    # the reference's statements for the right shift: output[i][0] == input[i][0]
    if (dx, dy) == (-1, 0):
        assert np.array_equal(got[:, 0], inp[:, 0])
        assert np.array_equal(got[:, 1:], inp[:, :-1])


def test_getdata_neutral():
    inp = np.arange(20, dtype=np.int32).reshape(4, 5)
    assert oracle.getdata_i32(inp, 0, 0, -1, 0, padding=0, neutral=-7) == -7
    assert oracle.getdata_i32(inp, 3, 4, 0, 1, padding=0, neutral=-7) == -7
    assert oracle.getdata_i32(inp, 1, 1, 1, 1, padding=0, neutral=-7) == inp[2, 2]


def test_stencil_delta_reproduces_weights():
    R = 2
    w = oracle.gen_f32_sym(31, (2 * R + 1) ** 2).reshape(2 * R + 1, 2 * R + 1)
    img = np.zeros((11, 13), np.float32)
    img[5, 6] = 1.0
    out = oracle.mapoverlap_stencil_f32(img, w, R, oracle.NEUTRAL)
    # out[r][c] = sum w[dy+R][dx+R] * img[r+dy][c+dx] = w[5-r+R][6-c+R]
    for r in range(3, 8):
        for c in range(4, 9):
            assert out[r, c] == w[5 - r + R, 6 - c + R]
    assert out[0, 0] == 0


@pytest.mark.parametrize("padding", [0, 1])
def test_stencil_equals_textbook_correlation(padding):
    from scipy.signal import correlate2d
    R = 3
    w = oracle.gaussian_weights(R, 1.5)
    img = oracle.gen_f32_unit(41, 37 * 53, 255.0).reshape(37, 53)
    got = oracle.mapoverlap_stencil_f32(img, w, R, padding)
    if padding == 0:
        exp = correlate2d(img.astype(np.float64), w.astype(np.float64), mode="same",
                          boundary="fill", fillvalue=0)
    else:
        p = np.pad(img.astype(np.float64), R, mode="edge")
        exp = correlate2d(p, w.astype(np.float64), mode="valid")
    assert np.allclose(got, exp, rtol=1e-5, atol=1e-3)
    pix = np.arange(0, 37 * 53, 7, dtype=np.int64)
    f64 = oracle.mapoverlap_stencil_f64(img, w, R, pix, padding)
    assert np.allclose(f64, exp.ravel()[pix], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("padding", [0, 1])
def test_stencil_adjoint_dot_product(padding):
    """<W, A x> == <A^T W, x> in FP64 over a whole small image."""
    R = 2
    w = oracle.gen_f32_sym(51, 25).reshape(5, 5)
    x = oracle.gen_f32_sym(52, 19 * 23).reshape(19, 23)
    g = oracle.gen_f32_sym(53, 19 * 23).reshape(19, 23)
    pix = np.arange(19 * 23, dtype=np.int64)
    Ax = oracle.mapoverlap_stencil_f64(x, w, R, pix, padding)
    ATg = oracle.mapoverlap_stencil_adjoint_f64(g, w, R, pix, padding)
    lhs = np.dot(g.ravel().astype(np.float64), Ax)
    rhs = np.dot(ATg, x.ravel().astype(np.float64))
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs))


# ---------------------------------------------------------------- AllPairs
@pytest.mark.parametrize("M,K,N", [(64, 64, 64), (32, 128, 32), (100, 60, 11)])
def test_allpairs_kats(M, K, N):  # AllPairsTests.cpp:77-133 (integer-valued)
    rng = np.random.default_rng(1)
    A = rng.integers(0, 10, (M, K)).astype(np.float32)
    B = rng.integers(0, 10, (K, N)).astype(np.float32)
    C = oracle.allpairs_mm_f32_seqk(A, B)
    assert np.array_equal(C, (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32))
    assert np.array_equal(oracle.allpairs_mm_f64(A, B), A.astype(np.float64) @ B)
    rows = np.array([0, M - 1, M // 2], np.int32)
    assert np.array_equal(oracle.allpairs_mm_f32_seqk(A, B, rows=rows), C[rows])


def test_allpairs_seqk_is_sequential():
    """The fold runs from the identity in k order (not pairwise): check on a
    case where order changes the fp32 result."""
    A = np.array([[1.0, 1e8, -1e8]], np.float32)
    B = np.array([[1.0], [1.0], [1.0]], np.float32)
    # ((0 + 1) + 1e8) - 1e8 = 0 in fp32; a pairwise order would give 1
    assert oracle.allpairs_mm_f32_seqk(A, B)[0, 0] == np.float32(0.0)


def test_sqdist_closed_form():
    P = oracle.gen_f32_sym(61, 40 * 16).reshape(40, 16)
    D64 = oracle.allpairs_sqdist_f64(P, np.ascontiguousarray(P.T))
    Pd = P.astype(np.float64)
    exp = (Pd ** 2).sum(1)[:, None] + (Pd ** 2).sum(1)[None, :] - 2 * Pd @ Pd.T
    assert np.allclose(D64, exp, atol=1e-12)
    D32 = oracle.allpairs_sqdist_f32(P, np.ascontiguousarray(P.T))
    assert np.allclose(D32, D64, rtol=1e-5, atol=1e-5)
    assert np.all(np.diag(D32) == 0)


# ---------------------------------------------------------------- segmented
def test_segmented():
    x = oracle.gen_f32_sym(71, 1000)
    off = np.array([0, 0, 10, 400, 401, 1000], np.int64)
    s = oracle.segmented_sum_f64(x, off)
    sc = oracle.segmented_scan_f64(x, off, True)
    for k in range(5):
        seg = x[off[k]:off[k + 1]].astype(np.float64)
        assert abs(s[k] - seg.sum()) < 1e-12
        assert np.allclose(sc[off[k]:off[k + 1]], np.cumsum(seg), atol=1e-12)


def test_segmented_u32():
    """Integer segments modulo 2^32: against a Python loop (tiny case, values
    that wrap) and against orc_scan_add_u32 / numpy int64 sums per segment."""
    x = np.array([0xFFFFFFF0, 0x20, 7, 0x80000000, 0x80000000, 1, 2, 3], np.uint32)
    off = np.array([0, 2, 2, 5, 8], np.int64)
    s = oracle.segmented_sum_u32(x, off)
    assert s.tolist() == [0x10, 0, 7, 6]  # 0xFFFFFFF0 + 0x20 wraps to 0x10
    for incl in (False, True):
        sc = oracle.segmented_scan_u32(x, off, incl, fill=0xDEAD)
        exp = []
        for k in range(4):
            acc = 0
            for i in range(off[k], off[k + 1]):
                if incl:
                    acc = (acc + int(x[i])) & 0xFFFFFFFF
                    exp.append(acc)
                else:
                    exp.append(acc)
                    acc = (acc + int(x[i])) & 0xFFFFFFFF
        assert sc.tolist() == exp
    y = oracle.gen_i32_below(73, 5000, 1 << 30)
    o2 = np.array([100, 100, 101, 2500, 4000], np.int64)
    s2 = oracle.segmented_sum_u32(y, o2)
    sc2 = oracle.segmented_scan_u32(y, o2, True, fill=9)
    for k in range(4):
        seg = y[o2[k]:o2[k + 1]].astype(np.int64)
        assert s2[k] == np.uint32(seg.sum() & 0xFFFFFFFF)
        assert np.array_equal(sc2[o2[k]:o2[k + 1]],
                              oracle.scan_add_u32(y[o2[k]:o2[k + 1]], True))
    assert np.all(sc2[:100] == 9) and np.all(sc2[4000:] == 9)


# ---- exactly rounded references (Python Fractions) ------------------------
from fractions import Fraction  # noqa: E402


def _round_f32(q):
    """The float32 nearest the rational q, ties to even: float(q) is the
    nearest double, so the answer is within one float32 step of it."""
    f = np.float32(float(q))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - q)
        key = (d, int(np.array(c, np.float32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return np.float32(best[1])


def _fma_f32(a, b, c):
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_round_f32_matches_numpy_on_sums():
    """The exact rounding agrees with float32 addition, which IEEE rounds
    correctly (a check of the helper itself)."""
    a, b = oracle.gen_f32_sym(71, 300), oracle.gen_f32_sym(72, 300)
    for x, y in zip(a, b):
        assert _round_f32(Fraction(float(x)) + Fraction(float(y))) == np.float32(x + y)


@pytest.mark.parametrize("padding,R", [(0, 2), (1, 2), (0, 1), (1, 3)])
def test_stencil_fma_form_is_the_exact_fma_chain(padding, R):
    """orc_mapoverlap_stencil_f32_fma is sum = fma(getData(dx, dy), w, sum)
    with dy outer and dx inner, every step rounded once: rebuilt here step by
    step with exact rational arithmetic on a non-square image, all pixels
    (the border ones included)."""
    rows, cols = 9, 13
    img = (oracle.gen_f32_unit(0x3501, rows * cols, 255.0)).reshape(rows, cols)
    w = oracle.gen_f32_sym(0x3502, (2 * R + 1) ** 2).reshape(2 * R + 1, 2 * R + 1)
    pix = np.arange(rows * cols, dtype=np.int64)
    got = oracle.mapoverlap_stencil_f32_fma(img, w, R, pix, padding, 0.0)
    for p in range(rows * cols):
        r, c = divmod(p, cols)
        s = np.float32(0.0)
        for dy in range(-R, R + 1):
            for dx in range(-R, R + 1):
                rr, cc = r + dy, c + dx
                if padding == 1:
                    rr, cc = min(max(rr, 0), rows - 1), min(max(cc, 0), cols - 1)
                    v = img[rr, cc]
                else:
                    v = img[rr, cc] if 0 <= rr < rows and 0 <= cc < cols else np.float32(0)
                s = _fma_f32(v, w[dy + R, dx + R], s)
        assert got[p] == s, (p, got[p], s)
    # and it differs from the two-rounding form somewhere (so the test can
    # tell the forms apart)
    two = oracle.mapoverlap_stencil_f32(img, w, R, padding, 0.0)
    assert not np.array_equal(got, two.ravel())


def test_map_zip_helpers_against_numpy_and_exact_fma():
    """saxpy_f32 / scale_f32 / zip_mul_f32 round each operation once, as
    numpy float32 does; saxpy_f32_fma rounds a*x+y once (exact rationals)."""
    x, y = oracle.gen_f32_sym(0x2003, 5000), oracle.gen_f32_sym(0x2004, 5000)
    a = np.float32(2.5)
    assert np.array_equal(oracle.saxpy_f32(a, x, y), (a * x + y).astype(np.float32))
    assert np.array_equal(oracle.saxpy_f32(a, x, y, nthreads=4), oracle.saxpy_f32(a, x, y))
    assert np.array_equal(oracle.scale_f32(a, x), (a * x).astype(np.float32))
    assert np.array_equal(oracle.zip_mul_f32(x, y), (x * y).astype(np.float32))
    a3 = np.float32(1.0 / 3.0)  # a*x is inexact, so one and two roundings differ
    fma = oracle.saxpy_f32_fma(a3, x, y)
    for i in range(0, 5000, 7):
        assert fma[i] == _fma_f32(a3, x[i], y[i]), i
    assert not np.array_equal(fma, oracle.saxpy_f32(a3, x, y))


def test_loguniform_generator_matches_the_committed_fixture():
    """The batched configuration's 4096 segment lengths, written from Python
    integers by tests/golden/make_fixtures.py, are what orc_gen_loguniform
    draws from seed 0x5001 (floor(10^(3 + 3 d)))."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "batched_shapes.json")) as f:
        fx = json.load(f)
    got = oracle.gen_loguniform(0x5001, 4096, 3.0, 3.0)
    assert np.array_equal(got, np.array(fx["segment_lengths"], np.int64))
    assert got.min() >= 1000 and got.max() < 1000000
    # the product shapes: 64 * (1 + U{0..15}) from seed 0x5003, three draws each
    u = ((oracle.gen_u64(0x5003, 3 * 512) >> np.uint64(32)) * np.uint64(16)) >> np.uint64(32)
    shapes = (64 * (1 + u.astype(np.int64))).reshape(512, 3)
    assert np.array_equal(shapes, np.array(fx["gemm_shapes"], np.int64))

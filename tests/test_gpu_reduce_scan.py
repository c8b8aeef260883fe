"""Reduce / Scan on the GPU: the reference's known-answer tests
(test/ReduceTests.cpp, test/ScanTests.cpp) restated, bit-exact integer parity
with the oracle (mod 2^32), float parity against FP64 within a stated bound,
and size-independent properties at the full 2^28 size."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ADD_F = "float func(float x, float y){ return x+y; }"
ADD_I = "int func(int x, int y){ return x+y; }"


def test_reduce_kat_4950(skl):  # ReduceTests.cpp:77-90
    r = skl.Reduce("float(float)", ADD_F)
    out = r(skl.Vector(np.arange(100, dtype=np.float32)))
    assert len(out) >= 1 and out[0] == 4950


def test_reduce_additional_arg_kat(skl):  # ReduceTests.cpp:92-106
    r = skl.Reduce("float(float)", "float func(float x, float y, float a){ return a*(x+y); }")
    out = r(skl.Vector(np.arange(100, dtype=np.float32)), 1.0)
    assert out[0] == 4950


def test_reduce_int_kat(skl):  # ReduceTests.cpp:108-121
    r = skl.Reduce("int(int)", ADD_I)
    assert r(skl.Vector(np.arange(1587, dtype=np.int32)))[0] == 1258491


def test_reduce_long_kat(skl):  # ReduceTests.cpp:123-136
    r = skl.Reduce("int(int)", ADD_I)
    assert r(skl.Vector(np.ones(100000000, np.int32)))[0] == 100000000


@pytest.mark.parametrize("n", list(range(1, 256)))
def test_reduce_sizes_kat(skl, n):  # ReduceTests.cpp:138-153 (ones, as float); 155-170
    # (every size 8192..8447) is test_gpu_full_size.py::test_reduce_all_sizes_8192_to_8447
    r = skl.Reduce("float(float)", ADD_F)
    assert r(skl.Vector(np.ones(n, np.float32)))[0] == n


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4097, 1 << 20, (1 << 20) + 7,
                               3 * (1 << 22) + 5])
def test_reduce_int_parity(skl, n):
    """int32 Reduce(+) is exact mod 2^32: bit-equal to the oracle's
    reference-order fold."""
    x = oracle.gen_i32_below(0x1003, n, 1000)
    r = skl.Reduce("int(int)", ADD_I)
    got = np.int32(r(skl.Vector(x))[0]).view(np.uint32)
    assert got == oracle.reduce_add_u32_reforder(x)


def test_reduce_int_wraps(skl):
    """Sums past 2^31 wrap exactly like uint32 arithmetic."""
    n = 1 << 22
    x = np.full(n, 1000, np.int32)
    r = skl.Reduce("int(int)", ADD_I)
    got = np.int32(r(skl.Vector(x))[0]).view(np.uint32)
    assert got == np.uint32((1000 * n) % (1 << 32))


@pytest.mark.parametrize("n", [1, 100, 4099, 1 << 20, 3 * (1 << 21) + 1])
@pytest.mark.parametrize("prec", ["DEFAULT", "FP64"])
def test_reduce_float_compensated_vs_fp64(skl, n, prec):
    """float Reduce(+), DEFAULT (compensated partials, FP64 cross-block
    combine) and FP64 (FP64 accumulators): |gpu - fp64| <= 2 ulp(|fp64|) +
    1e-6*(sum|x|)*2^-24, and never worse than the reference-order fp32
    fold's own error."""
    x = oracle.gen_f32_sym(0x2002, n)
    r = skl.Reduce("float(float)", ADD_F, precision=getattr(skl, "PREC_" + prec))
    got = float(r(skl.Vector(x))[0])
    ref = oracle.sum_f64(x)
    refo = float(oracle.reduce_add_f32_reforder(x))
    err, err_ref = abs(got - ref), abs(refo - ref)
    tol = 2 * np.spacing(np.float32(abs(ref))) + 1e-6 * np.abs(x).sum() * 2.0 ** -24
    print(f"n={n} gpu_err={err:.3e} reforder_err={err_ref:.3e} tol={tol:.3e}")
    assert err <= tol
    assert err <= max(err_ref, tol)


def test_reduce_float_plain_fp32_user_order(skl):
    """precision FP32 runs the user function itself (no compensation)."""
    x = oracle.gen_f32_sym(0x2002, 1 << 20)
    r = skl.Reduce("float(float)", ADD_F, precision=skl.PREC_FP32)
    got = float(r(skl.Vector(x))[0])
    ref = oracle.sum_f64(x)
    assert abs(got - ref) <= 1e-6 * np.abs(x).sum()


def test_reduce_max_double(skl):
    x = oracle.gen_f32_sym(7, 100001).astype(np.float64)
    r = skl.Reduce("double(double)", "double func(double x, double y){ return x > y ? x : y; }")
    assert r(skl.Vector(x))[0] == x.max()


def test_reduce_empty_raises(skl):
    r = skl.Reduce("float(float)", ADD_F)
    with pytest.raises(skl.SkelError):
        r(skl.Vector(np.zeros(0, np.float32)))


def test_zip_reduce_fused_dot(skl):
    """Zip(x*y) -> Reduce(+) fused into one kernel (config 1 dot product)."""
    from skelcl_amd import _lib
    import ctypes as C
    n = 1 << 20
    x = oracle.gen_f32_sym(0x1001, n)
    y = oracle.gen_f32_sym(0x1002, n)
    z = skl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    r = skl.Reduce("float(float)", ADD_F)
    X, Y = skl.Vector(x), skl.Vector(y)
    o = skl.DeviceBuffer(4)
    _lib.check(_lib.lib().skl_zip_reduce_run(z.handle, r.handle, o.ptr, X.device_ptr(),
                                             Y.device_ptr(), n, None))
    got = float(o.to_numpy(np.float32, 1)[0])
    ref = oracle.dot_f64(x, y)
    assert abs(got - ref) <= 2 * np.spacing(np.float32(abs(ref))) + 1e-12 * n
    # and the unfused pipeline agrees to the same bound
    got2 = float(r(z(X, Y))[0])
    assert abs(got2 - ref) <= 2 * np.spacing(np.float32(abs(ref))) + 1e-12 * n


# ------------------------------------------------------------------ Scan
def test_scan_kat_1024(skl):  # ScanTests.cpp:71-88
    s = skl.Scan("int(int)", ADD_I)
    out = np.asarray(s(skl.Vector(np.ones(1024, np.int32))))
    assert np.array_equal(out, np.arange(1024, dtype=np.int32))


def test_scan_kat_10(skl):  # ScanTests.cpp:90-107
    s = skl.Scan("int(int)", ADD_I)
    out = np.asarray(s(skl.Vector(np.ones(10, np.int32))))
    assert np.array_equal(out, np.arange(10, dtype=np.int32))


def test_scan_float_identity_string(skl):  # ScanTests.cpp:66-69
    s = skl.Scan("float(float)", ADD_F, "0.0f")
    out = np.asarray(s(skl.Vector(np.ones(5, np.float32))))
    assert np.array_equal(out, np.arange(5, dtype=np.float32))


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 4095, 4096, 4097, 65536 + 3,
                               1 << 20, 5 * (1 << 20) + 11])
@pytest.mark.parametrize("inclusive", [False, True])
def test_scan_int_parity(skl, n, inclusive):
    """int32 Scan(+) bit-exact (mod 2^32) vs the oracle; the exclusive form
    also equals the oracle's reference-order multi-pass Blelloch scan."""
    x = oracle.gen_i32_below(0x2001, n, 1000)
    s = skl.Scan("int(int)", ADD_I)
    out = np.asarray(s(skl.Vector(x), inclusive=inclusive)).view(np.uint32)
    exp = oracle.scan_add_u32(x, inclusive)
    assert np.array_equal(out, exp)
    if not inclusive and n <= (1 << 20):
        assert np.array_equal(out, oracle.scan_excl_add_u32_reforder(x))


def test_scan_in_place(skl):
    x = oracle.gen_i32_below(5, 100000, 1000)
    s = skl.Scan("int(int)", ADD_I)
    v = skl.Vector(x)
    s(v, out=skl.out(v), inclusive=True)
    assert np.array_equal(np.asarray(v).view(np.uint32), oracle.scan_add_u32(x, True))


def test_scan_repeated_launches(skl):
    """The look-back state is re-armed by every launch (epoch tags)."""
    s = skl.Scan("int(int)", ADD_I)
    for k in range(5):
        n = 300000 - 17 * k
        x = oracle.gen_i32_below(100 + k, n, 1000)
        out = np.asarray(s(skl.Vector(x))).view(np.uint32)
        assert np.array_equal(out, oracle.scan_add_u32(x, False)), k


@pytest.mark.parametrize("n", [1, 1000, 4097, (1 << 20) + 3])
@pytest.mark.parametrize("inclusive", [False, True])
@pytest.mark.parametrize("prec", ["DEFAULT", "FP64"])
def test_scan_float_vs_fp64(skl, n, inclusive, prec):
    """float Scan(+), DEFAULT (compensated FP32 tile prefixes, FP64 carries)
    and FP64 (FP64 prefixes): every prefix within 1 ulp of the correctly
    rounded FP64 prefix, plus a tiny absolute term; the reference-order
    Blelloch result is reported."""
    x = oracle.gen_f32_sym(0x5002, n)
    s = skl.Scan("float(float)", ADD_F, "0.0f", precision=getattr(skl, "PREC_" + prec))
    out = np.asarray(s(skl.Vector(x), inclusive=inclusive)).astype(np.float64)
    ref = oracle.scan_add_f64(x, inclusive)
    err = np.abs(out - ref)
    tol = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
    if not inclusive:
        ro = oracle.scan_excl_add_f32_reforder(x).astype(np.float64)
        print(f"n={n} max_err gpu={err.max():.3e} reforder={np.abs(ro - ref).max():.3e}")
    assert np.all(err <= tol)


def test_scan_double_max(skl):
    x = oracle.gen_f32_sym(9, 50001).astype(np.float64)
    s = skl.Scan("double(double)", "double func(double x, double y){ return fmax(x,y); }",
                 "-1e300")
    out = np.asarray(s(skl.Vector(x), inclusive=True))
    assert np.array_equal(out, np.maximum.accumulate(x))


def test_scan_full_size_properties(skl):
    """2^28 int32 (the bandwidth configuration): inclusive - exclusive == x
    (exact, element-wise), last inclusive == Reduce(+) (mod 2^32), and a
    strided sample equals the closed form."""
    from skelcl_amd import _lib
    n = 1 << 28
    L = _lib.lib()
    x = skl.DeviceBuffer(4 * n)
    _lib.check(L.skl_generate_i32_below(x.ptr, n, 0x2001, 1000, None))
    ex, inc = skl.DeviceBuffer(4 * n), skl.DeviceBuffer(4 * n)
    s = skl.Scan("int(int)", ADD_I)
    import ctypes as C
    _lib.check(L.skl_scan_run(s.handle, ex.ptr, x.ptr, n, 0, None, 0, None))
    _lib.check(L.skl_scan_run(s.handle, inc.ptr, x.ptr, n, 1, None, 0, None))
    r = skl.Reduce("int(int)", ADD_I)
    tot = skl.DeviceBuffer(4)
    _lib.check(L.skl_reduce_run(r.handle, tot.ptr, x.ptr, n, None, 0, None))
    skl.synchronize()
    hx = x.to_numpy(np.uint32, n)
    he = ex.to_numpy(np.uint32, n)
    hi = inc.to_numpy(np.uint32, n)
    assert np.array_equal(hi - he, hx)
    assert hi[-1] == tot.to_numpy(np.uint32, 1)[0]
    assert he[0] == 0
    # closed form at sampled positions
    cs = np.cumsum(hx.astype(np.uint64)) % (1 << 32)
    idx = np.arange(0, n, 9973)
    assert np.array_equal(hi[idx], cs[idx].astype(np.uint32))


# Pinned launch shapes (unroll, blocks per CU, packs per lane): the Scan tile
# is 256 x packs x V elements, so these cover tiles of 2048 to 16384 int32.
SCAN_SHAPES = [(4, 8, 2), (4, 8, 4), (4, 8, 8), (4, 8, 16)]


def _pin(s, unroll, blocks, packs):
    import ctypes as C
    from skelcl_amd import _lib
    sh = (C.c_int * 4)(unroll, blocks, packs, 4)
    _lib.check(_lib.lib().skl_skeleton_set_launch_shape(s.handle, sh))
    got = (C.c_int * 4)()
    _lib.check(_lib.lib().skl_skeleton_launch_shape(s.handle, got))
    assert tuple(got) == (unroll, blocks, packs, 4)


@pytest.mark.parametrize("shape", SCAN_SHAPES)
@pytest.mark.parametrize("n", [1, 1000, 65536 + 3, 5 * (1 << 20) + 11])
@pytest.mark.parametrize("inclusive", [False, True])
def test_scan_pinned_tile_int_parity(skl, shape, n, inclusive):
    """int32 Scan(+) bit-exact vs the oracle for every tile size."""
    x = oracle.gen_i32_below(0x2101, n, 1000)
    s = skl.Scan("int(int)", ADD_I)
    _pin(s, *shape)
    out = np.asarray(s(skl.Vector(x), inclusive=inclusive)).view(np.uint32)
    assert np.array_equal(out, oracle.scan_add_u32(x, inclusive))


def test_scan_tile_switches_keep_lookback_state(skl):
    """Launches of different sizes and tile sizes in a row on one skeleton:
    the look-back words of earlier launches (other tile counts) never match."""
    s = skl.Scan("int(int)", ADD_I)
    for k, shape in enumerate(SCAN_SHAPES * 2):
        _pin(s, *shape)
        n = 3000000 - 100003 * k if k % 2 == 0 else 1000 + k
        x = oracle.gen_i32_below(200 + k, n, 1000)
        out = np.asarray(s(skl.Vector(x), inclusive=k % 3 == 0)).view(np.uint32)
        assert np.array_equal(out, oracle.scan_add_u32(x, k % 3 == 0)), k


@pytest.mark.parametrize("shape", SCAN_SHAPES[:2])
@pytest.mark.parametrize("n", [4097, 300001])
def test_scan_pinned_tile_backward(skl, shape, n):
    """The reversed scan (Scan(+) adjoint) with small tiles."""
    s = skl.Scan("float(float)", ADD_F, "0.0f")
    _pin(s, *shape)
    g = oracle.gen_f32_sym(0x5105, n)
    d = np.asarray(s.backward(skl.Vector(g), inclusive=True)).astype(np.float64)
    ref = oracle.scan_add_adjoint_f64(g, True)
    tol = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-9
    assert np.all(np.abs(d - ref) <= tol)


def test_fresh_skeleton_uses_default_shape(skl):
    """A new skeleton's first program is generated with its default launch
    shape (the shape is compiled into the kernel)."""
    import ctypes as C
    from skelcl_amd import _lib
    r = skl.Reduce("float(float)", ADD_F)
    sp = C.c_void_p()
    _lib.check(_lib.lib().skl_skeleton_source(r.handle, C.byref(sp)))
    src = _lib.take_string(sp)
    got = (C.c_int * 4)()
    _lib.check(_lib.lib().skl_skeleton_launch_shape(r.handle, got))
    assert f"#define SKL_UNROLL {got[0]}\n" in src


def test_launch_shape_rejects_out_of_range(skl):
    import ctypes as C
    from skelcl_amd import _lib
    s = skl.Scan("int(int)", ADD_I)
    for bad in [(0, 1, 4, 4), (2, 0, 4, 4), (2, 1, 33, 4), (2, 1, 4, 0), (17, 1, 4, 4)]:
        sh = (C.c_int * 4)(*bad)
        assert _lib.lib().skl_skeleton_set_launch_shape(s.handle, sh) != 0, bad

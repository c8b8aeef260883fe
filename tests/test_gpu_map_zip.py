"""Map / Zip on the GPU: the reference's known-answer tests (test/MapTests.cpp,
test/ZipTests.cpp) restated, plus oracle parity at ragged and large sizes."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_map_negate_kat(skl):  # MapTests.cpp:96-130
    m = skl.Map("float(float)", "float func(float f){ return -f; }")
    x = np.arange(10, dtype=np.float32) * 2.5
    out = np.asarray(m(skl.Vector(x)))
    assert np.array_equal(out, -x)


def test_map_add_args_kat(skl):  # MapTests.cpp:158-178
    m = skl.Map("float(float)", "float func(float f, float add, float add2)\n"
                "{ return f+add+add2; }")
    x = np.arange(10, dtype=np.float32) * 2.5
    out = np.asarray(m(skl.Vector(x), 5.25, 7.75))
    exp = (x + np.float32(5.25)) + np.float32(7.75)
    assert np.array_equal(out, exp)


def test_map_int_of_float_device_id(skl):  # MapTests.cpp:132-156 (one device)
    m = skl.Map("int(float)", "int func(float f) { return skelcl_get_device_id(); }")
    out = np.asarray(m(skl.Vector(np.zeros(10, np.float32))))
    assert out.dtype == np.int32 and np.array_equal(out, np.zeros(10, np.int32))


def test_map_void_with_out_arg(skl):  # MapTests.cpp:179-195
    m = skl.Map("void(float)", "void func(float f, __global float* out) \n"
                "{ out[get_global_id(0)] = f; }")
    x = np.arange(10, dtype=np.float32) * 2.5
    o = skl.Vector(size=10, etype="float")
    m(skl.Vector(x), skl.out(o))
    assert np.array_equal(np.asarray(o), x)


def test_map_matrix_args_kat(skl):  # MapTests.cpp:277-325
    m = skl.Map("float(float)", """
float func( float f, float_matrix_t mat, float add2, int_matrix_t mat2 )
{
  return f + get(mat, 1, 1) + add2 + get(mat2, 1, 2);
}
""")
    inp = skl.Matrix(np.arange(12, dtype=np.float32).reshape(4, 3) * 2.5)
    mat = skl.Matrix(np.arange(12, dtype=np.float32).reshape(4, 3))
    m2 = np.zeros((2, 3), np.int32)
    m2[1, 2] = 5
    out = np.asarray(m(inp, mat, 7.75, skl.Matrix(m2)))
    exp = ((np.asarray(inp) + np.float32(4.0)) + np.float32(7.75)) + np.float32(5)
    assert out.shape == (4, 3) and np.array_equal(out, exp)


def test_map_local_memory_kat(skl):  # MapTests.cpp:348-365
    m = skl.Map("int(int)", "int func(int i, __local int* lp) \n"
                "{ lp[0] = i; barrier(CLK_LOCAL_MEM_FENCE); return -lp[0]; }")
    m.set_work_group_size(16)
    x = np.full(32, 5, np.int32)
    out = np.asarray(m(skl.Vector(x), skl.Local(4)))
    assert np.array_equal(out, -x)


def test_map_float4_kat(skl):  # MapTests.cpp:367-390
    m = skl.Map("float4(float4)", "float4 func(float4 f) { return -f; }")
    x = np.tile(np.array([1, 2, 3, 4], np.float32), (10, 1))
    out = np.asarray(m(skl.Vector(x, etype="float4")))
    assert out.shape == (10, 4) and np.array_equal(out, -x)


def test_map_double4_kat(skl):  # MapTests.cpp:392-415
    m = skl.Map("double4(double4)", "double4 func(double4 f) { return -f; }")
    x = np.tile(np.array([1, 2, 3, 4], np.float64), (10, 1))
    out = np.asarray(m(skl.Vector(x, etype="double4")))
    assert np.array_equal(out, -x)


@pytest.mark.parametrize("n", [1, 3, 4, 1000, 65537, (1 << 20) + 3])
def test_map_scale_parity(skl, n):
    """Map(a*x) vs the oracle: bit-exact (one rounding either way)."""
    m = skl.Map("float(float)", "float func(float x, float a){ return a*x; }")
    x = oracle.gen_f32_sym(0x2003, n)
    out = np.asarray(m(skl.Vector(x), np.float32(2.5)))
    assert np.array_equal(out, oracle.scale_f32(2.5, x))


def test_zip_add_kat(skl):  # ZipTests.cpp:82-104
    z = skl.Zip("float(float, float)", "float func(float x, float y){ return x+y; }")
    l = np.arange(100, dtype=np.float32)
    r = np.arange(100, dtype=np.float32) * 2
    assert np.array_equal(np.asarray(z(skl.Vector(l), skl.Vector(r))), l + r)


def test_zip_add_args_kat(skl):  # ZipTests.cpp:106-131
    z = skl.Zip("float(float, float)",
                "float func(float x, float y, float a, float b){ return x*y+a+b; }")
    l = np.arange(100, dtype=np.float32)
    r = np.arange(100, dtype=np.float32) * 2
    out = np.asarray(z(skl.Vector(l), skl.Vector(r), 5.5, 3.25))
    # user function x*y+a+b: fma(x, y, a) + b under contraction; exact for
    # these small integer-valued products
    assert np.array_equal(out, (l * r + np.float32(5.5)) + np.float32(3.25))


def test_zip_out_aliases_left_kat(skl):  # ZipTests.cpp:133-157
    z = skl.Zip("float(float, float)", "float func(float x, float y) { return x; }")
    l = skl.Vector(np.arange(50, dtype=np.float32))
    r = skl.Vector(np.full(50, 7, np.float32))
    z(l, r, out=skl.out(l))
    assert np.array_equal(np.asarray(l), np.arange(50, dtype=np.float32))


def test_zip_matrix_kat(skl):  # ZipTests.cpp:159-186
    z = skl.Zip("float(float,float)", "float func(float mat1,float mat2)"
                "{ return mat1 + mat2; }")
    a = np.arange(64, dtype=np.float32).reshape(8, 8)
    out = np.asarray(z(skl.Matrix(a), skl.Matrix(a)))
    assert out.shape == (8, 8) and np.array_equal(out, a + a)


def test_zip_void_kat(skl):  # ZipTests.cpp:188-214
    z = skl.Zip("void(float, float)",
                "void func(float x, float y, __global float* out)"
                "{ out[get_global_id(0)] = x+y; }")
    l = np.arange(20, dtype=np.float32)
    o = skl.Vector(size=20, etype="float")
    z(skl.Vector(l), skl.Vector(l), skl.out(o))
    assert np.array_equal(np.asarray(o), l + l)


@pytest.mark.parametrize("n", [1, 5, 4099, (1 << 20) + 1])
def test_zip_saxpy_parity(skl, n):
    """saxpy in the reference's Zip form (examples/saxpy/main.cpp:77-88):
    bit-exact vs fma(a,x,y) (contracted), <= 1 ulp vs the two-rounding form."""
    z = skl.Zip("float(float, float)",
                "float func(float x, float y, float a){ return a*x+y; }")
    x = oracle.gen_f32_sym(0x2003, n)
    y = oracle.gen_f32_sym(0x2004, n)
    out = np.asarray(z(skl.Vector(x), skl.Vector(y), np.float32(2.5)))
    assert np.array_equal(out, oracle.saxpy_f32_fma(2.5, x, y))
    # the two-rounding form differs from the fused one by the rounding of a*x
    # (<= ulp(a*x)/2) plus the final roundings: many ulps of the result when
    # a*x + y cancels
    two = oracle.saxpy_f32(2.5, x, y)
    ax = np.abs(np.float32(2.5) * x)
    bound = np.spacing(ax).astype(np.float64) + np.spacing(
        np.maximum(np.abs(out), np.abs(two))).astype(np.float64)
    assert np.all(np.abs(out.astype(np.float64) - two) <= bound)


def test_zip_size_mismatch_raises(skl):  # ZipDef.h:101,126
    z = skl.Zip("float(float, float)", "float func(float x, float y){ return x+y; }")
    with pytest.raises(skl.SkelError):
        z(skl.Vector(np.zeros(4, np.float32)), skl.Vector(np.zeros(5, np.float32)))


def test_map_index_kat(skl):  # IndexVectorTests.cpp:61-98
    m = skl.Map("int(Index)", "int func(Index i){ return (int)i; }")
    out = np.asarray(m(skl.IndexVector(1000)))
    assert np.array_equal(out, np.arange(1000, dtype=np.int32))


def test_map_indexpoint_kat(skl):  # IndexMatrixTests.cpp:61-107
    m = skl.Map("int(IndexPoint)", "int func(IndexPoint p){ return p.x + 1000*p.y; }")
    out = np.asarray(m(skl.IndexMatrix((37, 129))))
    r, c = np.mgrid[0:37, 0:129]
    assert np.array_equal(out, (c + 1000 * r).astype(np.int32))


@pytest.mark.parametrize("n", [1, 255, 256, 100003])
def test_map_void_index_writes_through_an_out_buffer(skl, n):
    """Map<void(Index)> (src/Map.cpp:53-103): no output container; the user
    function stores through its extra buffer, one call per index."""
    o = skl.Vector(np.full(n, -1, np.int32))
    m = skl.Map("void(Index)", "void func(Index i, __global int* o){ o[i] = 3 * (int)i + 1; }")
    assert m(skl.IndexVector(n), skl.out(o)) is None
    assert np.array_equal(np.asarray(o), 3 * np.arange(n, dtype=np.int32) + 1)


@pytest.mark.parametrize("rows,cols", [(1, 1), (37, 129), (300, 65)])
def test_map_void_indexpoint_writes_through_an_out_buffer(skl, rows, cols):
    """Map<void(IndexPoint)> (src/Map.cpp:105-159): p.x is the column, p.y
    the row; every point is visited once."""
    o = skl.Vector(np.zeros(rows * cols, np.int32))
    m = skl.Map("void(IndexPoint)",
                "void func(IndexPoint p, __global int* o, int cols)"
                "{ o[p.y * cols + p.x] += p.x - 7 * p.y + 1; }")
    assert m(skl.IndexMatrix((rows, cols)), skl.out(o), cols) is None
    r, c = np.mgrid[0:rows, 0:cols]
    assert np.array_equal(np.asarray(o).reshape(rows, cols), (c - 7 * r + 1).astype(np.int32))


def test_generators_bit_exact(skl):
    """The device generators reproduce the host SplitMix64 restatement."""
    n = 100003
    b = skl.DeviceBuffer(4 * n)
    from skelcl_amd import _lib
    _lib.check(_lib.lib().skl_generate_f32_sym(b.ptr, n, 0x2002, None))
    assert np.array_equal(b.to_numpy(np.float32, n), oracle.gen_f32_sym(0x2002, n))
    _lib.check(_lib.lib().skl_generate_f32_unit(b.ptr, n, 0x3001, 255.0, None))
    assert np.array_equal(b.to_numpy(np.float32, n), oracle.gen_f32_unit(0x3001, n, 255.0))
    _lib.check(_lib.lib().skl_generate_i32_below(b.ptr, n, 0x2001, 1000, None))
    assert np.array_equal(b.to_numpy(np.int32, n), oracle.gen_i32_below(0x2001, n, 1000))


@pytest.mark.parametrize("wg", [64, 128, 512, 1024])
def test_pinned_work_group_sizes(skl, wg):
    """Every streaming skeleton launches with the work-group size it was
    compiled for (setWorkGroupSize): exact results at 64 .. 1024 threads."""
    import oracle
    n = 1000003
    a, b = oracle.gen_f32_sym(71, n), oracle.gen_f32_sym(72, n)
    m = skl.Map("float(float)", "float func(float x){ return -x; }")
    m.set_work_group_size(wg)
    assert np.array_equal(np.asarray(m(skl.Vector(a))), -a)
    z = skl.Zip("float(float, float)", "float func(float x, float y){ return x + y; }")
    z.set_work_group_size(wg)
    assert np.array_equal(np.asarray(z(skl.Vector(a), skl.Vector(b))), a + b)
    xi = oracle.gen_i32_below(73, n, 1000)
    r = skl.Reduce("int(int)", "int func(int x, int y){ return x + y; }")
    r.set_work_group_size(wg)
    assert np.int64(np.asarray(r(skl.Vector(xi)))[0]) == int(xi.astype(np.int64).sum())
    s = skl.Scan("int(int)", "int func(int x, int y){ return x + y; }")
    s.set_work_group_size(wg)
    out = np.asarray(s(skl.Vector(xi), inclusive=True)).astype(np.int64)
    assert np.array_equal(out, np.cumsum(xi.astype(np.int64)))


def test_reduce_rejects_ragged_work_group_size(skl):
    r = skl.Reduce("int(int)", "int func(int x, int y){ return x + y; }")
    r.set_work_group_size(100)
    with pytest.raises(Exception):
        r(skl.Vector(np.ones(1000, np.int32)))

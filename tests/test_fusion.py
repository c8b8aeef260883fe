"""User-function fusion on the CPU: the signature reader (in place of the
reference's libclang rewriter; its golden tests in
libraries/stooling/test/*.cpp restated), op detection, and a hiprtc compile
of every skeleton kind's fused gfx950 kernel (no GPU needed)."""
import ctypes as C

import pytest

import skelcl_amd as skl
from skelcl_amd import _lib


def describe(src, func="func"):
    p = C.c_void_p()
    _lib.check(_lib.lib().skl_describe_user_function(src.encode(), func.encode(), C.byref(p)))
    lines = _lib.take_string(p).strip().split("\n")
    out = {"return": None, "params": [], "op": None, "work_item": None}
    for ln in lines:
        key, _, rest = ln.partition(" ")
        if key == "return":
            out["return"] = rest
        elif key == "param":
            kind, size, decl = rest.split(" ", 2)
            t, name = decl.split("|")
            out["params"].append((int(kind), int(size), t, name))
        elif key == "op":
            out["op"] = rest
        elif key == "work_item":
            out["work_item"] = rest == "1"
        elif key == "stencil":
            out["stencil"] = int(rest)
    return out


STENCIL = """float func(input_matrix_t img, __global float* w) {
  float sum = 0.0f;
  for (int dy = -5; dy <= 5; ++dy)
    for (int dx = -5; dx <= 5; ++dx)
      sum += getData(img, dx, dy) * w[(dy + 5) * 11 + (dx + 5)];
  return sum;
}"""


def test_weighted_stencil_recognised():
    """The weighted-stencil form (any names, either operand order, braces,
    `s = s + p`, postfix increments, const weights, zero spellings) is
    recognised with its radius; each listed change to the arithmetic or
    the window is not."""
    assert describe(STENCIL)["stencil"] == 5
    same = [
        STENCIL.replace("sum", "acc").replace("img", "m").replace("dy", "y").replace("dx", "x"),
        STENCIL.replace("getData(img, dx, dy) * w[(dy + 5) * 11 + (dx + 5)]",
                        "w[(dy + 5) * 11 + (dx + 5)] * getData(img, dx, dy)"),
        STENCIL.replace("(dy + 5) * 11", "11 * (dy + 5)").replace("+ (dx + 5)]", "+ dx + 5]"),
        STENCIL.replace("sum += ", "sum = sum + "),
        STENCIL.replace("++dy", "dy++").replace("++dx", "dx++"),
        STENCIL.replace("__global float* w", "const __global float* w"),
        STENCIL.replace("0.0f", "0.f"),
        STENCIL.replace("0.0f", "0"),
        STENCIL.replace("    for (int dx", "  {  for (int dx").replace("  return", "  }\n  return"),
        "/* lead */ " + STENCIL.replace("return sum;", "return sum; // out"),
    ]
    for src in same:
        assert describe(src)["stencil"] == 5, src
    three = (STENCIL.replace("-5", "-3").replace("<= 5", "<= 3").replace("+ 5)", "+ 3)")
             .replace("* 11", "* 7"))
    assert describe(three)["stencil"] == 3
    zero = (STENCIL.replace("-5", "-0").replace("<= 5", "<= 0").replace("+ 5)", "+ 0)")
            .replace("* 11", "* 1"))
    assert describe(zero)["stencil"] == 0
    other = [
        STENCIL.replace("float sum = 0.0f", "float sum = 1.0f"),
        STENCIL.replace("sum += ", "sum -= "),
        STENCIL.replace("getData(img, dx, dy)", "getData(img, dy, dx)"),
        STENCIL.replace("(dy + 5) * 11 + (dx + 5)", "(dx + 5) * 11 + (dy + 5)"),
        STENCIL.replace("* 11", "* 12"),
        STENCIL.replace("<= 5; ++dx", "<= 4; ++dx"),
        STENCIL.replace("for (int dy = -5; dy <= 5; ++dy)\n    for (int dx = -5; dx <= 5; ++dx)",
                        "for (int dx = -5; dx <= 5; ++dx)\n    for (int dy = -5; dy <= 5; ++dy)"),
        STENCIL.replace("return sum;", "return sum * 2.0f;"),
        STENCIL.replace("float func", "double func"),
        STENCIL.replace("__global float* w", "__global double* w"),
        STENCIL.replace("__global float* w", "__constant float* w"),
        "#define SCALE 1\n" + STENCIL,
        STENCIL.replace("__global float* w)", "__global float* w, int k)"),
        STENCIL.replace("for (int dx", "for (int dy"),  # the inner loop shadows dy
        STENCIL.replace("w[", "img["),
    ]
    for src in other:
        assert describe(src)["stencil"] == -1, src
    assert describe("float func(float a, float b) { return a + b; }")["stencil"] == -1


# --- TransferParametersTest.cpp:46-117: the parameters from index k are the
#     ones appended to the kernel ("int y, float z" from foo(int x, ...))
def test_transfer_single_parameter():
    d = describe("void foo(int x, int y) {}", "foo")
    assert [p[3] for p in d["params"][1:]] == ["y"]
    assert [p[2] for p in d["params"][1:]] == ["int"]


def test_transfer_two_parameters():
    d = describe("void bar(int i);\nvoid foo(int x, int y, float z){ }", "foo")
    assert [(p[2], p[3]) for p in d["params"][1:]] == [("int", "y"), ("float", "z")]


def test_opencl_qualified_parameters():
    d = describe("__kernel void bar(__global int* i, __local float* l, const unsigned int n)"
                 "{ }", "bar")
    kinds = [p[0] for p in d["params"]]
    assert kinds == [_lib.SKL_ARG_BUFFER, _lib.SKL_ARG_LOCAL, _lib.SKL_ARG_VALUE]
    assert d["params"][2][1] == 4


def test_parameter_type_names():  # GetParameterTypeNamesTest.cpp:46-75
    d = describe("void foo(const int x, int* y) { }", "foo")
    assert d["params"][0][2] == "const int" and d["params"][0][1] == 4
    assert d["params"][1][2] == "int *" and d["params"][1][0] == _lib.SKL_ARG_BUFFER


def test_matrix_parameter_kind():
    d = describe("float func( float f, float_matrix_t mat, float add2, int_matrix_t mat2 )"
                 "{ return f + get(mat, 1, 1) + add2 + get(mat2, 1, 2); }")
    assert [p[0] for p in d["params"]] == [0, _lib.SKL_ARG_MATRIX, 0, _lib.SKL_ARG_MATRIX]


def test_definition_not_call_and_comments():
    src = """
    // float func(float a) { return a; } is a comment
    /* float func(int q) {} */
    float helper(float v) { return func2(v); }
    float func(float x, float y) { return helper(x) + y; }
    """
    d = describe(src)
    assert d["return"] == "float" and [p[3] for p in d["params"]] == ["x", "y"]


def test_missing_function_is_parse_error():
    p = C.c_void_p()
    st = _lib.lib().skl_describe_user_function(b"int foo(int x){return x;}", b"func",
                                                C.byref(p))
    assert st == _lib.SKL_ERR_PARSE


@pytest.mark.parametrize("src,op", [
    ("float func(float x, float y){ return x+y; }", "add"),
    ("int func(int a, int b) { return (b + a); }", "add"),
    ("float func(float x, float y){ return x*y; }", "mul"),
    ("float func(float x, float y){ return fmax(x,y); }", "max"),
    ("float func(float x, float y){ return x < y ? x : y; }", "min"),
    ("float func(float x, float y){ return x+y+1; }", "other"),
    ("float func(float x, float y, float a){ return a*(x+y); }", "other"),
])
def test_op_detection(src, op):
    assert describe(src)["op"] == op


def test_work_item_detection():
    assert describe("void func(float f, __global float* out) "
                    "{ out[get_global_id(0)] = f; }")["work_item"]
    assert not describe("float func(float f) { return -f; }")["work_item"]


# --------------------------------------------------------------- hiprtc builds
GAUSS = """float func(input_matrix_t img, __global float* w) {
  float sum = 0.0f;
  for (int dy = -5; dy <= 5; ++dy)
    for (int dx = -5; dx <= 5; ++dx)
      sum += getData(img, dx, dy) * w[(dy + 5) * 11 + (dx + 5)];
  return sum; }"""
SQDIST = ("float func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim) { float s = 0;"
          " for (unsigned k = 0; k < dim; ++k) { float t = getElementFromRow(row, k) -"
          " getElementFromColumn(col, k); s += t*t; } return s; }")

CASES = [
    (_lib.SKL_MAP, ["float", "float"], "float func(float f){ return -f; }", {}),
    (_lib.SKL_MAP, ["float", "float"],
     "float func(float f, float add, float add2){ return f+add+add2; }", {}),
    (_lib.SKL_MAP, ["void", "float"],
     "void func(float f, __global float* out) { out[get_global_id(0)] = f; }", {}),
    (_lib.SKL_MAP, ["float", "float"], "float func( float f, float_matrix_t mat, float add2,"
     " int_matrix_t mat2 ) { return f + get(mat, 1, 1) + add2 + get(mat2, 1, 2); }", {}),
    (_lib.SKL_MAP, ["int", "int"], "int func(int i, __local int* lp) { lp[0] = i;"
     " barrier(CLK_LOCAL_MEM_FENCE); return -lp[0]; }", {}),
    (_lib.SKL_MAP, ["float4", "float4"], "float4 func(float4 f) { return -f; }", {}),
    (_lib.SKL_MAP, ["double4", "double4"], "double4 func(double4 f) { return -f; }", {}),
    (_lib.SKL_MAP_INDEX, ["float"], "float func(Index i){ return (float)i; }", {}),
    (_lib.SKL_MAP_INDEXPOINT, ["int"], "int func(IndexPoint p){ return p.x + 10*p.y; }", {}),
    (_lib.SKL_ZIP, ["float", "float", "float"],
     "float func(float x, float y, float a){ return a*x+y; }", {}),
    (_lib.SKL_ZIP, ["void", "float", "float"], "void func(float x, float y, __global float*"
     " out){ out[get_global_id(0)] = x+y; }", {}),
    (_lib.SKL_REDUCE, ["float"], "float func(float x, float y){ return x+y; }", {}),
    (_lib.SKL_REDUCE, ["int"], "int func(int x, int y){ return x+y; }", {}),
    (_lib.SKL_REDUCE, ["float"], "float func(float x, float y, float a){ return a*(x+y); }",
     {}),
    (_lib.SKL_SCAN, ["int"], "int func(int x, int y){ return x+y; }", {}),
    (_lib.SKL_SCAN, ["float"], "float func(float x, float y){ return x+y; }",
     {"identity": "0.0f"}),
    (_lib.SKL_SCAN, ["double"], "double func(double x, double y){ return fmax(x,y); }",
     {"identity": "-1e300"}),
    (_lib.SKL_MAPOVERLAP, ["int", "int"], "int func(input_matrix_t f){ return getData(f, -1, 0); }",
     {"overlap_range": 1}),
    (_lib.SKL_MAPOVERLAP, ["float", "float"], GAUSS, {"overlap_range": 5, "padding": 0}),
    (_lib.SKL_ALLPAIRS, ["float", "float", "float"], "float func(float x, float y){ return x+y; }",
     {"source2": "float func(float x, float y){ return x*y; }"}),
    (_lib.SKL_ALLPAIRS, ["float", "float", "float"],
     "float func(float x, float y, float b){ return x+y+b; }",
     {"source2": "float func(float x, float y, float a){ return x*y+a; }"}),
    (_lib.SKL_ALLPAIRS_GENERAL, ["float", "float", "float"], SQDIST, {}),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fused_kernel_compiles_for_gfx950(case):
    kind, types, src, kw = CASES[case]
    text = skl.check_source(kind, types, src, **kw)
    assert "extern \"C\" __global__" in text
    assert "#pragma clang force_cuda_host_device begin" in text


def test_user_compile_error_is_reported():
    with pytest.raises(skl.SkelError) as ei:
        skl.check_source(_lib.SKL_MAP, ["float", "float"], "float func(float f){ return f +; }")
    assert ei.value.status == _lib.SKL_ERR_COMPILE
    assert "user_source" in str(ei.value)


def test_empty_source_is_invalid():
    with pytest.raises(skl.SkelError) as ei:
        skl.check_source(_lib.SKL_MAP, ["float", "float"], "")
    assert ei.value.status == _lib.SKL_ERR_INVALID


def test_wrong_type_count_is_invalid():
    with pytest.raises(skl.SkelError):
        skl.check_source(_lib.SKL_ZIP, ["float", "float"],
                         "float func(float x, float y){ return x; }")

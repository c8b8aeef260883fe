"""The native weighted-stencil kernel behind MapOverlap (native/stencil.hip,
chosen when the user function is the weighted stencil, radius <= 5) against
the generated MapOverlap kernel on the same source (a preprocessor line in
front keeps the recogniser away, so the generic path runs) and against the
oracle: bit-exact on every pixel, for radii 0-5, both paddings, neutral
values, ragged shapes smaller and larger than a tile, every rows-per-thread
instantiation, contraction on and off, and a stencil narrower than the
skeleton's overlap range.  Weights are signed and asymmetric, so a
transposed or mirrored weight index fails."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (3, 200), (200, 3), (65, 130), (257, 129), (16, 64)]


def stencil_src(n):
    return f"""float func(input_matrix_t img, __global float* w) {{
  float sum = 0.0f;
  for (int dy = -{n}; dy <= {n}; ++dy)
    for (int dx = -{n}; dx <= {n}; ++dx)
      sum += getData(img, dx, dy) * w[(dy + {n}) * {2 * n + 1} + (dx + {n})];
  return sum;
}}"""


GENERIC = "#define SKL_TEST_GENERIC 1\n"


def _set_rb(skl, m, rb):
    from skelcl_amd import _lib
    sh = (C.c_int * 4)()
    _lib.check(_lib.lib().skl_skeleton_launch_shape(m.handle, sh))
    sh[3] = rb
    _lib.check(_lib.lib().skl_skeleton_set_launch_shape(m.handle, sh))


def _run(skl, m, img, w):
    return np.asarray(m(skl.Matrix(img), skl.Vector(w.ravel()))).astype(np.float32)


def _weights(n, seed):
    return oracle.gen_f32_sym(seed, (2 * n + 1) ** 2).reshape(2 * n + 1, 2 * n + 1)


def test_recogniser_sees_the_stencil(skl):
    from skelcl_amd import _lib
    p = C.c_void_p()
    _lib.check(_lib.lib().skl_describe_user_function(stencil_src(5).encode(), b"func",
                                                     C.byref(p)))
    assert "stencil 5" in _lib.take_string(p)
    p = C.c_void_p()
    _lib.check(_lib.lib().skl_describe_user_function((GENERIC + stencil_src(5)).encode(),
                                                     b"func", C.byref(p)))
    assert "stencil -1" in _lib.take_string(p)


@pytest.mark.parametrize("R", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("padding", [0, 1])
def test_native_matches_generic_and_oracle(skl, R, padding):
    w = _weights(R, 0x5100 + R)
    nat = skl.MapOverlap("float(float)", stencil_src(R), R, padding, 0)
    gen = skl.MapOverlap("float(float)", GENERIC + stencil_src(R), R, padding, 0)
    cases = []
    for k, (rows, cols) in enumerate(SHAPES):
        img = oracle.gen_f32_sym(0x5200 + 16 * R + k, rows * cols).reshape(rows, cols)
        ref = _run(skl, gen, img, w)
        pix = np.arange(rows * cols, dtype=np.int64)
        fma = oracle.mapoverlap_stencil_f32_fma(img, w, R, pix, padding, 0.0).reshape(rows, cols)
        assert np.array_equal(ref.view(np.uint32), fma.view(np.uint32)), (rows, cols)
        cases.append((img, ref))
    for rb in (4, 6, 8):
        _set_rb(skl, nat, rb)
        for img, ref in cases:
            print(f"R={R} padding={padding} shape={img.shape} rb={rb}", flush=True)
            got = _run(skl, nat, img, w)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (img.shape, rb)


@pytest.mark.parametrize("neutral", [-7.5, 3, 1e30])
def test_native_neutral_values(skl, neutral):
    R, rows, cols = 3, 90, 150
    w = _weights(R, 0x5301)
    img = oracle.gen_f32_sym(0x5302, rows * cols).reshape(rows, cols)
    nat = skl.MapOverlap("float(float)", stencil_src(R), R, skl.NEUTRAL, neutral)
    gen = skl.MapOverlap("float(float)", GENERIC + stencil_src(R), R, skl.NEUTRAL, neutral)
    got, ref = _run(skl, nat, img, w), _run(skl, gen, img, w)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    pix = np.arange(rows * cols, dtype=np.int64)
    fma = oracle.mapoverlap_stencil_f32_fma(img, w, R, pix, 0, float(np.float32(neutral)))
    assert np.array_equal(got.ravel().view(np.uint32), fma.view(np.uint32))


@pytest.mark.parametrize("R", [2, 5])
@pytest.mark.parametrize("padding", [0, 1])
def test_native_contraction_off(skl, R, padding):
    """-ffp-contract=off: a rounded product and a rounded sum per term, as
    the oracle's two-rounding loop."""
    rows, cols = 131, 197
    w = _weights(R, 0x5401)
    img = oracle.gen_f32_sym(0x5402, rows * cols).reshape(rows, cols)
    nat = skl.MapOverlap("float(float)", stencil_src(R), R, padding, 0,
                         build_options="-ffp-contract=off")
    got = _run(skl, nat, img, w)
    o32 = oracle.mapoverlap_stencil_f32(img, w, R, padding)
    assert np.array_equal(got.view(np.uint32), o32.view(np.uint32))


def test_stencil_narrower_than_the_overlap(skl):
    """overlap 5, a radius-2 loop: the native kernel runs radius 2 and still
    reads what the generated kernel reads."""
    rows, cols = 100, 77
    w = _weights(2, 0x5501)
    img = oracle.gen_f32_sym(0x5502, rows * cols).reshape(rows, cols)
    for padding in (0, 1):
        nat = skl.MapOverlap("float(float)", stencil_src(2), 5, padding, 0)
        gen = skl.MapOverlap("float(float)", GENERIC + stencil_src(2), 5, padding, 0)
        got, ref = _run(skl, nat, img, w), _run(skl, gen, img, w)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_radius_beyond_the_native_kernel_runs_generic(skl):
    R, rows, cols = 6, 70, 90
    w = _weights(R, 0x5601)
    img = oracle.gen_f32_sym(0x5602, rows * cols).reshape(rows, cols)
    m = skl.MapOverlap("float(float)", stencil_src(R), R, 1, 0)
    got = _run(skl, m, img, w)
    pix = np.arange(rows * cols, dtype=np.int64)
    fma = oracle.mapoverlap_stencil_f32_fma(img, w, R, pix, 1, 0.0)
    assert np.array_equal(got.ravel().view(np.uint32), fma.view(np.uint32))

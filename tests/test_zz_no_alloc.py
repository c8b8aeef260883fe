"""skl_reduce_run_noalloc against skl_reduce_run through the C-ABI: the same
kernel on the same grid, so the results are bit-identical; two calls in a
row, the second on a smaller input, for float(+) (compensated partials),
int(+) (mod 2^32) and a user function (max).  Also: the no-allocation form
rejects a launch shape whose scratch exceeds its static array, and it can
be captured in a graph and replayed."""
import ctypes as C

import numpy as np
import pytest

import oracle
from skelcl_amd import _lib

pytestmark = pytest.mark.gpu

CASES = [
    ("float(float)", "float func(float x, float y){ return x+y; }", np.float32, 0x2002),
    ("int(int)", "int func(int x, int y){ return x+y; }", np.int32, 0x1003),
    ("float(float)", "float func(float x, float y){ return x > y ? x : y; }", np.float32, 0x2003),
]


def _data(dt, seed, n):
    if dt == np.int32:
        return oracle.gen_i32_below(seed, n, 1000)
    return oracle.gen_f32_sym(seed, n)


def _run(fn, r, out, x, n, stream=None):
    _lib.check(fn(r.handle, out.ptr, x.ptr, n, None, 0, stream))


@pytest.mark.parametrize("sig,src,dt,seed", CASES)
def test_noalloc_matches_reduce_run(skl, sig, src, dt, seed):
    L = _lib.lib()
    r = skl.Reduce(sig, src)
    for n in ((1 << 22) + 13, 100003):  # the second call is smaller
        print(f"noalloc reduce {sig} n={n}", flush=True)
        x = skl.DeviceBuffer.from_array(_data(dt, seed, n))
        o1 = skl.DeviceBuffer(16)
        o2 = skl.DeviceBuffer(16)
        _run(L.skl_reduce_run, r, o1, x, n)
        _run(L.skl_reduce_run_noalloc, r, o2, x, n)
        skl.synchronize()
        a = o1.to_numpy(dt, 1)
        b = o2.to_numpy(dt, 1)
        assert a.view(np.uint32)[0] == b.view(np.uint32)[0], (n, a, b)
        if dt == np.int32:
            assert a.view(np.uint32)[0] == oracle.reduce_add_u32_reforder(_data(dt, seed, n))


def test_noalloc_two_skeletons_share_the_scratch(skl):
    """Calls from two skeleton objects in a row on one stream use the same
    static array; each still gets its own correct result."""
    L = _lib.lib()
    rf = skl.Reduce(*CASES[0][:2])
    ri = skl.Reduce(*CASES[1][:2])
    xf = oracle.gen_f32_sym(0x2002, 77777)
    xi = oracle.gen_i32_below(0x1003, 55555, 1000)
    bf, bi = skl.DeviceBuffer.from_array(xf), skl.DeviceBuffer.from_array(xi)
    of, oi, ref_f = skl.DeviceBuffer(16), skl.DeviceBuffer(16), skl.DeviceBuffer(16)
    _run(L.skl_reduce_run_noalloc, rf, of, bf, len(xf))
    _run(L.skl_reduce_run_noalloc, ri, oi, bi, len(xi))
    _run(L.skl_reduce_run, rf, ref_f, bf, len(xf))
    skl.synchronize()
    assert of.to_numpy(np.float32, 1)[0] == ref_f.to_numpy(np.float32, 1)[0]
    assert oi.to_numpy(np.uint32, 1)[0] == oracle.reduce_add_u32_reforder(xi)


def test_noalloc_rejects_oversized_scratch(skl):
    """A launch shape with more work-group partials than the 64 KiB array
    holds is refused with a clear error (nothing runs)."""
    L = _lib.lib()
    r = skl.Reduce(*CASES[0][:2])
    shape = (C.c_int * 4)()
    _lib.check(L.skl_skeleton_launch_shape(r.handle, shape))
    shape[1] = 64  # 64 work-groups per CU: 16384 partials of 8 bytes
    _lib.check(L.skl_skeleton_set_launch_shape(r.handle, shape))
    n = 1 << 24
    x = skl.DeviceBuffer(n * 4)
    o = skl.DeviceBuffer(16)
    st = L.skl_reduce_run_noalloc(r.handle, o.ptr, x.ptr, n, None, 0, None)
    assert st == _lib.SKL_ERR_INVALID
    assert b"no-allocation scratch holds 65536" in L.skl_last_error()


def test_noalloc_graph_replay(skl):
    """The clear kernel + Reduce kernel pair captured in a graph gives the
    eager result on every replay, with eager work between replays."""
    L = _lib.lib()
    r = skl.Reduce(*CASES[0][:2])
    n = (1 << 20) + 5
    x = skl.DeviceBuffer.from_array(oracle.gen_f32_sym(0x2002, n))
    ref, o = skl.DeviceBuffer(16), skl.DeviceBuffer(16)
    _run(L.skl_reduce_run, r, ref, x, n)
    s = skl.Stream()
    _run(L.skl_reduce_run_noalloc, r, o, x, n, s.handle)  # warm: nothing left to set up
    s.synchronize()
    g = C.c_void_p()
    _lib.check(L.skl_graph_begin_capture(s.handle))
    _run(L.skl_reduce_run_noalloc, r, o, x, n, s.handle)
    _lib.check(L.skl_graph_end_capture(s.handle, C.byref(g)))
    try:
        want = ref.to_numpy(np.float32, 1)[0]
        for rep in range(4):
            o.fill(b"\0\0\xc0\x7f", 4, 4)  # NaN: a replay must overwrite it
            skl.synchronize()
            _lib.check(L.skl_graph_launch(g, s.handle))
            s.synchronize()
            assert o.to_numpy(np.float32, 1)[0] == want, rep
            _run(L.skl_reduce_run, r, ref, x, n)  # eager work between replays
            skl.synchronize()
    finally:
        L.skl_graph_destroy(g)

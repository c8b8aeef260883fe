"""The one-launch batch (skl_batch_run, an extension with no reference
counterpart): segment sums and the segmented scan of one array plus small
products in one kernel launch.  Segment results are checked against the
oracle's FP64 segmented sum and scan (within one float ulp), the products
bit for bit against the batched-product launch (skl_matmul_batched_f32) and
against FP64."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TILE = 24576  # elements per segment tile of a batch launch with products
TILE_SEG = 32768  # ... of a launch of segments alone
ADD = "float func(float x, float y){ return x+y; }"


def _fixture():
    with open(os.path.join(os.path.dirname(__file__), "golden", "batched_shapes.json")) as f:
        return json.load(f)


def _products(skl, shapes, seed):
    """Device descriptors of products with seeded operands, and FP64 references."""
    from skelcl_amd import _lib
    descs = (_lib.skl_gemm_desc * len(shapes))()
    keep, refs = [], []
    for i, (M, N, K) in enumerate(shapes):
        A = oracle.gen_f32_sym(seed + 2 * i, max(M * K, 1))[:M * K].reshape(M, K)
        B = oracle.gen_f32_sym(seed + 2 * i + 1, max(K * N, 1))[:K * N].reshape(K, N)
        dA = skl.DeviceBuffer.from_array(A) if M * K else skl.DeviceBuffer(16)
        dB = skl.DeviceBuffer.from_array(B) if K * N else skl.DeviceBuffer(16)
        dC = skl.DeviceBuffer(max(4 * M * N, 16))
        keep += [dA, dB, dC]
        descs[i].A, descs[i].B, descs[i].C = dA.value, dB.value, dC.value
        descs[i].M, descs[i].N, descs[i].K = M, N, K
        refs.append((dC, M, N, A.astype(np.float64) @ B.astype(np.float64)))
    dd = skl.DeviceBuffer(C.sizeof(descs))
    _lib.check(_lib.lib().skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs),
                                         None, 1))
    return dd, keep, refs


def _run(skl, red, sums, scan, out, inclusive, X, n, O, nseg, dd=None, count=0, prec=1,
         workers=0):
    from skelcl_amd import _lib
    _lib.check(_lib.lib().skl_batch_run(
        red.handle if red else None, sums.ptr if sums else None,
        scan.handle if scan else None, out.ptr if out else None, int(inclusive),
        X.ptr if X else None, n, O.ptr if O else None, nseg, dd.ptr if dd else None, count,
        prec, workers, None))


def _check_sums(got, x, off):
    ref = oracle.segmented_sum_f64(x, np.asarray(off, np.int64))
    got = np.asarray(got, np.float64)
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref).astype(np.float32)) + 1e-12)


def _check_scan(got, x, off, inclusive, n, fill=7.0):
    off = np.asarray(off, np.int64)
    ref = oracle.segmented_scan_f64(x, off, inclusive)
    inside = np.zeros(n, bool)
    inside[off[0]:off[-1]] = True
    tol = np.spacing(np.abs(ref[inside]).astype(np.float32)).astype(np.float64) + 1e-9
    assert np.all(np.abs(got[inside].astype(np.float64) - ref[inside]) <= tol)
    assert np.all(got[~inside] == fill)


@pytest.mark.parametrize("seg_prec", ["DEFAULT", "FP64"])
@pytest.mark.parametrize("mm_prec,workers", [(1, 1), (2, 1), (2, 2)])
def test_batch_segments_and_products(skl, seg_prec, mm_prec, workers):
    """The batched workload's segment lengths (first 256 of the committed
    fixture) and products (first 40): sums, inclusive scan and products from
    one launch."""
    from skelcl_amd import _lib
    fx = _fixture()
    lens = np.array(fx["segment_lengths"][:256], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    x = oracle.gen_f32_sym(0x5002, n)
    P = getattr(skl, "PREC_" + seg_prec)
    red = skl.Reduce("float(float)", ADD, precision=P)
    scan = skl.Scan("float(float)", ADD, "0.0f", precision=P)
    X, O = skl.DeviceBuffer.from_array(x), skl.DeviceBuffer.from_array(off)
    S = skl.DeviceBuffer.from_array(np.full(256, 7.0, np.float32))
    Y = skl.DeviceBuffer.from_array(np.full(n, 7.0, np.float32))
    shapes = fx["gemm_shapes"][:40]
    dd, keep, refs = _products(skl, shapes, 0x5004)
    _run(skl, red, S, scan, Y, True, X, n, O, 256, dd, len(shapes), mm_prec, workers)
    _check_sums(S.to_numpy(np.float32, 256), x, off)
    _check_scan(Y.to_numpy(np.float32, n), x, off, True, n)
    got = [dC.to_numpy(np.float32, M * N).copy() for dC, M, N, _ in refs]
    for g, (dC, M, N, ref) in zip(got, refs):
        assert np.linalg.norm(g.reshape(M, N) - ref) / np.linalg.norm(ref) < 1e-6
    # the same products from the batched-product launch: the same bits
    _lib.check(_lib.lib().skl_matmul_batched_f32(dd.ptr, len(shapes), mm_prec, None))
    for g, (dC, M, N, _) in zip(got, refs):
        assert np.array_equal(g.view(np.uint32), dC.to_numpy(np.float32, M * N).view(np.uint32))
    bad = C.c_int(-1)
    _lib.check(_lib.lib().skl_batch_check(scan.handle, C.byref(bad)))
    assert bad.value == 0


@pytest.mark.parametrize("case", ["inner_range", "dense_tiny", "tile_edges", "ragged_tail",
                                  "empty_runs"])
@pytest.mark.parametrize("inclusive", [False, True])
@pytest.mark.parametrize("alone", [True, False])
def test_batch_segment_layouts(skl, case, inclusive, alone):
    """Segment layouts against the tiles (32768 elements in a launch of
    segments alone, 24576 beside products): uncovered ends (outputs left as
    they are), runs of empty and one-element segments (more segment ends per
    tile than the LDS slots hold: the rest are found by searching the
    offsets), heads on, before and after tile starts, a last tile cut short,
    empty segments at both ends of the array."""
    rng = np.random.default_rng({"inner_range": 1, "dense_tiny": 2, "tile_edges": 3,
                                 "ragged_tail": 4, "empty_runs": 5}[case])
    T = TILE_SEG if alone else TILE
    if case == "inner_range":
        n = 5 * T + 123
        off = [1000, 1000, 20000, 20001, 60000, 4 * T + 5]
    elif case == "dense_tiny":
        lens = rng.integers(0, 4, 30000)
        off = np.concatenate([[0], np.cumsum(lens)]).tolist()
        n = off[-1] + 50
    elif case == "tile_edges":
        off = [0, T - 1, T, T + 1, 2 * T, 3 * T, 3 * T, 5 * T - 1, 7 * T, 7 * T + 3]
        n = off[-1]
    elif case == "ragged_tail":
        n = 9 * T + 7
        lens = rng.integers(1, 40000, 12)
        off = np.concatenate([[0], np.minimum(np.cumsum(lens), n)]).tolist()
        off[-1] = n
    else:
        n = 3 * T
        off = [0, 0, 0, 5, T, T, T, 2 * T + 9, n, n, n]
    nseg = len(off) - 1
    x = oracle.gen_f32_sym(0x5010, n)
    red = skl.Reduce("float(float)", ADD)
    scan = skl.Scan("float(float)", ADD, "0.0f")
    X, O = skl.DeviceBuffer.from_array(x), skl.DeviceBuffer.from_array(np.array(off, np.int64))
    S = skl.DeviceBuffer.from_array(np.full(nseg, 7.0, np.float32))
    Y = skl.DeviceBuffer.from_array(np.full(n, 7.0, np.float32))
    dd, count = None, 0
    if not alone:
        dd, keep, refs = _products(skl, [(32, 48, 16)], 0x5400)
        count = 1
    _run(skl, red, S, scan, Y, inclusive, X, n, O, nseg, dd, count, 2, 1)
    _check_sums(S.to_numpy(np.float32, nseg), x, off)
    _check_scan(Y.to_numpy(np.float32, n), x, off, inclusive, n)


def test_batch_parts_alone_and_repeated(skl):
    """Sums only, scan only and products only, and alternating sizes on the
    same skeletons (the ticket and status words carry over between calls)."""
    fx = _fixture()
    red = skl.Reduce("float(float)", ADD)
    scan = skl.Scan("float(float)", ADD, "0.0f")
    for it, n in enumerate([300000, 40000, 17, 1 << 20, 0]):
        lens = oracle.gen_loguniform(0x5020 + it, 40, 1.0, 4.0)
        off = np.minimum(np.concatenate([[0], np.cumsum(lens)]), n).astype(np.int64)
        off[-1] = n
        x = oracle.gen_f32_sym(0x5030 + it, max(n, 1))[:n]
        X = skl.DeviceBuffer.from_array(x) if n else None
        O = skl.DeviceBuffer.from_array(off)
        S = skl.DeviceBuffer.from_array(np.full(40, 7.0, np.float32))
        _run(skl, red, S, None, None, False, X, n, O, 40)
        _check_sums(S.to_numpy(np.float32, 40), x, off)
        if n:
            Y = skl.DeviceBuffer.from_array(np.full(n, 7.0, np.float32))
            _run(skl, None, None, scan, Y, it % 2 == 0, X, n, O, 40)
            _check_scan(Y.to_numpy(np.float32, n), x, off, it % 2 == 0, n)
    shapes = [(129, 260, 36), (3, 130, 7), (1, 4, 4), (200, 8, 64), (64, 64, 1),
              (130, 132, 100), (0, 8, 8), (257, 12, 4)] + fx["gemm_shapes"][:8]
    for prec in (1, 2):
        dd, keep, refs = _products(skl, shapes, 0x5200)
        _run(skl, None, None, None, None, False, None, 0, None, 0, dd, len(shapes), prec, 2)
        for dC, M, N, ref in refs:
            if M * N:
                c = dC.to_numpy(np.float32, M * N).reshape(M, N)
                assert np.linalg.norm(c - ref) / max(np.linalg.norm(ref), 1e-30) < 1e-6, (M, N)


def test_batch_invalid_offsets_reported(skl):
    """Device offsets checked by the launch's tiles: decreasing, past n and
    negative offsets are reported by skl_batch_check; a valid call on the
    same skeletons afterwards is correct and reports nothing."""
    from skelcl_amd import _lib
    n = 64
    x = oracle.gen_f32_sym(0x5007, n)
    X = skl.DeviceBuffer.from_array(x)
    red = skl.Reduce("float(float)", ADD)
    scan = skl.Scan("float(float)", ADD, "0.0f")
    bad = C.c_int(-1)
    for off in ([0, 40, 20, 64], [0, 10, 65], [-1, 10, 20]):
        O = skl.DeviceBuffer.from_array(np.array(off, np.int64))
        S = skl.DeviceBuffer(4 * (len(off) - 1))
        Y = skl.DeviceBuffer(4 * n)
        _run(skl, red, S, scan, Y, True, X, n, O, len(off) - 1)
        _lib.check(_lib.lib().skl_batch_check(scan.handle, C.byref(bad)))
        assert bad.value == 1, off
    off = np.array([0, 30, 64], np.int64)
    O = skl.DeviceBuffer.from_array(off)
    S = skl.DeviceBuffer(8)
    Y = skl.DeviceBuffer.from_array(np.full(n, 7.0, np.float32))
    _run(skl, red, S, scan, Y, False, X, n, O, 2)
    _lib.check(_lib.lib().skl_batch_check(scan.handle, C.byref(bad)))
    assert bad.value == 0
    _check_sums(S.to_numpy(np.float32, 2), x, off)
    _check_scan(Y.to_numpy(np.float32, n), x, off, False, n)


def test_batch_graph_replays(skl):
    """The one-launch batch captured into a graph and replayed four times,
    with the outputs poisoned and eager work between replays: every replay
    equals the eager launch bit for bit."""
    from skelcl_amd import _lib
    L = _lib.lib()
    fx = _fixture()
    lens = np.array(fx["segment_lengths"][:64], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    x = oracle.gen_f32_sym(0x5002, n)
    red = skl.Reduce("float(float)", ADD)
    scan = skl.Scan("float(float)", ADD, "0.0f")
    X, O = skl.DeviceBuffer.from_array(x), skl.DeviceBuffer.from_array(off)
    S, Y = skl.DeviceBuffer(4 * 64), skl.DeviceBuffer(4 * n)
    shapes = fx["gemm_shapes"][:16]
    dd, keep, refs = _products(skl, shapes, 0x5004)
    st = skl.Stream()

    def launch():
        _lib.check(L.skl_batch_run(red.handle, S.ptr, scan.handle, Y.ptr, 1, X.ptr, n, O.ptr, 64,
                                   dd.ptr, len(shapes), 2, 1, st.handle))

    launch()
    st.synchronize()
    eager = (S.to_numpy(np.float32, 64).copy(), Y.to_numpy(np.float32, n).copy(),
             [dC.to_numpy(np.float32, M * N).copy() for dC, M, N, _ in refs])
    _check_sums(eager[0], x, off)
    g = C.c_void_p()
    _lib.check(L.skl_graph_begin_capture(st.handle))
    launch()
    _lib.check(L.skl_graph_end_capture(st.handle, C.byref(g)))
    other = skl.Scan("float(float)", ADD, "0.0f")
    try:
        for rep in range(4):
            S.upload(np.full(64, np.nan, np.float32))
            Y.upload(np.full(n, np.nan, np.float32))
            _ = np.asarray(other.segmented(skl.Vector(x[:5000]), [0, 100, 5000]))
            _lib.check(L.skl_graph_launch(g, st.handle))
            st.synchronize()
            assert np.array_equal(S.to_numpy(np.float32, 64).view(np.uint32),
                                  eager[0].view(np.uint32)), rep
            assert np.array_equal(Y.to_numpy(np.float32, n).view(np.uint32),
                                  eager[1].view(np.uint32)), rep
            for c0, (dC, M, N, _) in zip(eager[2], refs):
                assert np.array_equal(dC.to_numpy(np.float32, M * N).view(np.uint32),
                                      c0.view(np.uint32)), rep
    finally:
        _lib.check(L.skl_graph_destroy(g))


@pytest.mark.parametrize("case", ["fixture", "dense_tiny", "tile_edges", "wrap"])
@pytest.mark.parametrize("with_products", [False, True])
def test_batch_int_segments_exact(skl, case, with_products):
    """int Reduce(+) / Scan(+) segments (exact modulo 2^32) against the
    oracle, bit for bit: the batched workload's lengths (first 64), runs of
    empty and one-element segments, heads at the tile edges of both tile
    sizes (24576 beside products, 32768 alone), and values whose sums wrap."""
    from skelcl_amd import _lib
    rng = np.random.default_rng({"fixture": 11, "dense_tiny": 12, "tile_edges": 13,
                                 "wrap": 14}[case])
    if case == "fixture":
        lens = np.array(_fixture()["segment_lengths"][:64], np.int64)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        n = int(off[-1])
        x = oracle.gen_i32_below(0x5101, n, 1000)
    elif case == "dense_tiny":
        lens = rng.integers(0, 4, 20000)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        n = int(off[-1]) + 9
        x = oracle.gen_i32_below(0x5102, n, 1000)
    elif case == "tile_edges":
        T, U = TILE, 2 * TILE
        off = np.array([0, T - 1, T, T + 1, U - 1, U, U + 1, 3 * T, 4 * T, 5 * T + 3], np.int64)
        n = int(off[-1]) + 100
        x = oracle.gen_i32_below(0x5103, n, 1000)
    else:
        off = np.array([0, 3, 3, 70000, 70001, 150000], np.int64)
        n = 150000
        x = oracle.gen_u64(0x5104, n).astype(np.uint32).view(np.int32)
    nseg = len(off) - 1
    red = skl.Reduce("int(int)", "int func(int x, int y){ return x+y; }")
    scan = skl.Scan("int(int)", "int func(int x, int y){ return x+y; }", "0")
    X, O = skl.DeviceBuffer.from_array(x), skl.DeviceBuffer.from_array(off)
    dd, count = None, 0
    if with_products:
        dd, keep, refs = _products(skl, [(64, 64, 64), (130, 7, 33)], 0x5300)
        count = 2
    for incl in (False, True):
        S = skl.DeviceBuffer.from_array(np.full(nseg, 0x5EED, np.int32))
        Y = skl.DeviceBuffer.from_array(np.full(n, 0x5EED, np.int32))
        _run(skl, red, S, scan, Y, incl, X, n, O, nseg, dd, count, 2, 1)
        assert np.array_equal(S.to_numpy(np.uint32, nseg), oracle.segmented_sum_u32(x, off))
        assert np.array_equal(Y.to_numpy(np.uint32, n),
                              oracle.segmented_scan_u32(x, off, incl, fill=0x5EED))
    bad = C.c_int(-1)
    _lib.check(_lib.lib().skl_batch_check(scan.handle, C.byref(bad)))
    assert bad.value == 0
    if with_products:
        for dC, M, N, ref in refs:
            c = dC.to_numpy(np.float32, M * N).reshape(M, N)
            assert np.linalg.norm(c - ref) / np.linalg.norm(ref) < 1e-6


def test_batch_rejects_mixed_element_types(skl):
    """A float Reduce with an int Scan, or a Reduce(*), is refused."""
    from skelcl_amd import _lib
    x = skl.DeviceBuffer.from_array(np.zeros(16, np.float32))
    O = skl.DeviceBuffer.from_array(np.array([0, 16], np.int64))
    S, Y = skl.DeviceBuffer(4), skl.DeviceBuffer(64)
    rf = skl.Reduce("float(float)", ADD)
    si = skl.Scan("int(int)", "int func(int x, int y){ return x+y; }", "0")
    rm = skl.Reduce("float(float)", "float func(float x, float y){ return x*y; }")
    L = _lib.lib()
    assert L.skl_batch_run(rf.handle, S.ptr, si.handle, Y.ptr, 1, x.ptr, 16, O.ptr, 1, None, 0,
                           1, 0, None) == _lib.SKL_ERR_INVALID
    assert L.skl_batch_run(rm.handle, S.ptr, None, None, 1, x.ptr, 16, O.ptr, 1, None, 0, 1, 0,
                           None) == _lib.SKL_ERR_UNSUPPORTED


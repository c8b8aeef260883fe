#!/usr/bin/env python3
"""Headline benchmark: the bandwidth skeletons at scale on one MI355X.

One step = one pass of the three skeletons of the bandwidth workload, each
over N = 2^28 elements already resident in HBM (synthetic SplitMix64 inputs
generated on the GPU):
  * inclusive Scan(+) on int32 uniform[0,1000)      8 B/elem algorithmic
  * Reduce(+) on float32 uniform[-1,1) (compensated  4 B/elem
    per-block partials, FP64 cross-block combine)
  * saxpy a=2.5 on float32 (Zip form of            12 B/elem
    examples/saxpy/main.cpp:77-88)
value = total algorithmic bytes / time (GB/s, 1e9 B/s).  The kernels run on
one stream; HIP events on that stream time every launch.

Prints ONE JSON line.  `--gpus N` with N > 1 exits: the benchmark drives one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 1 << 28
SEED_SCAN, SEED_RED, SEED_X, SEED_Y = 0x2001, 0x2002, 0x2003, 0x2004
SAXPY_A = 2.5
# the 11x11 weighted stencil of the MapOverlap configuration
GAUSS_SRC = """float func(input_matrix_t img, __global float* w) {
  float sum = 0.0f;
  for (int dy = -5; dy <= 5; ++dy)
    for (int dx = -5; dx <= 5; ++dx)
      sum += getData(img, dx, dy) * w[(dy + 5) * 11 + (dx + 5)];
  return sum;
}"""


def progress(msg: str) -> None:
    """One line per finished stage on stderr (the JSON result goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def host_cpus():
    """(threads the CPU baseline uses, CPUs this process may run on, CPU
    model).  The threads are the process's CPU share: OMP_NUM_THREADS when
    set (the share a one-GPU machine grants), else every CPU in its
    affinity mask."""
    avail = len(os.sched_getaffinity(0))
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return (min(share, avail) if share > 0 else avail), avail, model


def cpu_baseline(nthreads: int, reps: int = 5):
    """The oracle (reference-order CPU restatement) on the host cores: the
    same workload at full size, each skeleton run once untimed and then
    `reps` times; the median per skeleton."""
    import oracle
    t0 = time.perf_counter()
    xs = oracle.gen_i32_below(SEED_SCAN, N, 1000)
    xr = oracle.gen_f32_sym(SEED_RED, N)
    x = oracle.gen_f32_sym(SEED_X, N)
    y = oracle.gen_f32_sym(SEED_Y, N)
    tgen = time.perf_counter() - t0

    def timed(fn):
        out = fn()
        ts = []
        for _ in range(reps):
            t1 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t1)
        return statistics.median(ts), out

    t = {}
    t["scan"], ex = timed(lambda: oracle.scan_excl_add_u32_reforder(xs, 256, nthreads))
    inc = ex + xs.view(np.uint32)  # inclusive from the reference's exclusive form
    t["reduce"], red = timed(lambda: oracle.reduce_add_f32_reforder(xr, 256, nthreads))
    t["saxpy"], sax = timed(lambda: oracle.saxpy_f32(SAXPY_A, x, y, nthreads))
    ref64 = oracle.sum_f64(xr)
    total = sum(t.values())
    return {
        "times_s": t, "gen_s": tgen, "total_s": total, "reduce_ref": float(red),
        "reduce_fp64": ref64, "scan_last": int(inc[-1]), "saxpy_head": sax[:4].tolist(),
        "xr": xr,
    }


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    counter summary (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled by
    what the counter reads for a copy kernel of known size); None if absent."""
    path = os.path.join(ROOT, "profiles", "bandwidth_n2e28_hbm_counters.md")
    try:
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic" and w[1] == kernel:
                    return int(w[2])
    except OSError:
        pass
    return None


# device time spent in back-to-back calls before each extras measurement
SETTLE_MS = 100.0


def _time(stream, fn, reps=20, warm=3, settle_ms=SETTLE_MS):
    """Median of `reps` HIP-event timings (ms) of fn() on `stream`, after
    `warm` calls and then back-to-back calls for at least `settle_ms` of
    device time: under a sudden compute load the clocks dip for several
    milliseconds before they settle (the 8192^2 stencil forward runs
    0.24-0.29 ms during the first 20-40 launches and 0.21 ms afterwards), and
    the figures here are for a device in steady use."""
    import skelcl_amd as skelcl
    for _ in range(warm):
        fn()
    done, batch = 0.0, 1
    while done < settle_ms:
        a, b = skelcl.Event(), skelcl.Event()
        a.record(stream)
        for _ in range(batch):
            fn()
        b.record(stream)
        b.synchronize()
        done += a.elapsed_ms(b)
        batch = min(batch * 2, 32)  # few launches queued at a time
    ts = []
    for _ in range(reps):
        a, b = skelcl.Event(), skelcl.Event()
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        ts.append(a.elapsed_ms(b))
    return statistics.median(ts)


# Plain streaming probes for the practical HBM ceiling: grid-stride loops of
# 16-byte non-temporal accesses, 4 in flight per thread, integer work only
# (the read probe folds its loads with XOR and almost never stores).
CEILING_SRC = r"""
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
extern "C" __global__ __launch_bounds__(256) void probe_read(const u4* __restrict__ in,
    unsigned long long n, unsigned int* __restrict__ sink) {
  const unsigned long long G = (unsigned long long)gridDim.x * 256;
  unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  u4 acc = {0u, 0u, 0u, 0u};
  for (; i + 3 * G < n; i += 4 * G) {
    const u4 a = __builtin_nontemporal_load(in + i), b = __builtin_nontemporal_load(in + i + G);
    const u4 c = __builtin_nontemporal_load(in + i + 2 * G);
    const u4 d = __builtin_nontemporal_load(in + i + 3 * G);
    acc ^= (a ^ b) ^ (c ^ d);
  }
  for (; i < n; i += G) acc ^= __builtin_nontemporal_load(in + i);
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u) sink[0] = 1u;
}
extern "C" __global__ __launch_bounds__(256) void probe_write(u4* __restrict__ out,
    unsigned long long n) {
  const unsigned long long G = (unsigned long long)gridDim.x * 256;
  unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  const u4 v = {(unsigned)i, 1u, 2u, 3u};
  for (; i + 3 * G < n; i += 4 * G) {
    __builtin_nontemporal_store(v, out + i);
    __builtin_nontemporal_store(v, out + i + G);
    __builtin_nontemporal_store(v, out + i + 2 * G);
    __builtin_nontemporal_store(v, out + i + 3 * G);
  }
  for (; i < n; i += G) __builtin_nontemporal_store(v, out + i);
}
extern "C" __global__ __launch_bounds__(256) void probe_copy(const u4* __restrict__ in,
    u4* __restrict__ out, unsigned long long n) {
  const unsigned long long G = (unsigned long long)gridDim.x * 256;
  unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * G < n; i += 4 * G) {
    const u4 a = __builtin_nontemporal_load(in + i), b = __builtin_nontemporal_load(in + i + G);
    const u4 c = __builtin_nontemporal_load(in + i + 2 * G);
    const u4 d = __builtin_nontemporal_load(in + i + 3 * G);
    __builtin_nontemporal_store(a, out + i);
    __builtin_nontemporal_store(b, out + i + G);
    __builtin_nontemporal_store(c, out + i + 2 * G);
    __builtin_nontemporal_store(d, out + i + 3 * G);
  }
  for (; i < n; i += G) __builtin_nontemporal_store(__builtin_nontemporal_load(in + i), out + i);
}
extern "C" __global__ __launch_bounds__(256) void probe_zip(const u4* __restrict__ a,
    const u4* __restrict__ b, u4* __restrict__ out, unsigned long long n) {
  const unsigned long long G = (unsigned long long)gridDim.x * 256;
  unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * G < n; i += 4 * G) {
    const u4 a0 = __builtin_nontemporal_load(a + i), a1 = __builtin_nontemporal_load(a + i + G);
    const u4 a2 = __builtin_nontemporal_load(a + i + 2 * G);
    const u4 a3 = __builtin_nontemporal_load(a + i + 3 * G);
    const u4 b0 = __builtin_nontemporal_load(b + i), b1 = __builtin_nontemporal_load(b + i + G);
    const u4 b2 = __builtin_nontemporal_load(b + i + 2 * G);
    const u4 b3 = __builtin_nontemporal_load(b + i + 3 * G);
    __builtin_nontemporal_store(a0 ^ b0, out + i);
    __builtin_nontemporal_store(a1 ^ b1, out + i + G);
    __builtin_nontemporal_store(a2 ^ b2, out + i + 2 * G);
    __builtin_nontemporal_store(a3 ^ b3, out + i + 3 * G);
  }
  for (; i < n; i += G)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i) ^ __builtin_nontemporal_load(b + i),
                                out + i);
}
"""


# LDS read probes: conflict-free 16- and 8-byte reads (ds_read_b128 /
# ds_read_b64) of a 32 KiB buffer in a loop, 8 reads in flight per lane; the
# ceiling the LDS-fed kernels (MFMA products, stencil) are compared with.
LDS_SRC = r"""
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
typedef unsigned int u2 __attribute__((ext_vector_type(2)));
extern "C" __global__ __launch_bounds__(512) void lds_read_b128(unsigned* sink, int iters) {
  __shared__ u4 buf[2048];
  const int t = threadIdx.x;
  for (int i = t; i < 2048; i += blockDim.x) buf[i] = u4{(unsigned)i, 1u, 2u, 3u};
  __syncthreads();
  u4 acc = {0u, 0u, 0u, 0u};
  int idx = t & 2047;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc ^= buf[(idx + k * 64) & 2047];
    idx = (idx + 512) & 2047;
  }
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u) sink[0] = 1u;
}
extern "C" __global__ __launch_bounds__(512) void lds_read_b64(unsigned* sink, int iters) {
  __shared__ u2 buf[4096];
  const int t = threadIdx.x;
  for (int i = t; i < 4096; i += blockDim.x) buf[i] = u2{(unsigned)i, 1u};
  __syncthreads();
  u2 acc = {0u, 0u};
  int idx = t & 4095;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc ^= buf[(idx + k * 64) & 4095];
    idx = (idx + 512) & 4095;
  }
  if ((acc.x ^ acc.y) == 0x9E3779B9u) sink[0] = 1u;
}
"""


# FP32 FMA issue probes: 16 independent accumulators per lane, fused
# multiply-adds whose operands are all vector registers, or whose multiplier
# is wave-uniform (a scalar register), in a loop the compiler cannot fold.
VALU_SRC = r"""
extern "C" __global__ __launch_bounds__(256) void fma_vv(float* sink, int iters, float m) {
  float acc[16], x = (float)threadIdx.x * 1e-3f, y = m + (float)(threadIdx.x & 7) * 1e-4f;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = x + (float)i;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = __builtin_fmaf(acc[i], y, x);
  }
  float t = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) t += acc[i];
  if (t == 1234.5f) sink[0] = t;
}
extern "C" __global__ __launch_bounds__(256) void fma_sv(float* sink, int iters, float m) {
  float acc[16], x = (float)threadIdx.x * 1e-3f;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = x + (float)i;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = __builtin_fmaf(acc[i], m, x);
  }
  float t = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) t += acc[i];
  if (t == 1234.5f) sink[0] = t;
}
"""


def run_valu_ceiling(skelcl, L, ck, stream):
    """FP32 FMA rate with vector-only operands and with a scalar multiplier."""
    import ctypes as C
    st = stream.handle
    prog, log = C.c_void_p(), C.c_void_p()
    ck(L.skl_program_build(VALU_SRC.encode(), b"-ffp-contract=off", C.byref(prog), C.byref(log)))
    skelcl._lib.take_string(log)
    cus = skelcl.device_info()["compute_units"]
    sink = skelcl.DeviceBuffer(256)
    sp, iters, m = C.c_void_p(sink.ptr.value), C.c_int(2048), C.c_float(0.999)
    args = (C.c_void_p * 3)(*[C.cast(C.pointer(v), C.c_void_p).value for v in (sp, iters, m)])
    out = {}
    for name in ("fma_vv", "fma_sv"):
        k = C.c_void_p()
        ck(L.skl_kernel_get(prog, name.encode(), C.byref(k)))
        rows = []
        for per_cu in (4, 8, 16):
            grid = (C.c_uint * 3)(cus * per_cu, 1, 1)
            block = (C.c_uint * 3)(256, 1, 1)
            t = _time(stream, lambda: ck(L.skl_kernel_launch(k, grid, block, 0, args, st)), reps=10)
            flop = 2.0 * cus * per_cu * 256 * iters.value * 8 * 16
            rows.append({"work_groups_per_cu": per_cu, "ms": round(t, 4),
                         "TFLOPs": round(flop / (t * 1e-3) / 1e12, 1)})
        out[name] = {"sweep": rows, "best_TFLOPs": max(r["TFLOPs"] for r in rows)}
        progress(f"{name} done")
    ck(L.skl_program_release(prog))
    return out


def run_lds_ceiling(skelcl, L, ck, stream):
    """LDS read bandwidth per CU at a few work-group counts and sizes."""
    import ctypes as C
    st = stream.handle
    prog, log = C.c_void_p(), C.c_void_p()
    ck(L.skl_program_build(LDS_SRC.encode(), b"", C.byref(prog), C.byref(log)))
    skelcl._lib.take_string(log)
    cus = skelcl.device_info()["compute_units"]
    sink = skelcl.DeviceBuffer(256)
    sp, iters = C.c_void_p(sink.ptr.value), C.c_int(4096)
    args = (C.c_void_p * 2)(C.cast(C.pointer(sp), C.c_void_p).value,
                            C.cast(C.pointer(iters), C.c_void_p).value)
    out = {}
    for name, width in (("lds_read_b128", 16), ("lds_read_b64", 8)):
        k = C.c_void_p()
        ck(L.skl_kernel_get(prog, name.encode(), C.byref(k)))
        rows = []
        for per_cu, threads in ((1, 256), (1, 512), (2, 512), (4, 256)):
            grid = (C.c_uint * 3)(cus * per_cu, 1, 1)
            block = (C.c_uint * 3)(threads, 1, 1)
            t = _time(stream, lambda: ck(L.skl_kernel_launch(k, grid, block, 0, args, st)), reps=10)
            moved = cus * per_cu * threads * iters.value * 8 * width
            rows.append({"work_groups_per_cu": per_cu, "threads": threads, "ms": round(t, 4),
                         "bytes_per_ns_per_cu": round(moved / (t * 1e6) / cus, 1)})
        out[name] = {"sweep": rows,
                     "best_bytes_per_ns_per_cu": max(r["bytes_per_ns_per_cu"] for r in rows)}
        progress(f"{name} done")
    ck(L.skl_program_release(prog))
    return out


def run_ceiling(skelcl, L, ck, stream):
    """Read-only, write-only, copy and two-read-one-write (the saxpy pattern)
    streams over 1 GiB per array at a sweep of grid sizes (work-groups of 256
    threads per CU): the practical HBM ceiling the skeleton legs are
    compared with."""
    import ctypes as C
    st = stream.handle
    prog, log = C.c_void_p(), C.c_void_p()
    ck(L.skl_program_build(CEILING_SRC.encode(), b"", C.byref(prog), C.byref(log)))
    skelcl._lib.take_string(log)
    kern = {}
    for name in ("probe_read", "probe_write", "probe_copy", "probe_zip"):
        k = C.c_void_p()
        ck(L.skl_kernel_get(prog, name.encode(), C.byref(k)))
        kern[name] = k
    nbytes = 1 << 30
    n4 = nbytes // 16
    src, src2, dst, sink = (skelcl.DeviceBuffer(nbytes), skelcl.DeviceBuffer(nbytes),
                            skelcl.DeviceBuffer(nbytes), skelcl.DeviceBuffer(256))
    ck(L.skl_generate_f32_sym(src.ptr, nbytes // 4, 0x6001, st))
    ck(L.skl_generate_f32_sym(src2.ptr, nbytes // 4, 0x6002, st))
    cus = skelcl.device_info()["compute_units"]
    n_arg = C.c_ulonglong(n4)
    sp, dp, kp = C.c_void_p(src.ptr.value), C.c_void_p(dst.ptr.value), C.c_void_p(sink.ptr.value)
    sp2 = C.c_void_p(src2.ptr.value)
    argv = {"probe_read": ([C.cast(C.pointer(sp), C.c_void_p), C.cast(C.pointer(n_arg), C.c_void_p),
                            C.cast(C.pointer(kp), C.c_void_p)], nbytes),
            "probe_write": ([C.cast(C.pointer(dp), C.c_void_p),
                             C.cast(C.pointer(n_arg), C.c_void_p)], nbytes),
            "probe_copy": ([C.cast(C.pointer(sp), C.c_void_p), C.cast(C.pointer(dp), C.c_void_p),
                            C.cast(C.pointer(n_arg), C.c_void_p)], 2 * nbytes),
            "probe_zip": ([C.cast(C.pointer(sp), C.c_void_p), C.cast(C.pointer(sp2), C.c_void_p),
                           C.cast(C.pointer(dp), C.c_void_p), C.cast(C.pointer(n_arg), C.c_void_p)],
                          3 * nbytes)}
    out = {}
    for name, (a, moved) in argv.items():
        args = (C.c_void_p * len(a))(*[x.value for x in a])
        rows = []
        for per_cu in (1, 1.5, 2, 3, 4, 8, 16, 64):
            grid = (C.c_uint * 3)(int(cus * per_cu), 1, 1)
            block = (C.c_uint * 3)(256, 1, 1)
            t = _time(stream, lambda: ck(L.skl_kernel_launch(kern[name], grid, block, 0, args, st)))
            rows.append({"work_groups_per_cu": per_cu, "ms": round(t, 4),
                         "GBps": round(moved / (t * 1e-3) / 1e9, 1)})
        best = max(rows, key=lambda r: r["GBps"])
        out[name] = {"bytes": moved, "sweep": rows, "best_GBps": best["GBps"],
                     "best_frac_of_8TBps": round(best["GBps"] / HBM_PEAK_GBS, 4)}
        progress(f"{name} done")
    ck(L.skl_program_release(prog))
    return out


def run_extras(skelcl, L, ck, stream, args):
    """The other configurations, one measurement each (not part of `value`)."""
    import ctypes as C
    import oracle
    st = stream.handle
    ex = {}
    only = set(args.only.split(",")) if args.only else None

    def want(name):
        return only is None or name in only

    add = skelcl.Reduce("float(float)", "float func(float x, float y){ return x+y; }", "0")

    # ---- configuration 1: dot product N=2^20 (Zip x*y then Reduce +), int Reduce
    if want("dot"):
        n1 = 1 << 20
        x1, y1 = oracle.gen_f32_sym(0x1001, n1), oracle.gen_f32_sym(0x1002, n1)
        X1, Y1 = skelcl.DeviceBuffer.from_array(x1), skelcl.DeviceBuffer.from_array(y1)
        T1, o1 = skelcl.DeviceBuffer(4 * n1), skelcl.DeviceBuffer(16)
        mult = skelcl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
        addi = skelcl.Reduce("int(int)", "int func(int x, int y){ return x+y; }", "0")
        k1 = oracle.gen_i32_below(0x1003, n1, 1000)
        K1 = skelcl.DeviceBuffer.from_array(k1)

        def dot_unfused():
            ck(L.skl_zip_run(mult.handle, T1.ptr, X1.ptr, Y1.ptr, n1, None, 0, st))
            ck(L.skl_reduce_run(add.handle, o1.ptr, T1.ptr, n1, None, 0, st))

        def dot_fused():
            ck(L.skl_zip_reduce_run(mult.handle, add.handle, o1.ptr, X1.ptr, Y1.ptr, n1, st))

        t_unf = _time(stream, dot_unfused, reps=20)
        got_unf = float(o1.to_numpy(np.float32, 1)[0])
        t_fus = _time(stream, dot_fused, reps=20)
        got_fus = float(o1.to_numpy(np.float32, 1)[0])
        # graph-captured fused pipeline
        g = C.c_void_p()
        ck(L.skl_graph_begin_capture(st))
        dot_fused()
        ck(L.skl_graph_end_capture(st, C.byref(g)))
        t_graph = _time(stream, lambda: ck(L.skl_graph_launch(g, st)), reps=20, settle_ms=0)
        got_graph = float(o1.to_numpy(np.float32, 1)[0])
        ck(L.skl_graph_destroy(g))
        t_int = _time(stream, lambda: ck(L.skl_reduce_run(addi.handle, o1.ptr, K1.ptr, n1, None,
                                                          0, st)), reps=20)
        got_int = int(o1.to_numpy(np.int32, 1)[0])
        ref = oracle.dot_f64(x1, y1)
        ex["dot_2^20"] = {
            "unfused_ms": round(t_unf, 4), "fused_ms": round(t_fus, 4),
            "graph_ms": round(t_graph, 4),
            "fused_GBps": round(8 * n1 / (t_fus * 1e-3) / 1e9, 1),
            "err_vs_fp64": {"unfused": abs(got_unf - ref), "fused": abs(got_fus - ref),
                            "graph": abs(got_graph - ref),
                            "cpu_reference_order": abs(float(oracle.reduce_add_f32_reforder(
                                oracle.zip_mul_f32(x1, y1))) - ref)},
            "int_reduce_ms": round(t_int, 4),
            "int_reduce_exact": bool(np.uint32(got_int & 0xFFFFFFFF)
                                     == oracle.reduce_add_u32_reforder(k1)),
        }
        progress("dot_2^20 done")

    # ---- store-only streams: Map<int(Index)> over 2^28 indices and
    # Map<int(IndexPoint)> over 16384 x 16384 points (4 B written per element)
    if want("index"):
        ni = 1 << 28
        oi = skelcl.DeviceBuffer(4 * ni)
        mi = skelcl.Map("int(Index)", "int func(Index i){ return (int)(i * 2654435761u); }")
        mp = skelcl.Map("int(IndexPoint)", "int func(IndexPoint p){ return p.x - 3 * p.y; }")
        t_i = _time(stream, lambda: ck(L.skl_map_index_run(mi.handle, oi.ptr, ni, 0, None, 0, st)))
        hi = oi.to_numpy(np.int32)
        idx = np.arange(0, ni, 4099, dtype=np.uint64)
        ok_i = bool(np.array_equal(hi[::4099], ((idx * 2654435761) & 0xFFFFFFFF)
                                   .astype(np.uint32).view(np.int32)))
        side = 1 << 14
        t_p = _time(stream, lambda: ck(L.skl_map_indexpoint_run(mp.handle, oi.ptr, side, side, 0,
                                                                None, 0, st)))
        hp = oi.to_numpy(np.int32).reshape(side, side)
        rr = np.arange(0, side, 97)
        ok_p = bool(np.array_equal(hp[rr], (np.arange(side)[None, :] - 3 * rr[:, None])
                                   .astype(np.int32)))
        ex["index_map_2^28"] = {
            "index_ms": round(t_i, 4), "index_GBps": round(4 * ni / (t_i * 1e-3) / 1e9, 1),
            "indexpoint_ms": round(t_p, 4),
            "indexpoint_GBps": round(4 * ni / (t_p * 1e-3) / 1e9, 1),
            "sampled_exact": ok_i and ok_p,
        }
        progress("index_map_2^28 done")
        del oi, hi, hp

    # ---- host buffers handed over: saxpy 2^26 through the Vector API
    if want("saxpy_host"):
        # (upload, kernel, download; pageable numpy memory), wall clock
        n2 = 1 << 26
        hx, hy = oracle.gen_f32_sym(SEED_X, n2), oracle.gen_f32_sym(SEED_Y, n2)
        sax_v = skelcl.Zip("float(float, float)",
                           "float func(float x, float y, float a){ return a*x+y; }")
        np.asarray(sax_v(skelcl.Vector(hx[:1024]), skelcl.Vector(hy[:1024]), np.float32(SAXPY_A)))
        walls = []
        for _ in range(3):
            t0 = time.perf_counter()
            res = np.asarray(sax_v(skelcl.Vector(hx), skelcl.Vector(hy), np.float32(SAXPY_A)))
            walls.append(time.perf_counter() - t0)
        w = min(walls)
        ex["saxpy_2^26_host_buffers"] = {
            "wall_ms": round(w * 1e3, 3), "GBps_algorithmic": round(12 * n2 / w / 1e9, 1),
            "bit_exact_vs_fma": bool(np.array_equal(res, oracle.saxpy_f32_fma(SAXPY_A, hx, hy))),
        }
        progress("saxpy_2^26_host_buffers done")
        del hx, hy, res

    # ---- configuration 3: MapOverlap 8192^2 Gaussian R=5 NEUTRAL + adjoint
    if want("mapoverlap"):
        R, S = 5, 8192
        w = oracle.gaussian_weights(R, 2.0)
        img = skelcl.DeviceBuffer(4 * S * S)
        Wb = skelcl.DeviceBuffer(4 * S * S)
        ck(L.skl_generate_f32_unit(img.ptr, S * S, 0x3001, 255.0, st))
        ck(L.skl_generate_f32_sym(Wb.ptr, S * S, 0x3002, st))
        outb, dinb = skelcl.DeviceBuffer(4 * S * S), skelcl.DeviceBuffer(4 * S * S)
        wv = skelcl.DeviceBuffer.from_array(w.ravel())
        gauss = skelcl.MapOverlap("float(float)", GAUSS_SRC, R, skelcl.NEUTRAL, 0)
        # the same user function behind a preprocessor line: not recognised
        # as the weighted stencil, so it runs the generated MapOverlap kernel
        gauss_gen = skelcl.MapOverlap("float(float)", "#define GENERIC 1\n" + GAUSS_SRC, R,
                                      skelcl.NEUTRAL, 0)
        warg, nw, keep, _ = gauss._marshal((wv,))
        if args.autotune:
            ck(L.skl_autotune_args(gauss.handle, S * S, warg, nw))
        t_gen = _time(stream, lambda: ck(L.skl_mapoverlap_run(gauss_gen.handle, outb.ptr, img.ptr,
                                                              S, S, warg, nw, st)))
        o_gen = outb.to_numpy(np.float32)
        t_mo = _time(stream, lambda: ck(L.skl_mapoverlap_run(gauss.handle, outb.ptr, img.ptr, S, S,
                                                             warg, nw, st)))
        t_adj = _time(stream, lambda: ck(L.skl_mapoverlap_stencil_backward(
            gauss.handle, dinb.ptr, Wb.ptr, wv.ptr, S, S, st)))
        himg = oracle.gen_f32_unit(0x3001, S * S, 255.0).reshape(S, S)
        rng = np.random.default_rng(0x3003)
        pix = rng.integers(0, S * S, 64).astype(np.int64)
        o = outb.to_numpy(np.float32)
        ref = oracle.mapoverlap_stencil_f64(himg, w, R, pix, 0)
        hW = oracle.gen_f32_sym(0x3002, S * S).reshape(S, S)
        dref = oracle.mapoverlap_stencil_adjoint_f64(hW, w, R, pix, 0)
        d = dinb.to_numpy(np.float32)
        flop = 2 * (2 * R + 1) ** 2 * S * S
        ex["mapoverlap_8192^2_r5"] = {
            "forward_ms": round(t_mo, 4), "forward_GBps": round(8 * S * S / (t_mo * 1e-3) / 1e9, 1),
            "forward_hbm_frac": round(8 * S * S / (t_mo * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "forward_TFLOPs": round(flop / (t_mo * 1e-3) / 1e12, 2),
            "forward_fp32_frac_of_157.3": round(flop / (t_mo * 1e-3) / 1e12 / 157.3, 4),
            "generated_kernel_ms": round(t_gen, 4),
            "bit_exact_vs_generated_kernel": bool(np.array_equal(o.view(np.uint32),
                                                                 o_gen.view(np.uint32))),
            "adjoint_ms": round(t_adj, 4),
            "max_rel_err_vs_fp64": float(np.max(np.abs(o[pix] - ref) / np.maximum(1, np.abs(ref)))),
            "adjoint_max_abs_err_vs_fp64": float(np.max(np.abs(d[pix] - dref))),
        }
        progress("mapoverlap_8192^2_r5 done")
        del img, Wb, outb, dinb, o, o_gen, d, himg, hW

    # ---- configuration 4: AllPairs 8192^3 (FP32 / BF16X3 / FP8), sq-Euclidean
    if want("allpairs"):
        M = 8192
        A = skelcl.DeviceBuffer(4 * M * M)
        B = skelcl.DeviceBuffer(4 * M * M)
        Cb = skelcl.DeviceBuffer(4 * M * M)
        ck(L.skl_generate_f32_sym(A.ptr, M * M, 0x4001, st))
        ck(L.skl_generate_f32_sym(B.ptr, M * M, 0x4002, st))
        hA = oracle.gen_f32_sym(0x4001, M * M).reshape(M, M)
        hB = oracle.gen_f32_sym(0x4002, M * M).reshape(M, M)
        rows = np.sort(np.random.default_rng(7).choice(M, 32, replace=False)).astype(np.int32)
        ref = oracle.allpairs_mm_f64(hA, hB, rows=rows)
        flop = 2.0 * M * M * M
        ap = {}
        for name, prec, peak in (("fp32", skelcl.PREC_FP32, 157.3), ("bf16x3", skelcl.PREC_BF16X3, 2500.0),
                                 ("fp8", skelcl.PREC_FP8, 5000.0)):
            t = _time(stream, lambda: ck(L.skl_matmul_f32(Cb.ptr, A.ptr, B.ptr, M, M, M, 0, 0, prec, st)))
            c = Cb.to_numpy(np.float32).reshape(M, M)[rows].astype(np.float64)
            err = float(np.linalg.norm(c - ref) / np.linalg.norm(ref))
            tf = flop / (t * 1e-3) / 1e12
            mfma_work = flop * (6 if name == "bf16x3" else 1)
            ap[name] = {"ms": round(t, 3), "TFLOPs_effective": round(tf, 1),
                        "unit_utilisation": round(mfma_work / (t * 1e-3) / 1e12 / peak, 4),
                        "normwise_err_vs_fp64": err}
        ap["bf16x3_speedup_vs_fp32"] = round(ap["fp32"]["ms"] / ap["bf16x3"]["ms"], 2)
        dC = skelcl.DeviceBuffer(4 * M * M)
        ck(L.skl_generate_f32_sym(dC.ptr, M * M, 0x4004, st))
        zip_ = skelcl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
        apk = skelcl.AllPairs("float(float, float)", reduce=add, zip=zip_,
                              precision=skelcl.PREC_BF16X3)
        t_bw = _time(stream, lambda: ck(L.skl_allpairs_backward(apk.handle, A.ptr, B.ptr, Cb.ptr,
                                                                Cb.ptr, dC.ptr, M, M, M, st)))
        ap["backward_bf16x3_ms"] = round(t_bw, 3)
        del A, B, Cb, dC, hA, hB
        n, d = 16384, 256
        P = skelcl.DeviceBuffer(4 * n * d)
        ck(L.skl_generate_f32_sym(P.ptr, n * d, 0x4003, st))
        hP = oracle.gen_f32_sym(0x4003, n * d).reshape(n, d)
        Pt = skelcl.DeviceBuffer.from_array(np.ascontiguousarray(hP.T))
        D = skelcl.DeviceBuffer(4 * n * n)
        sq = skelcl.AllPairs("float(float, float)",
                             "float func(lmatrix_t* row, rmatrix_t* col, const unsigned int dim) {"
                             " float s = 0; for (unsigned k = 0; k < dim; ++k) {"
                             " float t = getElementFromRow(row, k) - getElementFromColumn(col, k);"
                             " s += t*t; } return s; }")
        t_sq = _time(stream, lambda: ck(L.skl_allpairs_run(sq.handle, D.ptr, P.ptr, Pt.ptr, n, n, d,
                                                           None, 0, st)))
        rows = np.array([0, 1, 4097, 16383], np.int32)
        dref = oracle.allpairs_sqdist_f64(hP, np.ascontiguousarray(hP.T), rows=rows)
        dg = D.to_numpy(np.float32).reshape(n, n)[rows]
        ap["sqeuclid_16384x256"] = {"ms": round(t_sq, 3),
                                   "TFLOPs": round(3.0 * n * n * d / (t_sq * 1e-3) / 1e12, 1),
                                   "max_rel_err_vs_fp64": float(np.max(np.abs(dg - dref) /
                                                                       np.maximum(1, dref)))}
        ex["allpairs_8192^3"] = ap
        progress("allpairs_8192^3 done")
        del P, Pt, D

    # ---- configuration 5: batched segmented Reduce/Scan + 512 small products
    if want("batched"):
        import json as _json
        with open(os.path.join(ROOT, "tests", "golden", "batched_shapes.json")) as f:
            fx = _json.load(f)
        lens = np.array(fx["segment_lengths"], np.int64)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        tot = int(off[-1])
        segx = skelcl.DeviceBuffer(4 * tot)
        segy = skelcl.DeviceBuffer(4 * tot)
        ck(L.skl_generate_f32_sym(segx.ptr, tot, 0x5002, st))
        offb = skelcl.DeviceBuffer.from_array(off)
        sums = skelcl.DeviceBuffer(4 * len(lens))
        # the precision forms of float (+): DEFAULT (compensated FP32 inside a
        # block, FP64 across blocks) and FP64 (FP64 throughout)
        ADD = "float func(float x, float y){ return x+y; }"
        forms = {}
        for pname, prec in (("default", skelcl.PREC_DEFAULT), ("fp64", skelcl.PREC_FP64)):
            red_p = skelcl.Reduce("float(float)", ADD, "0", precision=prec)
            scan_p = skelcl.Scan("float(float)", ADD, "0.0f", precision=prec)
            f = {}
            f["segmented_reduce_ms"] = _time(stream, lambda: ck(L.skl_reduce_segmented_run(
                red_p.handle, sums.ptr, segx.ptr, tot, offb.ptr, len(lens), st)))
            f["segmented_scan_ms"] = _time(stream, lambda: ck(L.skl_scan_segmented_run(
                scan_p.handle, segy.ptr, segx.ptr, tot, offb.ptr, len(lens), 1, st)))
            f["scan_ms"] = _time(stream, lambda: ck(L.skl_scan_run(
                scan_p.handle, segy.ptr, segx.ptr, tot, 1, None, 0, st)))
            f["scan_adjoint_ms"] = _time(stream, lambda: ck(L.skl_scan_backward(
                scan_p.handle, segy.ptr, segx.ptr, tot, 1, st)))
            f["reduce_ms"] = _time(stream, lambda: ck(L.skl_reduce_run(
                red_p.handle, sums.ptr, segx.ptr, tot, None, 0, st)))
            forms[pname] = {k: round(v, 4) for k, v in f.items()}
        # the DEFAULT form's segment sums are the ones checked below
        ck(L.skl_reduce_segmented_run(add.handle, sums.ptr, segx.ptr, tot, offb.ptr, len(lens),
                                      st))
        t_sr = forms["default"]["segmented_reduce_ms"]
        t_ss = forms["default"]["segmented_scan_ms"]
        t_sb = forms["default"]["scan_adjoint_ms"]
        shapes = fx["gemm_shapes"]
        descs = (skelcl._lib.skl_gemm_desc * len(shapes))()
        bufs = []
        flops = 0.0
        for i, (m, nn, k) in enumerate(shapes):
            a_, b_, c_ = (skelcl.DeviceBuffer(4 * m * k), skelcl.DeviceBuffer(4 * k * nn),
                          skelcl.DeviceBuffer(4 * m * nn))
            ck(L.skl_generate_f32_sym(a_.ptr, m * k, 0x5004 + 2 * i, st))
            ck(L.skl_generate_f32_sym(b_.ptr, k * nn, 0x5005 + 2 * i, st))
            bufs += [a_, b_, c_]
            descs[i].A, descs[i].B, descs[i].C = a_.value, b_.value, c_.value
            descs[i].M, descs[i].N, descs[i].K = m, nn, k
            flops += 2.0 * m * nn * k
        dd = skelcl.DeviceBuffer(C.sizeof(descs))
        ck(L.skl_memcpy_h2d(dd.ptr, 0, C.cast(descs, C.c_void_p), C.sizeof(descs), None, 1))
        bt = {}
        for name, prec in (("fp32", skelcl.PREC_FP32), ("bf16x3", skelcl.PREC_BF16X3)):
            t = _time(stream, lambda: ck(L.skl_matmul_batched_f32(dd.ptr, len(shapes), prec, st)))
            bt[name] = {"ms": round(t, 3), "TFLOPs": round(flops / (t * 1e-3) / 1e12, 1)}
        hs = sums.to_numpy(np.float32)
        # the whole batched pipeline (segmented Reduce: 4 launches, segmented
        # Scan: 2, batched BF16X3 products: 3) eager and as one captured
        # hipGraph, at full size and on the first 64 / 8 segments and products
        scan_d = skelcl.Scan("float(float)", ADD, "0.0f")
        subsets = {}
        for nsm in (64, 8):
            dd_sm = skelcl.DeviceBuffer(C.sizeof(descs))
            ck(L.skl_memcpy_h2d(dd_sm.ptr, 0, C.cast(descs, C.c_void_p),
                                C.sizeof(skelcl._lib.skl_gemm_desc) * nsm, None, 1))
            subsets[f"first{nsm}"] = (nsm, skelcl.DeviceBuffer.from_array(off[:nsm + 1]),
                                      int(off[nsm]), dd_sm, nsm)

        def pipeline(nseg, ob, ne, descs_dev, ng):
            ck(L.skl_reduce_segmented_run(add.handle, sums.ptr, segx.ptr, ne, ob.ptr, nseg, st))
            ck(L.skl_scan_segmented_run(scan_d.handle, segy.ptr, segx.ptr, ne, ob.ptr, nseg, 1,
                                        st))
            ck(L.skl_matmul_batched_f32(descs_dev.ptr, ng, skelcl.PREC_BF16X3, st))

        pipe = {}
        for pname, argsp in [("full", (len(lens), offb, tot, dd, len(shapes)))] + \
                list(subsets.items()):
            t_e = _time(stream, lambda: pipeline(*argsp))
            ref_sums = sums.to_numpy(np.float32)
            g = C.c_void_p()
            ck(L.skl_graph_begin_capture(st))
            pipeline(*argsp)
            ck(L.skl_graph_end_capture(st, C.byref(g)))
            t_g = _time(stream, lambda: ck(L.skl_graph_launch(g, st)), settle_ms=0)
            same = bool(np.array_equal(sums.to_numpy(np.float32).view(np.uint32),
                                       ref_sums.view(np.uint32)))
            ck(L.skl_graph_destroy(g))
            pipe[pname] = {"eager_ms": round(t_e, 4), "graph_ms": round(t_g, 4),
                           "graph_sums_bit_identical": same}
        # The same work as ONE launch (skl_batch_run): the segment sums and the
        # inclusive segmented scan from one pass over the array, and the
        # BF16X3 products walked by 1 or 2 work-groups per CU meanwhile
        red_b = skelcl.Reduce("float(float)", ADD, "0")
        scan_b = skelcl.Scan("float(float)", ADD, "0.0f")
        sums_b = skelcl.DeviceBuffer(4 * len(lens))

        def batch_one(w, with_products=True):
            ck(L.skl_batch_run(red_b.handle, sums_b.ptr, scan_b.handle, segy.ptr, 1, segx.ptr,
                               tot, offb.ptr, len(lens), dd.ptr if with_products else None,
                               len(shapes) if with_products else 0, skelcl.PREC_BF16X3, w, st))

        one = {"segments_only_ms": round(_time(stream, lambda: batch_one(0, False)), 4)}
        for w in (1, 2):
            one[f"workers_per_cu_{w}_ms"] = round(_time(stream, lambda: batch_one(w)), 4)
        # the subsets of the eager pipeline above, as one launch each
        for pname, (nsm, ob, ne, dsm, ng) in subsets.items():
            one[f"{pname}_ms"] = round(_time(stream, lambda: ck(L.skl_batch_run(
                red_b.handle, sums_b.ptr, scan_b.handle, segy.ptr, 1, segx.ptr, ne, ob.ptr, nsm,
                dsm.ptr, ng, skelcl.PREC_BF16X3, 2, st))), 4)
        wbest = 1 if one["workers_per_cu_1_ms"] <= one["workers_per_cu_2_ms"] else 2
        batch_one(wbest)
        eager_sums = sums_b.to_numpy(np.float32)
        g = C.c_void_p()
        ck(L.skl_graph_begin_capture(st))
        batch_one(wbest)
        ck(L.skl_graph_end_capture(st, C.byref(g)))
        one["graph_ms"] = round(_time(stream, lambda: ck(L.skl_graph_launch(g, st)), settle_ms=0), 4)
        one["graph_sums_bit_identical"] = bool(np.array_equal(
            sums_b.to_numpy(np.float32).view(np.uint32), eager_sums.view(np.uint32)))
        ck(L.skl_graph_destroy(g))
        bad = C.c_int(-1)
        ck(L.skl_batch_check(scan_b.handle, C.byref(bad)))
        one["offsets_flagged"] = bool(bad.value)
        ck(L.skl_reduce_segmented_run(add.handle, sums.ptr, segx.ptr, tot, offb.ptr, len(lens),
                                      st))
        hs = sums.to_numpy(np.float32)
        chk = [0, 1, 2, 100]
        one["segment_sum_max_abs_err_vs_fp64"] = float(max(
            abs(float(eager_sums[s]) - oracle.sum_f64(
                oracle.gen_f32_sym(0x5002, int(off[s + 1]))[off[s]:off[s + 1]])) for s in chk))
        hx = {s: oracle.gen_f32_sym(0x5002, int(off[s + 1]))[off[s]:off[s + 1]] for s in chk}
        ex["batched"] = {
            "segments": len(lens), "elements": tot,
            "segmented_reduce_ms": round(t_sr, 3),
            "segmented_reduce_GBps": round(4 * tot / (t_sr * 1e-3) / 1e9, 1),
            "segmented_scan_ms": round(t_ss, 3),
            "segmented_scan_GBps": round(8 * tot / (t_ss * 1e-3) / 1e9, 1),
            "scan_adjoint_ms": round(t_sb, 3),
            "segment_sum_max_abs_err_vs_fp64": float(max(abs(float(hs[s]) - oracle.sum_f64(hx[s]))
                                                         for s in chk)),
            "gemms": len(shapes), "gemm_flop": flops, "gemm": bt,
            "float_add_precision_forms_ms": forms,
            "pipeline_eager_vs_graph": pipe,
            "one_launch": one,
        }
        progress("batched done")

    return ex


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--extras", type=int, default=1,
                    help="1: also measure the other configurations (AllPairs 8192^3 in "
                         "FP32 / BF16X3 / FP8 with errors vs FP64, MapOverlap 8192^2 R=5 "
                         "forward + adjoint, batched segmented Reduce/Scan + 512 GEMMs, "
                         "fused dot product)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: autotune the launch shapes of the timed skeletons first")
    ap.add_argument("--only", default="",
                    help="comma list of extra configurations (dot, index, saxpy_host, mapoverlap, "
                         "allpairs, batched): run only those and print their JSON (for "
                         "profiling); skips the headline legs and the CPU baseline")
    ap.add_argument("--ceiling", action="store_true",
                    help="only time plain read / write / copy / zip streaming probes over 1 GiB "
                         "at a sweep of grid sizes (the practical HBM ceiling) and print them")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also time an identity Map (copy) leg, for counter calibration")
    args = ap.parse_args()
    known = {"dot", "index", "saxpy_host", "mapoverlap", "allpairs", "batched"}
    if args.only and not set(args.only.split(",")) <= known:
        print(f"--only: choose from {', '.join(sorted(known))}", file=sys.stderr)
        return 2
    if args.gpus != 1:
        print(f"bench.py drives one GPU; --gpus {args.gpus} is not supported: the library "
              "runs every skeleton on one MI355X", file=sys.stderr)
        return 2

    import skelcl_amd as skelcl
    from skelcl_amd import _lib

    L = _lib.lib()
    ck = _lib.check
    skelcl.init()
    info = skelcl.device_info()
    stream = skelcl.Stream()
    st = stream.handle
    if args.ceiling:
        print(json.dumps({"ceiling": run_ceiling(skelcl, L, ck, stream),
                          "lds_ceiling": run_lds_ceiling(skelcl, L, ck, stream),
                          "valu_ceiling": run_valu_ceiling(skelcl, L, ck, stream),
                          "device": {"name": info["name"], "arch": info["arch"]}}), flush=True)
        return 0
    if args.only:
        print(json.dumps({"extras": run_extras(skelcl, L, ck, stream, args),
                          "device": {"name": info["name"], "arch": info["arch"]}}), flush=True)
        return 0

    # inputs resident in HBM
    xs, ys = skelcl.DeviceBuffer(4 * N), skelcl.DeviceBuffer(4 * N)
    xr = skelcl.DeviceBuffer(4 * N)
    x, y = skelcl.DeviceBuffer(4 * N), skelcl.DeviceBuffer(4 * N)
    sout = skelcl.DeviceBuffer(4 * N)
    rout = skelcl.DeviceBuffer(4)
    ck(L.skl_generate_i32_below(xs.ptr, N, SEED_SCAN, 1000, st))
    ck(L.skl_generate_f32_sym(xr.ptr, N, SEED_RED, st))
    ck(L.skl_generate_f32_sym(x.ptr, N, SEED_X, st))
    ck(L.skl_generate_f32_sym(y.ptr, N, SEED_Y, st))
    stream.synchronize()

    scan = skelcl.Scan("int(int)", "int func(int x, int y){ return x+y; }", "0")
    red = skelcl.Reduce("float(float)", "float func(float x, float y){ return x+y; }", "0")
    sax = skelcl.Zip("float(float, float)",
                     "float func(float x, float y, float a){ return a*x+y; }")
    a_arg, n_a, keep, _ = sax._marshal((np.float32(SAXPY_A),))
    shapes = {}
    if args.autotune:
        # launch shapes measured once per (skeleton, size class, device) and
        # cached; outside the timed region
        import ctypes as C
        ck(L.skl_autotune(scan.handle, N, 1))
        ck(L.skl_autotune(red.handle, N, 1))
        ck(L.skl_autotune_args(sax.handle, N, a_arg, n_a))
        for name, sk in (("scan", scan), ("reduce", red), ("saxpy", sax)):
            sh = (C.c_int * 4)()
            ck(L.skl_skeleton_launch_shape(sk.handle, sh))
            wg = C.c_int()
            ck(L.skl_skeleton_work_group_size(sk.handle, C.byref(wg)))
            shapes[name] = dict(zip(("unroll", "blocks_per_cu", "scan_j", "mo_rb"), list(sh)))
            shapes[name]["work_group_size"] = wg.value
        progress("launch shapes tuned")

    def run_scan():
        ck(L.skl_scan_run(scan.handle, sout.ptr, xs.ptr, N, 1, None, 0, st))

    def run_reduce():
        ck(L.skl_reduce_run(red.handle, rout.ptr, xr.ptr, N, None, 0, st))

    def run_saxpy():
        ck(L.skl_zip_run(sax.handle, ys.ptr, x.ptr, y.ptr, N, a_arg, n_a, st))

    legs = [("scan", run_scan, 8 * N), ("reduce", run_reduce, 4 * N),
            ("saxpy", run_saxpy, 12 * N)]
    if args.copy_calibration:
        # identity Map over N floats: a plain copy kernel of known size
        # (4N bytes read, 4N written) that calibrates the HBM byte counters
        copy = skelcl.Map("float(float)", "float func(float x){ return x; }")

        def run_copy():
            ck(L.skl_map_run(copy.handle, ys.ptr, x.ptr, N, None, 0, st))

        legs.append(("copy", run_copy, 8 * N))

    for _ in range(args.warmup):
        for _, f, _ in legs:
            f()
    stream.synchronize()

    # per-launch events on the launch stream
    ev = [[skelcl.Event() for _ in range(len(legs) + 1)] for _ in range(args.steps)]
    t_all0, t_all1 = skelcl.Event(), skelcl.Event()
    ck(L.skl_device_synchronize())
    wall0 = time.perf_counter()
    t_all0.record(stream)
    for s in range(args.steps):
        ev[s][0].record(stream)
        for i, (_, f, _) in enumerate(legs):
            f()
            ev[s][i + 1].record(stream)
    t_all1.record(stream)
    stream.synchronize()
    ck(L.skl_device_synchronize())
    wall = time.perf_counter() - wall0
    total_ms = t_all0.elapsed_ms(t_all1)
    per = {name: [ev[s][i].elapsed_ms(ev[s][i + 1]) for s in range(args.steps)]
           for i, (name, _, _) in enumerate(legs)}

    total_bytes = sum(b for name, _, b in legs if name != "copy")
    ms_per_step = total_ms / args.steps
    if args.copy_calibration:  # profiling run: the step is the workload legs only
        ms_per_step = sum(statistics.mean(per[name]) for name, _, _ in legs if name != "copy")
    value = total_bytes / (ms_per_step * 1e-3) / 1e9
    progress("timed steps done")
    legs_out = {}
    for name, _, b in legs:
        avg = statistics.mean(per[name])
        med = statistics.median(per[name])
        legs_out[name] = {"avg_ms": round(avg, 5), "median_ms": round(med, 5),
                          "GBps": round(b / (avg * 1e-3) / 1e9, 1),
                          "hbm_frac": round(b / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "bytes": b}
    dom = max((k for k in legs_out if k != "copy"), key=lambda k: legs_out[k]["avg_ms"])
    dl = legs_out[dom]
    dom_kernel = {"scan": "skl_scan", "reduce": "skl_reduce", "saxpy": "skl_zip"}[dom]

    # correctness of the timed outputs (checked after timing)
    got_red = float(rout.to_numpy(np.float32, 1)[0])
    tail = sout.to_numpy(np.int32, N)[-1:]
    result = {
        "metric": "GB/s & AllPairs TFLOP/s: MFMA/batched/adjoint vs plain FP32, err vs FP64",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f32",
        "data": "synthetic (SplitMix64 counter-based, generated in HBM)",
        "config": {"workload": "bandwidth skeletons N=2^28: inclusive Scan(+) int32, "
                               "Reduce(+) float32 (compensated), Zip saxpy a=2.5 float32",
                   "N": N, "global_batch": N, "parallelism": "single GPU"},
        "legs": legs_out,
        "roofline": {"kernel": dom_kernel,
                     "bound": "hbm", "achieved": round(dl["bytes"] / (dl["avg_ms"] * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dl["bytes"] / (dl["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom_kernel),
                     "algorithmic_bytes": dl["bytes"]},
        "launch_shapes": shapes,
        "device": {"name": info["name"], "arch": info["arch"], "cus": info["compute_units"]},
        "wall_s": round(wall, 3),
        "reduce_result": got_red,
        "scan_last": int(tail[0]),
    }
    if not args.no_cpu:
        nthreads, ncpu, model = host_cpus()
        cb = cpu_baseline(nthreads)
        progress("cpu baseline done")
        ref64 = cb["reduce_fp64"]
        result["err_vs_fp64"] = {
            "reduce_gpu_abs": abs(got_red - ref64), "reduce_cpu_reforder_abs":
                abs(cb["reduce_ref"] - ref64), "fp64": ref64}
        result["scan_last_matches_cpu"] = (np.uint32(cb["scan_last"] & 0xFFFFFFFF)
                                           == np.int32(tail[0]).view(np.uint32))
        cpu_gbs = total_bytes / cb["total_s"] / 1e9
        result["cpu_baseline"] = {
            "value": round(cpu_gbs, 2), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "host_cpus": ncpu, "cpu_model": model,
            "sample": "one full step (N=2^28 each): oracle reference-order multi-pass "
                      "Blelloch Scan + 8192-partial Reduce + saxpy, pthreads; each skeleton "
                      "run once untimed, then the median of 5; "
                      f"scan {cb['times_s']['scan']:.2f}s, reduce "
                      f"{cb['times_s']['reduce']:.2f}s, saxpy {cb['times_s']['saxpy']:.2f}s"}
        result["scan_last_matches_cpu"] = bool(result["scan_last_matches_cpu"])
    if args.extras:
        # free the bandwidth workload's buffers first (7 GiB)
        del xs, ys, xr, x, y, sout
        result["extras"] = run_extras(skelcl, L, ck, stream, args)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

// Batched segmented Reduce(+) / Scan(+) on float or int32 (beyond the reference):
// many uneven segments, given by device offsets, in one main launch after a
// small prologue launch.
//
// Reduce: a one-work-group prologue validates the offsets (non-decreasing,
// inside [0, n]), cuts segment s (length L) into ceil(L / 16384) chunks,
// records each segment's first chunk and each chunk's segment.  One
// work-group per chunk stores the chunk's sum as an FP64 partial; a second
// kernel, one wave per segment, adds the segment's partials in FP64 in a
// fixed order and writes out[s].
// Scan: the elements are cut into 32768-element tiles regardless of the
// segments and scanned in one pass with the segmented operator (see the
// flat Scan section below); a prologue, one thread per segment, validates
// the offsets and records each tile's first segment.  Every output is
// float(carry + local prefix), the look-back carries in FP64.
// Inside a chunk / tile the sums are compensated FP32 (double-float pairs,
// acc_df), or FP64 throughout when the caller asks for FP64.  int32 data
// (the same bits passed as float) sums exactly modulo 2^32 (acc_u32): chunk
// partials are held in FP64 as their unsigned 32-bit values, and the
// combine adds them modulo 2^32.
// Invalid offsets set an error flag: no chunk or tile runs, Reduce writes NaN
// (float) or INT_MIN (int32).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"
#include "ld_util.h"
#include "seg_tile.h"

namespace {

constexpr int WG = 256;
constexpr int SUM_J = 16;                                 // float4 packs per thread
constexpr long long SUM_CH = (long long)WG * SUM_J * 4;   // 16384 elements per chunk

struct SegCtl {
  unsigned long long total;  // number of chunks (0 after an error)
  unsigned int error;        // 1: negative length, 2: offsets outside [0, n]
};

// Reduce prologue, one work-group of 1024 threads: validates the offsets,
// cuts segment s (length L) into ceil(L / SUM_CH) chunks and writes
// cstart[s] = its first chunk (cstart[nseg] = total; the total is 0 after an
// error or beyond cap chunks).  Thread t owns the segments
// [t * per, (t + 1) * per): it reads their offsets (loads independent of one
// another, so they are in flight together), one block scan of the per-thread
// chunk counts places them, and a second pass over the same offsets writes.
__global__ __launch_bounds__(1024) void seg_sum_prologue(const int64_t* __restrict__ off,
                                                         int nseg, unsigned long long n,
                                                         long long cap,
                                                         long long* __restrict__ cstart,
                                                         SegCtl* ctl)
{
  __shared__ long long wsum[1024 / 64];
  __shared__ unsigned err;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int per = (nseg + 1023) / 1024;
  const int s0 = (int)threadIdx.x * per;
  const int s1 = s0 + per < nseg ? s0 + per : nseg;
  if (threadIdx.x == 0)
    err = (off[0] < 0 || off[nseg] < 0 || (unsigned long long)off[nseg] > n) ? 2u : 0u;
  __syncthreads();
  long long tot = 0;
  bool bad = false;
#pragma unroll 8
  for (int s = s0; s < s1; ++s) {
    const long long len = off[s + 1] - off[s];
    if (len < 0) bad = true;
    else tot += (len + SUM_CH - 1) / SUM_CH;
  }
  if (bad) atomicOr(&err, 1u);
  long long v = tot;  // inclusive scan over the threads
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[wv] = v;
  __syncthreads();
  long long c = v - tot;
  for (int w = 0; w < wv; ++w) c += wsum[w];
  for (int s = s0; s < s1; ++s) {
    const long long len = off[s + 1] - off[s];
    cstart[s] = c;
    c += len < 0 ? 0 : (len + SUM_CH - 1) / SUM_CH;
  }
  if (threadIdx.x == 1023) {
    cstart[nseg] = c;
    ctl->total = err || (unsigned long long)c > (unsigned long long)cap ? 0 : (unsigned long long)c;
    ctl->error = err;
  }
}

// one wave per segment: cseg[c] = s over the segment's chunks
__global__ __launch_bounds__(256) void seg_sum_fill(const long long* __restrict__ cstart,
                                                    int nseg, int* __restrict__ cseg,
                                                    const SegCtl* __restrict__ ctl)
{
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= nseg || ctl->total == 0) return;
  const long long c1 = cstart[s + 1];
  for (long long c = cstart[s] + (threadIdx.x & 63); c < c1; c += 64) cseg[c] = s;
}

template <typename A>
__device__ __forceinline__ A acc_shfl_down(const A& v, int o)
{
  int lo, hi;
  acc_bits(v, &lo, &hi);
  return acc_from_bits<A>(__shfl_down(lo, o, 64), __shfl_down(hi, o, 64));
}

// the work-group's sum, as FP64 (valid in thread 0)
template <typename A>
__device__ __forceinline__ double block_sum(A v, A* red)
{
  for (int o = 32; o >= 1; o >>= 1) v = acc_add(v, acc_shfl_down(v, o));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  A t = acc_zero<A>();
  if (threadIdx.x == 0)
    for (int i = 0; i < WG / 64; ++i) t = acc_add(t, red[i]);
  return acc_to_double(t);
}
template <>
__device__ __forceinline__ double block_sum<acc_u32>(acc_u32 v, acc_u32* red)
{
  for (int o = 32; o >= 1; o >>= 1) v.v += (unsigned)__shfl_down((int)v.v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  unsigned t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < WG / 64; ++i) t += red[i].v;
  return (double)t;  // exact: t < 2^32
}

// one work-group per chunk: sums in A (4 accumulators per thread, fixed
// order), a full chunk read as 16 float4 packs per thread all in flight
// (dword-aligned: a segment may start anywhere)
template <typename A>
__global__ __launch_bounds__(WG) void seg_sum_kernel(const float* __restrict__ in,
                                                     const int64_t* __restrict__ off,
                                                     const long long* __restrict__ cstart,
                                                     const int* __restrict__ cseg,
                                                     double* __restrict__ partials,
                                                     const SegCtl* __restrict__ ctl)
{
  typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
  __shared__ A red[WG / 64];
  const long long c = blockIdx.x;
  if (c >= (long long)ctl->total) return;
  const int s = cseg[c];
  const long long k = c - cstart[s];
  const long long lo = off[s] + k * SUM_CH;
  const long long end = off[s + 1];
  const long long hi = lo + SUM_CH < end ? lo + SUM_CH : end;
  A a0 = acc_zero<A>(), a1 = a0, a2 = a0, a3 = a0;
  if (hi - lo == SUM_CH) {
    f4u v[SUM_J];
#pragma unroll
    for (int j = 0; j < SUM_J; ++j)
      v[j] = __builtin_nontemporal_load(
          reinterpret_cast<const f4u*>(in + lo + ((long long)j * WG + threadIdx.x) * 4));
#pragma unroll
    for (int j = 0; j < SUM_J; ++j) {
      a0 = acc_add(a0, v[j][0]);
      a1 = acc_add(a1, v[j][1]);
      a2 = acc_add(a2, v[j][2]);
      a3 = acc_add(a3, v[j][3]);
    }
  } else {
    // the last chunk of a segment
    long long i = lo + threadIdx.x;
    for (; i + 3 * WG < hi; i += 4 * WG) {
      const float x0 = in[i], x1 = in[i + WG], x2 = in[i + 2 * WG], x3 = in[i + 3 * WG];
      a0 = acc_add(a0, x0);
      a1 = acc_add(a1, x1);
      a2 = acc_add(a2, x2);
      a3 = acc_add(a3, x3);
    }
    for (; i < hi; i += WG) a0 = acc_add(a0, in[i]);
  }
  const double tot = block_sum<A>(acc_add(acc_add(a0, a1), acc_add(a2, a3)), red);
  if (threadIdx.x == 0) partials[c] = tot;
}

// one wave per segment: its partials in a fixed order (lane-strided FP64
// sums, then a fixed shuffle tree); INT: the partials are unsigned 32-bit
// values, added modulo 2^32
template <bool INT>
__global__ __launch_bounds__(WG) void seg_combine_kernel(const long long* __restrict__ cstart,
                                                         int nseg,
                                                         const double* __restrict__ partials,
                                                         float* __restrict__ out,
                                                         const SegCtl* __restrict__ ctl)
{
  const int s = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= nseg) return;
  if (ctl->error) {
    if (lane == 0) out[s] = INT ? __uint_as_float(0x80000000u) : __builtin_nanf("");
    return;
  }
  const long long c0 = cstart[s], c1 = cstart[s + 1];
  if (INT) {
    unsigned u = 0;
    for (long long c = c0 + lane; c < c1; c += 64) u += (unsigned)partials[c];
    for (int o = 32; o >= 1; o >>= 1) u += (unsigned)__shfl_down((int)u, o, 64);
    if (lane == 0) out[s] = __uint_as_float(u);
    return;
  }
  double v = 0.0;
  for (long long c = c0 + lane; c < c1; c += 64) v += partials[c];
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) out[s] = (float)v;
}

// ---- flat segmented Scan ------------------------------------------------
// One pass over tiles of FTILE elements cut regardless of the segments, the
// segmented operator and a decoupled look-back (seg_tile.h); a prologue
// checks the offsets and writes each tile's first segment.
constexpr int FWG = 512;                                   // threads per tile
constexpr int FJ = 16;                                     // float4 packs per lane
constexpr long long FTILE = (long long)FWG * FJ * 4;       // 32768 elements

// Prologue, one thread per run s = -1 .. nseg (run -1 is [0, offsets[0]),
// run s < nseg is segment s, run nseg is [offsets[nseg], inf)):
//   * validates what it reads (offsets[0] >= 0, offsets[nseg] <= n,
//     offsets[s] <= offsets[s + 1]); a failure stores epoch << 2 | bit into
//     *err, epoch being the one the ticket word holds for the coming scan
//     launch, so a flag left by an earlier launch never matches;
//   * writes tfirst[t] = s for the tiles t = 0 .. ntiles starting inside a
//     valid run: tfirst[t] = the largest s with offsets[s] <= t * FTILE, -1
//     before offsets[0].  Valid runs are disjoint, so the writes total
//     ntiles + 1 whatever the segment count.
__global__ __launch_bounds__(256) void seg_flat_prologue(const int64_t* __restrict__ off,
                                                         int nseg, unsigned long long n,
                                                         long long ntiles,
                                                         int* __restrict__ tfirst,
                                                         const unsigned long long* ticket,
                                                         unsigned* err)
{
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x - 1;
  if (s > nseg) return;
  const long long a = s < 0 ? 0 : off[s];
  const long long b = s < nseg ? off[s + 1] : (long long)n;
  unsigned bad = 0;
  if (s < 0 && (b < 0 || (unsigned long long)off[nseg] > n || off[nseg] < 0)) bad = 2u;
  if (s >= 0 && s < nseg && b < a) bad = 1u;
  if (bad) {
    const unsigned ep = (unsigned)(__hip_atomic_load(ticket, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT) >> 32);
    __hip_atomic_store(err, (ep << 2) | bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (a < 0 || b < a || (unsigned long long)b > n) return;  // a neighbour flags it
  long long t1 = ntiles + 1;
  if (s < nseg) {
    const long long tb = (b + FTILE - 1) / FTILE;
    if (tb < t1) t1 = tb;
  }
  for (long long t = (a + FTILE - 1) / FTILE; t < t1; ++t) tfirst[t] = (int)s;
}

template <typename A>
__global__ __launch_bounds__(FWG) void seg_flat_scan_kernel(
    const float* __restrict__ in, float* __restrict__ out, unsigned long long n,
    const int64_t* __restrict__ off, int nseg, const int* __restrict__ tfirst,
    unsigned long long* __restrict__ ticket, unsigned long long* __restrict__ pub,
    int inclusive, const unsigned* __restrict__ err)
{
  __shared__ segtile::Smem<A, FWG, FJ, 0> sm;
  __shared__ unsigned long long s_tile;
  __shared__ unsigned s_epoch, s_err;

  // ticket word = epoch << 32 | tile; the work-group drawing the last tile
  // re-arms it for the next launch (every other draw has happened by then)
  if (threadIdx.x == 0) {
    const unsigned long long w =
        __hip_atomic_fetch_add(ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t = w & 0xFFFFFFFFull;
    const unsigned ep = (unsigned)(w >> 32);
    if (t == (unsigned long long)gridDim.x - 1)
      __hip_atomic_store(ticket, (unsigned long long)((ep + 1u) & 0x3FFFFFFFu) << 32,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_tile = t;
    s_epoch = ep;
    const unsigned e = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_err = (e >> 2) == ep && (e & 3u) != 0;
  }
  segtile::tile_clear(sm);
  __syncthreads();
  if (s_err) return;  // invalid offsets (every tile sees it): write nothing
  segtile::Job jb{in, out, nullptr, n, off, nseg, tfirst, pub, nullptr,
                  (long long)gridDim.x, inclusive};
  segtile::tile_run<A, FWG, FJ, 0, false>(jb, s_tile, s_epoch, sm);
}

int device_cus()
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  return cus;
}

size_t max_chunks(unsigned long long n, int nseg, long long ch)
{
  return (size_t)(n / (unsigned long long)ch) + (size_t)nseg + 1;
}

size_t pad256(size_t b) { return (b + 255) & ~(size_t)255; }

// Reduce workspace: SegCtl (256 B) | cstart over the segments (+1) | cseg
// over the chunks | FP64 partials over the chunks
size_t sum_bytes(unsigned long long n, int nseg)
{
  const size_t ch = max_chunks(n, nseg, SUM_CH);
  return 256 + pad256(((size_t)nseg + 1) * 8) + pad256(ch * 4) + ch * 8;
}

// flat Scan workspace: tile ticket (256 B) | error word (256 B) | tfirst over
// the tiles (+1), padded to 256 | two 64-bit status words per tile
long long flat_tiles(unsigned long long n) { return (long long)((n + FTILE - 1) / FTILE); }
size_t flat_table_bytes(long long ntiles) { return pad256((size_t)(ntiles + 1) * 4); }
size_t flat_bytes(unsigned long long n)
{
  const long long t = flat_tiles(n);
  return 512 + flat_table_bytes(t) + (size_t)t * 16;
}

}  // namespace

extern "C" size_t skl_internal_segmented_workspace(unsigned long long n, int nseg)
{
  const size_t sum = sum_bytes(n, nseg);
  const size_t scan = flat_bytes(n);
  return sum > scan ? sum : scan;
}

// form: 0 compensated FP32, 1 FP64, 2 int32 (exact modulo 2^32)
extern "C" int skl_internal_segmented_sum(float* out, const float* in, unsigned long long n,
                                          const int64_t* offsets, int nseg, int form, void* ws,
                                          hipStream_t s)
{
  const size_t cap = max_chunks(n, nseg, SUM_CH);
  if (cap > 0x7FFFFFFEull) return SKL_ERR_INVALID;
  char* p = static_cast<char*>(ws);
  SegCtl* ctl = reinterpret_cast<SegCtl*>(p);
  long long* cstart = reinterpret_cast<long long*>(p + 256);
  int* cseg = reinterpret_cast<int*>(p + 256 + pad256(((size_t)nseg + 1) * 8));
  double* partials = reinterpret_cast<double*>(reinterpret_cast<char*>(cseg) + pad256(cap * 4));
  seg_sum_prologue<<<1, 1024, 0, s>>>(offsets, nseg, n, (long long)cap, cstart, ctl);
  seg_sum_fill<<<(unsigned)((nseg + 3) / 4), 256, 0, s>>>(cstart, nseg, cseg, ctl);
  if (form == 2)
    seg_sum_kernel<acc_u32><<<(unsigned)cap, WG, 0, s>>>(in, offsets, cstart, cseg, partials, ctl);
  else if (form == 1)
    seg_sum_kernel<double><<<(unsigned)cap, WG, 0, s>>>(in, offsets, cstart, cseg, partials, ctl);
  else
    seg_sum_kernel<acc_df><<<(unsigned)cap, WG, 0, s>>>(in, offsets, cstart, cseg, partials, ctl);
  const unsigned cg = (unsigned)((nseg + WG / 64 - 1) / (WG / 64));
  if (form == 2)
    seg_combine_kernel<true><<<cg, WG, 0, s>>>(cstart, nseg, partials, out, ctl);
  else
    seg_combine_kernel<false><<<cg, WG, 0, s>>>(cstart, nseg, partials, out, ctl);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

extern "C" int skl_internal_segmented_scan(float* out, const float* in, unsigned long long n,
                                           const int64_t* offsets, int nseg, int inclusive,
                                           int form, void* ws, hipStream_t s)
{
  const long long ntiles = flat_tiles(n);
  if (ntiles > 0x7FFFFFFE) return SKL_ERR_INVALID;
  char* p = static_cast<char*>(ws);
  unsigned long long* ticket = reinterpret_cast<unsigned long long*>(p);
  unsigned* err = reinterpret_cast<unsigned*>(p + 256);
  int* tfirst = reinterpret_cast<int*>(p + 512);
  unsigned long long* pub =
      reinterpret_cast<unsigned long long*>(p + 512 + flat_table_bytes(ntiles));
  seg_flat_prologue<<<(unsigned)((nseg + 2 + 255) / 256), 256, 0, s>>>(offsets, nseg, n, ntiles,
                                                                      tfirst, ticket, err);
  if (form == 2)
    seg_flat_scan_kernel<acc_u32><<<(unsigned)ntiles, FWG, 0, s>>>(
        in, out, n, offsets, nseg, tfirst, ticket, pub, inclusive ? 1 : 0, err);
  else if (form == 1)
    seg_flat_scan_kernel<double><<<(unsigned)ntiles, FWG, 0, s>>>(
        in, out, n, offsets, nseg, tfirst, ticket, pub, inclusive ? 1 : 0, err);
  else
    seg_flat_scan_kernel<acc_df><<<(unsigned)ntiles, FWG, 0, s>>>(
        in, out, n, offsets, nseg, tfirst, ticket, pub, inclusive ? 1 : 0, err);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

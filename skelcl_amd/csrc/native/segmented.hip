// Batched segmented Reduce(+) / Scan(+) on float (beyond the reference):
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
// acc_df), or FP64 throughout when the caller asks for FP64.
// Invalid offsets set an error flag: no chunk or tile runs, Reduce writes NaN.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"
#include "ld_util.h"

namespace {

constexpr int WG = 256;
constexpr int SUM_J = 16;                                 // float4 packs per thread
constexpr long long SUM_CH = (long long)WG * SUM_J * 4;   // 16384 elements per chunk
constexpr unsigned ST_AGG = 1u, ST_INCL = 2u;

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
// sums, then a fixed shuffle tree)
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
    if (lane == 0) out[s] = __builtin_nanf("");
    return;
  }
  const long long c0 = cstart[s], c1 = cstart[s + 1];
  double v = 0.0;
  for (long long c = c0 + lane; c < c1; c += 64) v += partials[c];
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) out[s] = (float)v;
}

__device__ __forceinline__ void publish(unsigned long long* pub, long long c, unsigned state,
                                        double v, unsigned epoch)
{
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned long long tag = (unsigned long long)((epoch << 2) | state) << 32;
  __hip_atomic_store(&pub[2 * c], tag | (bits & 0xFFFFFFFFull), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&pub[2 * c + 1], tag | (bits >> 32), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// state (AGG / INCL) once both words carry the same tag of this launch, else 0
__device__ __forceinline__ unsigned read_status(const unsigned long long* pub, long long c,
                                                unsigned epoch, double* v)
{
  const unsigned long long w0 =
      __hip_atomic_load(&pub[2 * c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long w1 =
      __hip_atomic_load(&pub[2 * c + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned t0 = (unsigned)(w0 >> 32), t1 = (unsigned)(w1 >> 32);
  if (t0 != t1 || (t0 >> 2) != epoch) return 0;
  *v = __longlong_as_double((long long)((w1 << 32) | (w0 & 0xFFFFFFFFull)));
  return t0 & 3u;
}

// ---- flat segmented Scan -------------------------------------------------
// The n elements are cut into tiles of FTILE, independent of the segments.
// Element e is a head when e == offsets[s] for some s (offsets[nseg] marks
// the end of the last segment); a scan value restarts at every head.  The
// pair (open, v) -- "no head seen yet", partial sum -- combines as
//   (o1, v1) (+) (o2, v2) = (o1 && o2, o2 ? v1 + v2 : v2),
// which is associative, so the tile runs the same single-pass decoupled
// look-back as the plain Scan.  A tile that holds a head knows its inclusive
// value at once (the sum after its last head) and publishes it as INCL
// straight away, so a look-back never passes a tile with a head.
constexpr int FWG = 512;                                   // threads per tile
constexpr int FJ = 16;                                     // float4 packs per lane
constexpr long long FTILE = (long long)FWG * FJ * 4;       // 32768 elements
constexpr int FWAVES = FWG / 64;

// v of the lane named by the DPP control (zero where the control names no
// lane or the row is outside ROW_MASK), moved as two dwords
template <int CTRL, int ROW_MASK, typename A>
__device__ __forceinline__ A dpp_acc(const A& v)
{
  int lo, hi;
  acc_bits(v, &lo, &hi);
  return acc_from_bits<A>(__builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false),
                          __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false));
}

// inclusive segmented wave scan: lane l ends with the sum of lanes
// [hl, l], hl = the last lane <= l holding a head (0 when none).  Steps:
// row_shr 1, 2, 4, 8 inside each 16-lane row, then row_bcast 15 (lane 15 of
// each row into rows 1 and 3) and row_bcast 31 (lane 31 into rows 2 and 3);
// a step adds its source lane s only when s >= hl.  All VALU, no LDS.
template <typename A>
__device__ __forceinline__ A seg_wave_scan(A v, int lane, int hl)
{
  A t;
  t = dpp_acc<0x111, 0xf>(v);
  if (lane - 1 >= hl) v = acc_add(v, t);
  t = dpp_acc<0x112, 0xf>(v);
  if (lane - 2 >= hl) v = acc_add(v, t);
  t = dpp_acc<0x114, 0xf>(v);
  if (lane - 4 >= hl) v = acc_add(v, t);
  t = dpp_acc<0x118, 0xf>(v);
  if (lane - 8 >= hl) v = acc_add(v, t);
  t = dpp_acc<0x142, 0xa>(v);
  if ((lane & ~15) - 1 >= hl) v = acc_add(v, t);
  t = dpp_acc<0x143, 0xc>(v);
  if (31 >= hl) v = acc_add(v, t);
  return v;
}

// value of lane - 1 (zero in lane 0)
template <typename A>
__device__ __forceinline__ A wave_shr1(const A& v) { return dpp_acc<0x138, 0xf>(v); }

template <typename A>
__device__ __forceinline__ A readlane63(const A& v)
{
  int lo, hi;
  acc_bits(v, &lo, &hi);
  return acc_from_bits<A>(__builtin_amdgcn_readlane(lo, 63), __builtin_amdgcn_readlane(hi, 63));
}

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
  typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
  __shared__ unsigned long long s_tile;
  __shared__ unsigned s_epoch, s_err;
  __shared__ unsigned s_bits[FTILE / 32];
  __shared__ A s_lpv[FJ][FWG];  // per pack: wave-exclusive prefix (own thread only)
  __shared__ A s_wval[FWAVES];
  __shared__ int s_wopen[FWAVES];
  __shared__ double s_carry;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

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
  for (int i = threadIdx.x; i < (int)(FTILE / 32); i += FWG) s_bits[i] = 0;
  __syncthreads();
  if (s_err) return;  // invalid offsets: write nothing
  const unsigned long long tile = s_tile;
  const unsigned epoch = s_epoch;
  const unsigned long long tbase = tile * (unsigned long long)FTILE;
  const unsigned long long wbase = tbase + (unsigned long long)wv * (64 * FJ * 4);
  const bool full = tbase + FTILE <= n;

  float x[FJ][4];
  if (full) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      const f4u v = __builtin_nontemporal_load(
          reinterpret_cast<const f4u*>(in + wbase + ((unsigned long long)j * 64 + lane) * 4));
#pragma unroll
      for (int k = 0; k < 4; ++k) x[j][k] = v[k];
    }
  } else {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned long long e = wbase + ((unsigned long long)j * 64 + lane) * 4 + k;
        x[j][k] = ld_or(in, (size_t)e, 0, e < n, 0.0f);
      }
  }

  // head bits of the tile: offsets[s0 .. s1].  A few are read by wave 0
  // with scalar loads, which leave the data loads above in flight (a vector
  // load issued after them could only be waited for together with them).
  {
    const int s1 = tfirst[tile + 1];
    int s0 = tfirst[tile];
    if (s0 < 0) s0 = 0;
    if (s1 - s0 < 16) {
      if (wv == 0)
        for (int s = s0; s <= s1; ++s) {
          const long long p = off[s] - (long long)tbase;
          if (lane == 0 && p >= 0 && p < FTILE) atomicOr(&s_bits[p >> 5], 1u << (p & 31));
        }
    } else {
      for (int s = s0 + (int)threadIdx.x; s <= s1; s += FWG) {
        const long long p = off[s] - (long long)tbase;
        if (p >= 0 && p < FTILE) atomicOr(&s_bits[p >> 5], 1u << (p & 31));
      }
    }
  }
  __syncthreads();
  const long long lo = off[0], hi = off[nseg];

  // per pack: (open, sum after its last head); one segmented wave scan per
  // row of 64 packs
  unsigned lpo = 0;  // bit j: pack j's wave-exclusive prefix is open
  bool ropen = true;
  A rval = acc_zero<A>();
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const int p = (wv * 64 * FJ + j * 64 + lane) * 4;
    const unsigned hb = (s_bits[p >> 5] >> (p & 31)) & 0xFu;
    A ps = acc_zero<A>();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (hb & (1u << k)) ps = acc_zero<A>();
      ps = acc_add(ps, x[j][k]);
    }
    const unsigned long long M = __ballot(hb != 0);
    const unsigned long long le = M & ((2ull << lane) - 1ull);
    const int hl = le ? 63 - __builtin_clzll(le) : 0;
    const A inc = seg_wave_scan(ps, lane, hl);
    const A ex = wave_shr1(inc);
    const bool eopen = (M & ((1ull << lane) - 1ull)) == 0;
    const A rt = readlane63(inc);
    s_lpv[j][threadIdx.x] = eopen ? acc_add(rval, ex) : ex;
    if (eopen && ropen) lpo |= 1u << j;
    if (M) {
      ropen = false;
      rval = rt;
    } else {
      rval = acc_add(rval, rt);
    }
  }
  // keep the floats, not their FP64 conversions, live until the outputs
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int k = 0; k < 4; ++k) asm volatile("" : "+v"(x[j][k]));
  if (lane == 0) {
    s_wval[wv] = rval;
    s_wopen[wv] = ropen ? 1 : 0;
  }
  __syncthreads();
  bool wopen = true, topen = true;
  A wval = acc_zero<A>(), tval = acc_zero<A>();
#pragma unroll
  for (int w = 0; w < FWAVES; ++w) {
    const A v = s_wval[w];
    const bool o = s_wopen[w] != 0;
    if (w < wv) {
      wval = o ? acc_add(wval, v) : v;
      wopen = wopen && o;
    }
    tval = o ? acc_add(tval, v) : v;
    topen = topen && o;
  }
  const double tsum = acc_to_double(tval);

  // look-back (wave 0); a tile whose first element is a head needs no carry
  if (wv == 0) {
    const bool need = tile > 0 && (s_bits[0] & 1u) == 0;
    if (lane == 0) publish(pub, (long long)tile, topen ? ST_AGG : ST_INCL, tsum, epoch);
    double excl = 0.0;
    if (need) {
      long long wend = (long long)tile;  // window [wend - 64, wend)
      for (;;) {
        const long long idx = wend - 64 + lane;
        unsigned st = ST_INCL;
        double v = 0.0;
        if (idx >= 0) {
          do {
            st = read_status(pub, idx, epoch, &v);
          } while (st == 0);
        }
        const unsigned long long im = __ballot(st == ST_INCL);
        const int h = im ? 63 - __builtin_clzll(im) : -1;
        double r = lane >= h ? v : 0.0;  // lanes [max(h, 0), 63]
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o, 64);
        excl += r;
        if (h >= 0) break;
        wend -= 64;
      }
    }
    if (lane == 0) {
      if (topen) publish(pub, (long long)tile, ST_INCL, excl + tsum, epoch);
      s_carry = excl;
    }
  }
  __syncthreads();
  // an element's prefix = (the tile's carry, if no head precedes it in the
  // tile) + (its sum in A since the tile's start or its last head); in the
  // FP64 form the carry is folded into the sum at the pack's start
  const double carry = s_carry;

#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const int p = (wv * 64 * FJ + j * 64 + lane) * 4;
    const unsigned hb = (s_bits[p >> 5] >> (p & 31)) & 0xFu;
    const A lp = s_lpv[j][threadIdx.x];
    const bool open = (lpo >> j) & 1u;
    A r;
    double c = 0.0;
    if constexpr (__is_same(A, double)) {
      const double wcarry = wopen ? carry + wval : wval;
      r = open ? wcarry + lp : lp;
    } else {
      r = open ? acc_add(wval, lp) : lp;
      c = open && wopen ? carry : 0.0;
    }
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (hb & (1u << k)) {
        r = acc_zero<A>();
        c = 0.0;
      }
      const A ex = r;
      r = acc_add(r, x[j][k]);
      o[k] = (float)(c + acc_to_double(inclusive ? r : ex));
    }
    const unsigned long long e0 = wbase + ((unsigned long long)j * 64 + lane) * 4;
    if (full && (long long)e0 >= lo && (long long)e0 + 4 <= hi) {
      f4u w;
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = o[k];
      __builtin_nontemporal_store(w, reinterpret_cast<f4u*>(out + e0));
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const long long e = (long long)e0 + k;
        if (e >= lo && e < hi) out[e] = o[k];
      }
    }
  }
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

extern "C" int skl_internal_segmented_sum(float* out, const float* in, unsigned long long n,
                                          const int64_t* offsets, int nseg, int fp64, void* ws,
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
  if (fp64)
    seg_sum_kernel<double><<<(unsigned)cap, WG, 0, s>>>(in, offsets, cstart, cseg, partials, ctl);
  else
    seg_sum_kernel<acc_df><<<(unsigned)cap, WG, 0, s>>>(in, offsets, cstart, cseg, partials, ctl);
  seg_combine_kernel<<<(unsigned)((nseg + WG / 64 - 1) / (WG / 64)), WG, 0, s>>>(
      cstart, nseg, partials, out, ctl);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

extern "C" int skl_internal_segmented_scan(float* out, const float* in, unsigned long long n,
                                           const int64_t* offsets, int nseg, int inclusive,
                                           int fp64, void* ws, hipStream_t s)
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
  if (fp64)
    seg_flat_scan_kernel<double><<<(unsigned)ntiles, FWG, 0, s>>>(
        in, out, n, offsets, nseg, tfirst, ticket, pub, inclusive ? 1 : 0, err);
  else
    seg_flat_scan_kernel<acc_df><<<(unsigned)ntiles, FWG, 0, s>>>(
        in, out, n, offsets, nseg, tfirst, ticket, pub, inclusive ? 1 : 0, err);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

// One tile of the flat segmented Scan(+) on float, shared by the segmented
// Scan launch (segmented.hip) and the one-launch batch kernel (batch.hip).
//
// The n elements are cut into tiles of T = NT * FJ * 4, independent of the
// segments.  Element e is a head when e == offsets[s] for some s
// (offsets[nseg] marks the end of the last segment); a scan value restarts
// at every head.  The pair (open, v) -- "no head seen yet", partial sum --
// combines as
//   (o1, v1) (+) (o2, v2) = (o1 && o2, o2 ? v1 + v2 : v2),
// which is associative, so a tile runs the same single-pass decoupled
// look-back as the plain Scan.  A tile that holds a head knows its inclusive
// value at once (the sum after its last head) and publishes it as INCL
// straight away, so a look-back never passes a tile with a head.
//
// Optionally the tile also writes each segment's sum (the Reduce(+) of the
// segment: the inclusive prefix at its last element, carry included, in
// FP64 before the final rounding; 0 for an empty segment) -- one pass
// serves the segmented Reduce and Scan of the same array.  Float tiles sum
// in A (compensated pairs or FP64); int32 tiles (A = acc_u32) sum exactly
// modulo 2^32.
//
// The tile finds the segments it meets either in a table a prologue wrote
// (tfirst) or by searching the offsets itself; in the second form it also
// checks a slice of the offsets (every pair is checked by one tile) and
// reports invalid ones through the error word.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ld_util.h"

namespace segtile {

constexpr unsigned ST_AGG = 1u, ST_INCL = 2u;

// Carries (tile aggregates, look-back sums, segment totals): FP64 for float,
// the accumulator itself for int32.  A float tile adds its carry to each
// output's compensated sum at the end; where the carry and the accumulator
// are one type (FP64, int32) the carry is folded in at each pack's start.
template <typename A>
struct CarryOf {
  typedef double type;
};
template <>
struct CarryOf<acc_u32> {
  typedef acc_u32 type;
};

__device__ __forceinline__ double to_carry(double a) { return a; }
__device__ __forceinline__ double to_carry(const acc_df& a) { return acc_to_double(a); }
__device__ __forceinline__ acc_u32 to_carry(const acc_u32& a) { return a; }
__device__ __forceinline__ unsigned long long carry_bits(double v)
{
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ unsigned long long carry_bits(const acc_u32& v) { return v.v; }
template <typename C>
__device__ __forceinline__ C carry_from_bits(unsigned long long b);
template <>
__device__ __forceinline__ double carry_from_bits<double>(unsigned long long b)
{
  return __longlong_as_double((long long)b);
}
template <>
__device__ __forceinline__ acc_u32 carry_from_bits<acc_u32>(unsigned long long b)
{
  return acc_u32{(unsigned)b};
}
__device__ __forceinline__ double carry_add(double a, double b) { return a + b; }
__device__ __forceinline__ acc_u32 carry_add(const acc_u32& a, const acc_u32& b)
{
  return acc_u32{a.v + b.v};
}
__device__ __forceinline__ double carry_shfl_xor(double v, int o) { return __shfl_xor(v, o, 64); }
__device__ __forceinline__ acc_u32 carry_shfl_xor(const acc_u32& v, int o)
{
  return acc_u32{(unsigned)__shfl_xor((int)v.v, o, 64)};
}
// the stored bits of an output or a segment total
__device__ __forceinline__ float out_value(double v) { return (float)v; }
__device__ __forceinline__ float out_value(const acc_u32& v) { return __uint_as_float(v.v); }

template <typename C>
__device__ __forceinline__ void publish(unsigned long long* pub, long long c, unsigned state,
                                        const C& v, unsigned epoch)
{
  const unsigned long long bits = carry_bits(v);
  const unsigned long long tag = (unsigned long long)((epoch << 2) | state) << 32;
  __hip_atomic_store(&pub[2 * c], tag | (bits & 0xFFFFFFFFull), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&pub[2 * c + 1], tag | (bits >> 32), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// state (AGG / INCL) once both words carry the same tag of this launch, else 0
template <typename C>
__device__ __forceinline__ unsigned read_status(const unsigned long long* pub, long long c,
                                                unsigned epoch, C* v)
{
  const unsigned long long w0 =
      __hip_atomic_load(&pub[2 * c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long w1 =
      __hip_atomic_load(&pub[2 * c + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned t0 = (unsigned)(w0 >> 32), t1 = (unsigned)(w1 >> 32);
  if (t0 != t1 || (t0 >> 2) != epoch) return 0;
  *v = carry_from_bits<C>((w1 << 32) | (w0 & 0xFFFFFFFFull));
  return t0 & 3u;
}

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

// (number of s in [0, N) with off[s] <= pos) - 1 for non-decreasing
// offsets, for two positions at once: an 8-way search, each level eight
// loads per position at wave-uniform addresses (scalar loads, which leave
// the tile's vector loads in flight).  (16-way levels, one fewer at 4097
// offsets, ran slower.)  For any offsets the results stay in
// [-1, N - 1].
__device__ __forceinline__ void upper_index2(const int64_t* __restrict__ off, int N,
                                             long long pos0, long long pos1, int* r0, int* r1)
{
  int b0 = 0, b1 = 0;
  bool done0 = false, done1 = false;
  long long step = 1;
  while (step * 8 < N) step *= 8;
  for (;;) {
    int c0 = 0, c1 = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const long long i0 = b0 + k * step, i1 = b1 + k * step;
      const bool ok0 = i0 < N, ok1 = i1 < N;
      const long long v0 = off[ok0 ? i0 : b0];
      const long long v1 = off[ok1 ? i1 : b1];
      c0 += ok0 && v0 <= pos0 ? 1 : 0;
      c1 += ok1 && v1 <= pos1 ? 1 : 0;
    }
    if (!done0) {
      if (c0 == 0) {
        *r0 = b0 - 1;
        done0 = true;
      } else {
        b0 += (int)((c0 - 1) * step);
      }
    }
    if (!done1) {
      if (c1 == 0) {
        *r1 = b1 - 1;
        done1 = true;
      } else {
        b1 += (int)((c1 - 1) * step);
      }
    }
    if (step == 1) break;
    step /= 8;
  }
  if (!done0) *r0 = b0;
  if (!done1) *r1 = b1;
}

// LDS of one tile.  SLOTS > 0: segment sums are written, the first SLOTS
// of a tile's segment totals through LDS.
template <typename A, int NT, int FJ, int SLOTS>
struct Smem {
  static constexpr int T = NT * FJ * 4;
  static constexpr int NW = T / 32;  // bit words per tile
  unsigned bits[NW];                 // heads
  unsigned tails[SLOTS ? NW : 1];    // last elements of non-empty segments
  int tpre[SLOTS ? NW : 1];          // tails before each word
  A lpv[FJ][NT];                     // per pack: wave-exclusive prefix (own thread only)
  A wval[NT / 64];
  int wopen[NT / 64];
  int wcnt[NT / 64];
  typename CarryOf<A>::type carry;
  typename CarryOf<A>::type slot[SLOTS ? SLOTS : 1];
};

struct Job {
  const float* in;
  float* out;    // the scan, or NULL (segment sums only)
  float* sums;   // the segment sums (SLOTS > 0), or NULL
  unsigned long long n;
  const int64_t* off;  // nseg + 1 offsets
  int nseg;
  const int* tfirst;   // without SEARCH: tfirst[t] = the largest s with off[s] <= t * T (-1 if none)
  unsigned long long* pub;  // two status words per tile
  unsigned* err;       // with SEARCH: invalid offsets store epoch << 2 | bit
  long long ntiles;
  int inclusive;
};

// Clears the tile's bit words; a barrier must follow before tile_run.
template <typename A, int NT, int FJ, int SLOTS>
__device__ __forceinline__ void tile_clear(Smem<A, NT, FJ, SLOTS>& sm)
{
  for (int i = threadIdx.x; i < Smem<A, NT, FJ, SLOTS>::NW; i += NT) {
    sm.bits[i] = 0;
    if constexpr (SLOTS > 0) sm.tails[i] = 0;
  }
}

// Tile `tile` of the launch with epoch `epoch`, its bit words cleared
// (tile_clear, then a barrier).  SEARCH: the tile finds its segments in the
// offsets and checks its slice of them, else it reads tfirst.  Every tile publishes its status whatever it
// finds in the offsets, so no look-back waits forever.
template <typename A, int NT, int FJ, int SLOTS, bool SEARCH>
__device__ __forceinline__ void tile_run(const Job& jb, unsigned long long tile, unsigned epoch,
                                         Smem<A, NT, FJ, SLOTS>& sm)
{
  typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
  constexpr int T = NT * FJ * 4, NWAVES = NT / 64, NWORD = T / 32;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long n = jb.n;
  const int64_t* __restrict__ off = jb.off;
  const int nseg = jb.nseg;

  const unsigned long long tbase = tile * (unsigned long long)T;
  const long long tend = (long long)tbase + T;
  const unsigned long long wbase = tbase + (unsigned long long)wv * (64 * FJ * 4);
  const bool full = tbase + T <= n;

  float x[FJ][4];
  if (full) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      const f4u v = __builtin_nontemporal_load(
          reinterpret_cast<const f4u*>(jb.in + wbase + ((unsigned long long)j * 64 + lane) * 4));
#pragma unroll
      for (int k = 0; k < 4; ++k) x[j][k] = v[k];
    }
  } else {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned long long e = wbase + ((unsigned long long)j * 64 + lane) * 4 + k;
        x[j][k] = ld_or(jb.in, (size_t)e, 0, e < n, 0.0f);
      }
  }

  // the segments the tile meets: s0 .. s1 (tfirst of this tile and the next)
  int s0, s1;
  if constexpr (!SEARCH) {
    s1 = jb.tfirst[tile + 1];
    s0 = jb.tfirst[tile];
  } else {
    upper_index2(off, nseg + 1, (long long)tbase, tend, &s0, &s1);
  }
  // Head bits: offsets[s0 .. s1] inside the tile.  With segment sums, the
  // same pass marks the tails (last elements) of the non-empty segments
  // ending in the tile and writes 0 for the empty segments whose position
  // lies in (tbase, tend] (tile 0: [0, tend]).  A few offsets are read by
  // wave 0 with scalar loads, which leave the data loads above in flight (a
  // vector load issued after them could only be waited for together with
  // them).
  {
    const int sa = SLOTS && tile == 0 ? 0 : (s0 < 0 ? 0 : s0);
    auto visit = [&](int s, bool writer) {
      const long long a = off[s];
      const long long p = a - (long long)tbase;
      if (writer && p >= 0 && p < T) atomicOr(&sm.bits[p >> 5], 1u << (p & 31));
      if constexpr (SLOTS > 0) {
        if (s >= nseg) return;
        const long long b = off[s + 1];
        if (!jb.sums) return;
        if (a == b) {
          if (writer && (a > (long long)tbase || tile == 0) && a <= tend && a >= 0)
            jb.sums[s] = 0.0f;
        } else if (a < b && b > (long long)tbase && b <= tend) {
          const long long q = b - 1 - (long long)tbase;
          if (writer) atomicOr(&sm.tails[q >> 5], 1u << (q & 31));
        }
      }
    };
    if (s1 - sa < 16) {
      if (wv == 0)
        for (int s = sa; s <= s1; ++s) visit(s, lane == 0);
    } else {
      for (int s = sa + (int)threadIdx.x; s <= s1; s += NT) visit(s, true);
    }
  }
  __syncthreads();
  // outputs go to [lo, hi), inside [0, n) whatever the offsets hold
  const long long lo0 = off[0], hi0 = off[nseg];
  const long long lo = lo0 < 0 ? 0 : lo0;
  const long long hi = hi0 > (long long)n ? (long long)n : hi0;

  // tail ordinals: tpre[w] = tails in words [0, w)
  if constexpr (SLOTS > 0) {
    constexpr int WPT = NWORD / NT;  // words per thread
    static_assert(NWORD % NT == 0, "tail words per thread");
    int c = 0;
#pragma unroll
    for (int q = 0; q < WPT; ++q) c += __popc(sm.tails[threadIdx.x * WPT + q]);
    int v = c;  // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(v, o, 64);
      if (lane >= o) v += t;
    }
    if (lane == 63) sm.wcnt[wv] = v;
    __syncthreads();
    int base = v - c;
#pragma unroll
    for (int w = 0; w < NWAVES; ++w)
      if (w < wv) base += sm.wcnt[w];
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
      sm.tpre[threadIdx.x * WPT + q] = base;
      base += __popc(sm.tails[threadIdx.x * WPT + q]);
    }
  }

  // per pack: (open, sum after its last head); one segmented wave scan per
  // row of 64 packs
  unsigned lpo = 0;  // bit j: pack j's wave-exclusive prefix is open
  bool ropen = true;
  A rval = acc_zero<A>();
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const int p = (wv * 64 * FJ + j * 64 + lane) * 4;
    const unsigned hb = (sm.bits[p >> 5] >> (p & 31)) & 0xFu;
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
    sm.lpv[j][threadIdx.x] = eopen ? acc_add(rval, ex) : ex;
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
    sm.wval[wv] = rval;
    sm.wopen[wv] = ropen ? 1 : 0;
  }
  __syncthreads();
  bool wopen = true, topen = true;
  A wval = acc_zero<A>(), tval = acc_zero<A>();
#pragma unroll
  for (int w = 0; w < NWAVES; ++w) {
    const A v = sm.wval[w];
    const bool o = sm.wopen[w] != 0;
    if (w < wv) {
      wval = o ? acc_add(wval, v) : v;
      wopen = wopen && o;
    }
    tval = o ? acc_add(tval, v) : v;
    topen = topen && o;
  }
  typedef typename CarryOf<A>::type C;
  constexpr bool FOLD = __is_same(A, C);  // the carry joins each pack's sum
  const C tsum = to_carry(tval);

  // look-back (wave 0); a tile whose first element is a head needs no carry
  if (wv == 0) {
    const bool need = tile > 0 && (sm.bits[0] & 1u) == 0;
    if (lane == 0) publish(jb.pub, (long long)tile, topen ? ST_AGG : ST_INCL, tsum, epoch);
    C excl = carry_from_bits<C>(0);
    if (need) {
      long long wend = (long long)tile;  // window [wend - 64, wend)
      for (;;) {
        const long long idx = wend - 64 + lane;
        unsigned st = ST_INCL;
        C v = carry_from_bits<C>(0);
        if (idx >= 0) {
          do {
            st = read_status(jb.pub, idx, epoch, &v);
          } while (st == 0);
        }
        const unsigned long long im = __ballot(st == ST_INCL);
        const int h = im ? 63 - __builtin_clzll(im) : -1;
        C r = lane >= h ? v : carry_from_bits<C>(0);  // lanes [max(h, 0), 63]
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) r = carry_add(r, carry_shfl_xor(r, o));
        excl = carry_add(excl, r);
        if (h >= 0) break;
        wend -= 64;
      }
    }
    if (lane == 0) {
      if (topen) publish(jb.pub, (long long)tile, ST_INCL, carry_add(excl, tsum), epoch);
      sm.carry = excl;
    }
  }
  __syncthreads();
  // an element's prefix = (the tile's carry, if no head precedes it in the
  // tile) + (its sum in A since the tile's start or its last head); in the
  // FP64 form the carry is folded into the sum at the pack's start
  const C carry = sm.carry;

  // The outputs, and the segment totals at the tails: the first SLOTS of
  // them (in element order) through LDS slots to the segment side below;
  // the rest (a tile with more segment ends than slots) straight to sums[s],
  // s found by searching the offsets.
  const int sa = s0 < 0 ? 0 : s0;
  const int sb = s1 < nseg ? s1 : nseg;  // segments [sa, sb) may end here
  const bool store = jb.out != nullptr;
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const int p = (wv * 64 * FJ + j * 64 + lane) * 4;
    const unsigned hb = (sm.bits[p >> 5] >> (p & 31)) & 0xFu;
    const A lp = sm.lpv[j][threadIdx.x];
    const bool open = (lpo >> j) & 1u;
    A r;
    double c = 0.0;  // the carry still to add (compensated float form)
    if constexpr (FOLD) {
      const A wcarry = wopen ? acc_add(carry, wval) : wval;
      r = open ? acc_add(wcarry, lp) : lp;
    } else {
      r = open ? acc_add(wval, lp) : lp;
      c = open && wopen ? carry : 0.0;
    }
    unsigned tb = 0;
    int tord = 0;
    if constexpr (SLOTS > 0) {
      tb = (sm.tails[p >> 5] >> (p & 31)) & 0xFu;
      tord = sm.tpre[p >> 5] + __popc(sm.tails[p >> 5] & ((1u << (p & 31)) - 1u));
    }
    const unsigned long long e0 = wbase + ((unsigned long long)j * 64 + lane) * 4;
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (hb & (1u << k)) {
        r = acc_zero<A>();
        c = 0.0;
      }
      const A ex = r;
      r = acc_add(r, x[j][k]);
      if constexpr (FOLD)
        o[k] = out_value(jb.inclusive ? r : ex);
      else
        o[k] = (float)(c + acc_to_double(jb.inclusive ? r : ex));
      if constexpr (SLOTS > 0) {
        if (tb & (1u << k)) {
          C tot;
          if constexpr (FOLD) tot = r;
          else tot = c + acc_to_double(r);
          if (tord < SLOTS) {
            sm.slot[tord] = tot;
          } else {
            const long long e = (long long)e0 + k;
            int l = sa, h = sb;  // the last s in [sa, sb) with off[s] <= e
            while (h - l > 1) {
              const int m = (l + h) >> 1;
              if (off[m] <= e) l = m;
              else h = m;
            }
            jb.sums[l] = out_value(tot);
          }
          ++tord;
        }
      }
    }
    if (store) {
      if (full && (long long)e0 >= lo && (long long)e0 + 4 <= hi) {
        f4u w;
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = o[k];
        __builtin_nontemporal_store(w, reinterpret_cast<f4u*>(jb.out + e0));
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const long long e = (long long)e0 + k;
          if (e >= lo && e < hi) jb.out[e] = o[k];
        }
      }
    }
  }

  if constexpr (SLOTS > 0) {
    if (jb.sums) {
      __syncthreads();
      // each segment ending in the tile with a slot: its total from the slot
      auto take = [&](int s, bool writer) {
        const long long a = off[s], b = off[s + 1];
        if (!(a < b && b > (long long)tbase && b <= tend)) return;
        const int q = (int)(b - 1 - (long long)tbase);
        const int ord = sm.tpre[q >> 5] + __popc(sm.tails[q >> 5] & ((1u << (q & 31)) - 1u));
        if (writer && ord < SLOTS) jb.sums[s] = out_value(sm.slot[ord]);
      };
      if (sb - sa <= 16) {
        if (wv == 0)
          for (int s = sa; s < sb; ++s) take(s, lane == 0);
      } else {
        for (int s = sa + (int)threadIdx.x; s < sb; s += NT) take(s, true);
      }
    }
  }

  // Without a prologue the tiles check the offsets: tile t the pairs
  // [t * per, (t + 1) * per), tile 0 also the ends.  A failure is recorded
  // for this launch's epoch; the outputs of the launch are then unspecified.
  if constexpr (SEARCH) {
    const long long per = (nseg + jb.ntiles - 1) / jb.ntiles;
    const long long p0 = (long long)tile * per;
    const long long p1 = p0 + per < nseg ? p0 + per : nseg;
    unsigned bad = 0;
    if (p1 - p0 <= 16) {  // every thread the same pairs: scalar loads; thread 0 reports
      for (long long s = p0; s < p1; ++s)
        if (off[s] > off[s + 1]) bad = 1u;
      if (threadIdx.x != 0) bad = 0;
    } else {
      for (long long s = p0 + threadIdx.x; s < p1; s += NT)
        if (off[s] > off[s + 1]) bad = 1u;
    }
    if (tile == 0 && threadIdx.x == 0 && (off[0] < 0 || off[nseg] < 0 || (unsigned long long)off[nseg] > n))
      bad |= 2u;
    if (bad)
      __hip_atomic_store(jb.err, (epoch << 2) | bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace segtile

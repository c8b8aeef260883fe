// Dense C = A.B on float (the AllPairs(Reduce(+), Zip(*)) contraction,
// AllPairsKernel.cl:47-135) in three precision forms, built with hipcc for
// gfx950:
//
//  FP32    VALU, LDS-tiled: 256 threads, 128x128 block tile, 8x8 outputs per
//          thread, k-slices of 8; fmaf accumulation in k order per output.
//  BF16X3  MFMA v_mfma_f32_16x16x32_bf16.  Each operand is split into three
//          bf16 terms x = x0 + x1 + x2 (x1 = bf16(x - x0), x2 = bf16(x - x0 -
//          x1)), stored K-contiguous (B transposed) and zero-padded to the
//          tile grid by a split kernel; the GEMM kernel sums the six products
//          with i+j <= 2 (x0y0 into one FP32 accumulator, the five correction
//          products into a second one) and adds the two at the end.
//  FP8     MFMA v_mfma_scale_f32_16x16x128_f8f6f4 on OCP e4m3 operands, each
//          128x128 tile of A and of B^T quantised with its own power-of-two
//          scale (the smallest >= absmax/448); the scales go to the MFMA as
//          its E8M0 block scales, so it accumulates directly in FP32.
//
// Block tile 128x128, 256 threads = 4 waves as 2x2, each wave 64x64 = 4x4
// MFMA tiles of 16x16.  blockIdx.x walks N, blockIdx.y walks M.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "skelcl_skeletons.h"
#include "ld_util.h"
#include "gemm_tile.h"

namespace {

typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));


// ---------------------------------------------------------------------------
// bf16 round-to-nearest-even of a float (finite inputs; NaN stays NaN)
__device__ __forceinline__ uint16_t bf16_rne(float x)
{
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Split into three bf16 terms, each the rounding of what is left.
__device__ __forceinline__ void split3(float x, uint16_t& h0, uint16_t& h1, uint16_t& h2)
{
  h0 = bf16_rne(x);
  float r1 = x - bf16_to_f32(h0);
  h1 = bf16_rne(r1);
  float r2 = r1 - bf16_to_f32(h1);
  h2 = bf16_rne(r2);
}

// ---------------------------------------------------------------------------
// Tiled plane layout of the bf16x3 operands.  An operand X (R x K, read K
// contiguous; R padded to Rp % 128 == 0, K to Kp % 64 == 0, zeros) is stored
// as blocks of 128 rows x 32 k, one per (row tile t, k-slice s, plane p), at
// block index (t * (Kp / 32) + s) * 3 + p, 4096 bf16 each.  Inside a block,
// row r holds its 32 k as four 16-byte chunks, chunk c at position
// c ^ swz(r).  A block is the exact LDS image the GEMM kernel multiplies
// from, so one k-slice of one row tile (three planes) is a contiguous 24 KB
// run that global_load_lds copies in 1 KB pieces.  swz spreads the 16 rows
// of an MFMA fragment read over 16 distinct 16-byte bank groups (rows r,
// r+2, r+4, r+6 and r, r+4, r+8, r+12 differ in swz), for 32 or 64 banks.
constexpr int BK = 32;          // k per slice
constexpr int SLICE = BM * BK;  // bf16 per block
__device__ __forceinline__ int swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// Offset (bf16) of (row r, k) in plane 0; for k % 8 == 0 it starts an
// aligned 8-element chunk.
__device__ __forceinline__ size_t tiled_offset(int r, int k, int nk)
{
  const int rr = r & (BM - 1), kk = k & (BK - 1);
  return ((size_t)(r >> 7) * nk + (k >> 5)) * 3 * SLICE + rr * BK +
         ((((kk >> 3) ^ swz(rr)) << 3) | (kk & 7));
}

// Split eight values and store the three 16-byte chunks at out + o + p*SLICE.
__device__ __forceinline__ void split_store8(const float (&x)[8], uint16_t* out, size_t o)
{
  uint16_t h[3][8];
#pragma unroll
  for (int i = 0; i < 8; ++i) split3(x[i], h[0][i], h[1][i], h[2][i]);
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    i32x4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (int)((uint32_t)h[p][2 * i] | ((uint32_t)h[p][2 * i + 1] << 16));
    *reinterpret_cast<i32x4*>(out + o + (size_t)p * SLICE) = v;
  }
}

// X row-major R x Kc (ld = Kc) into the tiled planes, one thread per 8-k chunk.
// `vec`: X is 16-byte aligned and Kc % 4 == 0, so each half of a chunk is one
// float4 that is wholly inside or wholly outside the matrix.
__global__ __launch_bounds__(256) void split3_kernel(const float* __restrict__ X, int R, int Kc,
                                                     int vec, uint16_t* __restrict__ out,
                                                     int Rp, int Kp)
{
  const int nk = Kp / BK, cpr = Kp / 8;
  const size_t total = (size_t)Rp * cpr;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (size_t)gridDim.x * 256) {
    const int r = (int)(e / cpr), k0 = (int)(e - (size_t)r * cpr) * 8;
    const bool rok = r < R;
    const size_t base = (size_t)(rok ? r : 0) * Kc + k0;
    float x[8];
    if (vec) {
      const bool ok0 = rok && k0 + 4 <= Kc, ok1 = rok && k0 + 8 <= Kc;
      const f32x4* X4 = reinterpret_cast<const f32x4*>(X);
      const f32x4 a = as_global(X4)[ok0 ? base / 4 : 0];
      const f32x4 b = as_global(X4)[ok1 ? base / 4 + 1 : 0];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        x[i] = ok0 ? a[i] : 0.0f;
        x[4 + i] = ok1 ? b[i] : 0.0f;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) x[i] = ld_or(X, base + i, 0, rok && k0 + i < Kc, 0.0f);
    }
    split_store8(x, out, tiled_offset(r, k0, nk));
  }
}

// X row-major Kc x R (ld = R), read as its transpose (element (r, k) =
// X[k][r]), into the tiled planes: 64 r x 32 k per work-group through LDS,
// reads coalesced along r, one thread per 8-k chunk on the way out.
__global__ __launch_bounds__(256) void split3_transpose_kernel(const float* __restrict__ X,
                                                               int R, int Kc,
                                                               uint16_t* __restrict__ out,
                                                               int Rp, int Kp)
{
  __shared__ float t[32][65];
  const int nk = Kp / BK;
  const int r0 = blockIdx.x * 64, k0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
#pragma unroll
  for (int kk = ty; kk < 32; kk += 4) {
    const int k = k0 + kk, r = r0 + tx;
    t[kk][tx] = ld_or(X, (size_t)k * R + r, 0, k < Kc && r < R, 0.0f);
  }
  __syncthreads();
  const int r = threadIdx.x >> 2, q = threadIdx.x & 3;  // 64 rows x 4 chunks
  float x[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = t[q * 8 + i][r];
  split_store8(x, out, tiled_offset(r0 + r, k0 + q * 8, nk));
}

// ---------------------------------------------------------------------------
// XCD-aware tile order.  Work-groups are dispatched round-robin over the 8
// XCDs (work-group b runs on XCD b % 8), each with its own L2.  Linear
// work-group b is given tile xcd * (T / 8) + b / 8, so each XCD walks a
// contiguous range of tiles; inside a range tiles go 8 M-tiles at a time
// (column-major inside the group), so the ~32 tiles an XCD runs at once share
// their A row-panels and B column-panels in its L2.  T % 8 != 0 keeps the
// plain order (the map must be a bijection).
__device__ __forceinline__ void tile_coords(int& tm, int& tn)
{
  const int ntn = (int)gridDim.x, ntm = (int)gridDim.y;
  const int T = ntm * ntn;
  int t = (int)(blockIdx.y * gridDim.x + blockIdx.x);
  if ((T & 7) == 0) t = (t & 7) * (T >> 3) + (t >> 3);
  constexpr int GM = 8;
  const int group = t / (GM * ntn);
  const int first = group * GM;
  const int gsize = ntm - first < GM ? ntm - first : GM;
  const int r = t - group * GM * ntn;
  tm = first + r % gsize;
  tn = r / gsize;
}

// ---------------------------------------------------------------------------
// bf16x3 GEMM on the tiled planes of A (Mp rows) and B^T (Np rows).

// Six bf16 products of one k-step, ordered so that consecutive MFMAs write
// different accumulators: each product type sweeps all 16 tiles before the
// next, so a dependent MFMA is 16 instructions behind its predecessor.
#define SKL_BF16X3_STEP(acc0, acc1, a, b)                                           \
  do {                                                                               \
    _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_)                                 \
      _Pragma("unroll") for (int j_ = 0; j_ < 4; ++j_)                               \
        acc0[i_][j_] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i_][0], b[j_][0],   \
                                                               acc0[i_][j_], 0, 0, 0); \
    constexpr int pa_[5] = {1, 0, 2, 0, 1}, pb_[5] = {1, 2, 0, 1, 0};                \
    _Pragma("unroll") for (int t_ = 0; t_ < 5; ++t_)                                 \
      _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_)                               \
        _Pragma("unroll") for (int j_ = 0; j_ < 4; ++j_)                             \
          acc1[i_][j_] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(                    \
              a[i_][pa_[t_]], b[j_][pb_[t_]], acc1[i_][j_], 0, 0, 0);                \
  } while (0)

typedef __attribute__((address_space(3))) void lds_ptr_t;

// Fragments of one k-slice: a[i][p] is plane p of the 16-row A tile i, b
// likewise for B^T.
struct Bf16x3Frag {
  bf16x8 a[4][3], b[4][3];
};

__device__ __forceinline__ void bf16x3_read(Bf16x3Frag& f, const uint16_t* As,
                                            const uint16_t* Bs, int p, int arow, int brow,
                                            int pos)
{
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    f.a[i][p] = *reinterpret_cast<const bf16x8*>(As + p * SLICE + (arow + i * 16) * BK + pos);
    f.b[i][p] = *reinterpret_cast<const bf16x8*>(Bs + p * SLICE + (brow + i * 16) * BK + pos);
  }
}

__device__ __forceinline__ void bf16x3_mm(f32x4 (&acc)[4][4], const Bf16x3Frag& f, int pa, int pb)
{
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(f.a[i][pa], f.b[j][pb], acc[i][j], 0, 0, 0);
}

// MB 128-row blocks of A per work-group tile (MB x 128 rows by 128
// columns), 4 x MB waves of 64 x 64 (2 x MB down, 2 across).
template <int MB>
__global__ __launch_bounds__(256 * MB, 1) void gemm_bf16x3_kernel(
    const uint16_t* __restrict__ Ap, const uint16_t* __restrict__ Bp, float* __restrict__ C,
    int M, int N, int Kp, int ldc)
{
  // Two slice buffers (the tiled blocks of A and B^T), filled by
  // global_load_lds; separate objects, so that the compiler's wait for the
  // fill of one does not block reads of the other.
  __shared__ __attribute__((aligned(1024))) uint16_t As0[MB * 3 * SLICE], Bs0[3 * SLICE];
  __shared__ __attribute__((aligned(1024))) uint16_t As1[MB * 3 * SLICE], Bs1[3 * SLICE];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  int tm, tn;
  tile_coords(tm, tn);
  const int m0 = tm * (MB * BM), n0 = tn * BN;
  const int nk = Kp / BK;

  f32x4 acc0[4][4], acc1[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc0[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      acc1[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    }

  // Fill of slice s: per 128-row block the 24 KB run of its three planes, 24
  // pieces of 1 KB; each wave copies 6 pieces of A and 6 / MB of B; lane l
  // copies bytes 16 l .. 16 l + 15 of a piece.  Address = wave-uniform base +
  // 32-bit lane offset (saddr form).  Slices past the end are clamped to the
  // last one: a harmless reload into a buffer that is not read again.
  constexpr size_t RUN = 3 * SLICE * 2;  // bytes per block and slice
  const char* ga = reinterpret_cast<const char*>(Ap) + (size_t)tm * MB * nk * RUN;
  const char* gb = reinterpret_cast<const char*>(Bp) + (size_t)tn * nk * RUN;
  const unsigned loff = (unsigned)lane * 16u;
  auto fill = [&](uint16_t* As, uint16_t* Bs, int sl) {
    const size_t so = (size_t)(sl < nk ? sl : nk - 1) * RUN;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int piece = wave * 6 + j;  // 0 .. 24 MB - 1
      const int blk = piece / 24, q = piece % 24;
      __builtin_amdgcn_global_load_lds(ga + (size_t)blk * nk * RUN + so + q * 1024 + loff,
                                       (lds_ptr_t*)(As + piece * 512), 16, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 6 / MB; ++j) {
      const int piece = wave * (6 / MB) + j;  // 0 .. 23
      __builtin_amdgcn_global_load_lds(gb + so + piece * 1024 + loff,
                                       (lds_ptr_t*)(Bs + piece * 512), 16, 0, 0);
    }
  };

  // MFMA fragment: lane holds row lane % 16, k-chunk lane / 16 of a 16-row
  // tile; this wave's A rows lie in block wm / 2 of the tile
  const int fr = lane & 15;
  const int pos = (((lane >> 4) ^ swz(fr)) * 8);
  const int ablk = (wm >> 1) * 3 * SLICE;
  const int arow = (wm & 1) * 64 + fr, brow = wn * 64 + fr;

  // Slice k from (Ac, Bc) with plane 0 already in f; (An, Bn) receives
  // slice k+1 (its fill was issued one slice earlier).
  //  1. x0y0 with the plane-1 reads interleaved, x1y0 + x0y1 with the
  //     plane-2 reads interleaved, x1y1;
  //  2. wait for this wave's fill of slice k+1 and its LDS reads, barrier:
  //     slice k+1 is in LDS and nobody reads (Ac, Bc) any more;
  //  3. fill slice k+2 into (Ac, Bc); plane 0 of slice k+1 into g,
  //     interleaved with x2y0 + x0y2.
  // A fill thus has one slice of MFMA work to arrive.  (The compiler waits
  // for every outstanding fill before any LDS read that may alias one, so
  // fills further ahead would be waited for early.)  Scheduling barriers pin
  // the phases; group barriers (mask 0x008 MFMA, 0x100 LDS read) interleave
  // the reads with the MFMAs.
  auto slice = [&](uint16_t* Ac, uint16_t* Bc, const uint16_t* An, const uint16_t* Bn, int k,
                   Bf16x3Frag& f, Bf16x3Frag& g) {
    bf16x3_mm(acc0, f, 0, 0);
    bf16x3_read(f, Ac + ablk, Bc, 1, arow, brow, pos);
    bf16x3_mm(acc1, f, 1, 0);
    bf16x3_mm(acc1, f, 0, 1);
    bf16x3_read(f, Ac + ablk, Bc, 2, arow, brow, pos);
    bf16x3_mm(acc1, f, 1, 1);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    fill(Ac, Bc, k + 2);
    __builtin_amdgcn_sched_barrier(0);
    bf16x3_read(g, An + ablk, Bn, 0, arow, brow, pos);
    bf16x3_mm(acc1, f, 2, 0);
    bf16x3_mm(acc1, f, 0, 2);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  };

  Bf16x3Frag f0, f1;
  fill(As0, Bs0, 0);
  fill(As1, Bs1, 1);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  bf16x3_read(f0, As0 + ablk, Bs0, 0, arow, brow, pos);
  // nk is even (Kp % 64 == 0)
  for (int k = 0; k < nk; k += 2) {
    slice(As0, Bs0, As1, Bs1, k, f0, f1);
    slice(As1, Bs1, As0, Bs0, k + 1, f1, f0);
  }
  // no fill may still be writing LDS when the work-group ends
  __builtin_amdgcn_s_waitcnt(0);
  // C fragment of a 16x16 tile: lane holds column lane%16, rows 4*(lane/16)+0..3.
  // Non-temporal stores, as in the FP8 product below
  const int cc = lane & 15, cr = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + cc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + cr + r;
        if (row < M && col < N)
          __builtin_nontemporal_store(acc0[i][j][r] + acc1[i][j][r], C + (size_t)row * ldc + col);
      }
    }
}

// ---------------------------------------------------------------------------
// FP8 (OCP e4m3) with one scale per 128x128 tile.
// Quantise X (R x C view, optionally transposed) into Rp x Cp bytes,
// Rp % 128 == 0, Cp % 128 == 0; scales[(r/128) * (Cp/128) + c/128].
// Tiled FP8 operand layout: blocks of 128 rows x 128 k-bytes (16 KB), block
// index (row tile) * (Kp / 128) + (k block), so one k-block of one row tile
// is a contiguous 16 KB run that global_load_lds copies in 1 KB pieces.
// Inside a block a row is eight 16-byte chunks, chunk c at position
// c ^ swz8(r): the 16 lanes of an MFMA fragment read (rows r0 .. r0+15, the
// same chunk) then fall in 16 distinct 16-byte bank groups, for 32 or 64
// banks (swz8 is a bijection on rows 0-7 and on the even and the odd rows of
// 0-15).  It depends on bits 0-3 of the row only, so every 16-row fragment
// of a block sees the same pattern.
__device__ __forceinline__ int swz8(int r) { return (r ^ ((r >> 3) & 1)) & 7; }

// One FP8 operand to quantise: X is R x Cc row-major (trans = 0), or the
// transpose of the stored matrix (trans = 1: X^T is read from a Cc x R
// array); out / scales are the tiled bytes and the tile scales of an
// Rp x Cp (multiples of 128) operand.
struct QuantJob {
  const float* X;
  int R, Cc, trans, vec;
  uint8_t* out;
  float* scales;
  int Rp, Cp;
};

// One 128x128 tile per work-group; blockIdx.z picks the operand (both in one
// launch).  The tile is read once into registers, 16 bytes of X's
// contiguous dimension per load; absmax (wave shuffles, one barrier); a
// power-of-two scale s >= absmax/448 (stored as a float whose exponent
// field is the E8M0 scale the MFMA takes); e4m3 bytes by the hardware
// conversion (v_cvt_pk_fp8_f32, round to nearest even; |x / s| <= 448 stays
// in range), four to a dword, staged in LDS rows of the output and written
// row-wise in 16-byte chunks.  For trans each thread holds 4 x 4 blocks (4
// lines of 4 elements) and transposes them in registers, so its LDS stores
// are whole dwords; the 132-byte LDS row pitch spreads a wave's stores over
// the banks.  `vec`: X is 16-byte aligned and its row length a multiple of
// 4, so in-bounds quads load as one float4.
constexpr int QP = 132;  // LDS bytes per staged row (dword aligned)
__global__ __launch_bounds__(256) void quant_fp8_kernel(QuantJob ja, QuantJob jb)
{
  const QuantJob j = blockIdx.z ? jb : ja;
  // consecutive work-groups take neighbouring tiles along X's rows: for X
  // itself tc (its columns) varies fastest, for the transposed operand tr
  // (X's columns there).  At 8192^2 the transposed operand took 71.6 us with
  // the other order and takes 58 us with this one.
  const int L = blockIdx.y * gridDim.x + blockIdx.x;
  const int tr = j.trans ? L % gridDim.y : blockIdx.y;
  const int tc = j.trans ? L / gridDim.y : blockIdx.x;
  if (tr * 128 >= j.Rp || tc * 128 >= j.Cp) return;  // the grid covers the larger operand
  const int r0 = tr * 128, c0 = tc * 128;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ float red[4];
  __shared__ __attribute__((aligned(16))) uint8_t qt[128 * QP];
  f32x4 v[16];
  float amax = 0.0f;
  const float* X = j.X;
  const size_t ld = j.trans ? (size_t)j.R : (size_t)j.Cc;
  const int nl = j.trans ? j.Cc : j.R, nc = j.trans ? j.R : j.Cc;  // lines, length
  const int l0 = j.trans ? c0 : r0, e0 = j.trans ? r0 : c0;        // tile origin
  // element quad u: line l, elements e .. e + 3.  Not trans: 32 quads per
  // line, 128 lines.  Trans: u = 4 g + r is line 4 kq + r of block g, whose
  // quad index along the line is nq (lanes: nq = lane % 32, so a wave reads
  // two whole 512-byte lines per load).
  auto quad = [&](int u, int& l, int& e) {
    if (!j.trans) {
      const int q = threadIdx.x + 256 * u;
      l = l0 + (q >> 5);
      e = e0 + (q & 31) * 4;
    } else {
      const int g = u >> 2, r = u & 3;
      const int kq = (lane >> 5) + 2 * wv + 8 * g, nq = lane & 31;
      l = l0 + 4 * kq + r;
      e = e0 + 4 * nq;
    }
  };
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    int l, e;
    quad(u, l, e);
    const bool lok = l < nl;
    const size_t base = (size_t)(lok ? l : 0) * ld;
    if (j.vec && lok && e + 3 < nc) {
      // read once: a non-temporal load (the float operands need no room in
      // L2 or the MALL; 124 against 168 us at 8192^2 per operand pair)
      v[u] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(X + base + e));
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) v[u][i] = ld_or_g(X, base + e + i, 0, lok && e + i < nc, 0.0f);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) amax = fmaxf(amax, fabsf(v[u][i]));
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
  if (lane == 0) red[wv] = amax;
  __syncthreads();
  amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  int ex = 0;
  if (amax > 0.0f) frexpf(amax / 448.0f, &ex);  // absmax/448 = m * 2^ex, m in [0.5, 1)
  // 2^-127 is the smallest E8M0 scale (exponent field 0 of the float
  // below); tinier tiles flush towards zero in e4m3 instead of making
  // 1/scale overflow
  ex = ex < -127 ? -127 : ex;
  const float scale = ldexpf(1.0f, ex);
  const float inv = ldexpf(1.0f, -ex);
  if (threadIdx.x == 0) j.scales[tr * (j.Cp / 128) + tc] = scale;
  if (!j.trans) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int q = threadIdx.x + 256 * u;
      int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[u][0] * inv, v[u][1] * inv, 0, false);
      w = __builtin_amdgcn_cvt_pk_fp8_f32(v[u][2] * inv, v[u][3] * inv, w, true);
      *reinterpret_cast<int*>(&qt[(q >> 5) * QP + (q & 31) * 4]) = w;
    }
  } else {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int kq = (lane >> 5) + 2 * wv + 8 * g, nq = lane & 31;
#pragma unroll
      for (int i = 0; i < 4; ++i) {  // output row 4 nq + i: lines 4 kq .. 4 kq + 3
        int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * g][i] * inv, v[4 * g + 1][i] * inv, 0,
                                                false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * g + 2][i] * inv, v[4 * g + 3][i] * inv, w,
                                            true);
        *reinterpret_cast<int*>(&qt[(4 * nq + i) * QP + 4 * kq]) = w;
      }
    }
  }
  __syncthreads();
  // tiled output: this tile is block tr * (Cp / 128) + tc, rows of 128
  // bytes, chunk kc of row r at position kc ^ swz8(r)
  uint8_t* blk = j.out + ((size_t)tr * (j.Cp / 128) + tc) * (128 * 128);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = threadIdx.x + 256 * u;  // 128 rows x 8 chunks of 16 bytes
    const int lr = q >> 3, kc = q & 7;
    const int* src = reinterpret_cast<const int*>(&qt[lr * QP + kc * 16]);
    *reinterpret_cast<i32x4*>(blk + lr * 128 + ((kc ^ swz8(lr)) << 4)) =
        i32x4{src[0], src[1], src[2], src[3]};
  }
}

// FP8 product on the tiled operands: Aq (Mp/128 row blocks), Bq (B^T,
// Np/128 row blocks); sa: (Mp/128) x (Kp/128), sb: (Np/128) x (Kp/128).
// Work-group tile MB x 128 rows by NB x 128 columns; waves of WR x 16 rows by
// 64 columns (WR 16-row MFMA tiles, 4 16-column ones).  One 128-k block per
// slice, two LDS buffers filled by global_load_lds: after the barrier that
// starts slice k the fill of slice k+1 is issued into the buffer slice k-1
// read, so it has one slice of MFMA work to arrive.  With 8-row waves the
// fill's pieces are issued two after each of the slice's first rows of
// MFMAs rather than all at its start: the fill's LDS writes then spread over
// the slice instead of stalling every wave's first LDS reads at once (at
// 8192^3, 487 against 510 us).
template <int MB, int NB, int WR>
__global__ __launch_bounds__(MB * NB * 128 * 128 / (WR * 16 * 64) * 64, 1) void gemm_fp8_kernel(
    const uint8_t* __restrict__ Aq, const uint8_t* __restrict__ Bq,
    const float* __restrict__ sa, const float* __restrict__ sb, float* __restrict__ C, int M,
    int N, int Kp, int ldc)
{
  constexpr int KB = 128;
  constexpr int BLK = 128 * KB;                    // bytes per block
  constexpr int WAVES_N = NB * 2;                  // 64 columns per wave
  constexpr int WAVES = (MB * 128 / (WR * 16)) * WAVES_N;
  constexpr int PA = MB * (BLK / 1024) / WAVES;  // 1 KB pieces of A per wave per slice
  constexpr int PB = NB * (BLK / 1024) / WAVES;  // and of B
  static_assert(PA * WAVES == MB * (BLK / 1024) && PB * WAVES == NB * (BLK / 1024),
                "fill pieces must split evenly over the waves");
  // one named array per buffer: the compiler then knows that a fill of one
  // buffer does not write what the LDS reads of the other read, and waits
  // for the fills only where the code says so
  __shared__ __attribute__((aligned(1024))) uint8_t As0[MB * BLK], Bs0[NB * BLK];
  __shared__ __attribute__((aligned(1024))) uint8_t As1[MB * BLK], Bs1[NB * BLK];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave / WAVES_N, wn = wave % WAVES_N;
  int tm, tn;
  tile_coords(tm, tn);
  const int m0 = tm * (MB * 128), n0 = tn * (NB * 128);
  const int nkb = Kp / KB;
  f32x4 acc[WR][4];
#pragma unroll
  for (int i = 0; i < WR; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // fill of k-block s: the 16 MB pieces of 1 KB of A (block piece / 16) and
  // the 16 NB of B, PA and PB per wave; lane l copies bytes 16 l .. 16 l + 15
  // of a piece.  Slices past the end are clamped to the last one.
  const unsigned loff = (unsigned)lane * 16u;
  // piece p of this wave's part: p < PA is A piece wave * PA + p, the others
  // B piece wave * PB + p - PA
  auto fill_piece = [&](uint8_t* As, uint8_t* Bs, int sl, int p) {
    const size_t so = (size_t)(sl < nkb ? sl : nkb - 1) * BLK;
    if (p < PA) {
      const int piece = wave * PA + p, b = piece >> 4, q = piece & 15;
      __builtin_amdgcn_global_load_lds(
          Aq + ((size_t)(tm * MB + b) * nkb) * BLK + so + q * 1024 + loff,
          (lds_ptr_t*)(As + piece * 1024), 16, 0, 0);
    } else {
      const int piece = wave * PB + p - PA, b = piece >> 4, q = piece & 15;
      __builtin_amdgcn_global_load_lds(
          Bq + ((size_t)(tn * NB + b) * nkb) * BLK + so + q * 1024 + loff,
          (lds_ptr_t*)(Bs + piece * 1024), 16, 0, 0);
    }
  };
  auto fill = [&](uint8_t* As, uint8_t* Bs, int sl) {
#pragma unroll
    for (int p = 0; p < PA + PB; ++p) fill_piece(As, Bs, sl, p);
  };
  constexpr bool SPREAD = WR == 8;  // fill pieces issued between the MFMA rows

  // fragment: lane holds row lane % 16, k bytes 32 (lane / 16) .. + 31, i.e.
  // chunks 2 (lane / 16) and 2 (lane / 16) + 1 of that row
  const int fr = lane & 15, c0 = (lane >> 4) * 2;
  const int sw = swz8(fr);
  const int p0 = ((c0 ^ sw) << 4), p1 = (((c0 + 1) ^ sw) << 4);
  const int arow0 = wm * (WR * 16);                // within the A tile
  const int bblk = (wn >> 1) * BLK, brow0 = (wn & 1) * 64;
  auto rdA = [&](const uint8_t* As, int i) {
    const int r = arow0 + i * 16 + fr;
    const uint8_t* row = As + (r >> 7) * BLK + (r & 127) * 128;
    const i32x4 lo = *reinterpret_cast<const i32x4*>(row + p0);
    const i32x4 hi = *reinterpret_cast<const i32x4*>(row + p1);
    return i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  };
  auto rdB = [&](const uint8_t* Bs, int j) {
    const uint8_t* row = Bs + bblk + (brow0 + j * 16 + fr) * 128;
    const i32x4 lo = *reinterpret_cast<const i32x4*>(row + p0);
    const i32x4 hi = *reinterpret_cast<const i32x4*>(row + p1);
    return i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  };
  // the tile scales are powers of two: their exponent fields are the E8M0
  // block scales of the MFMA, which then accumulates directly
  const int ablock = tm * MB + (arow0 >> 7), bblock = tn * NB + (wn >> 1);
  // The exponents of this wave's A and B tile scales, slice k in lane
  // k % 64 of a register that holds slices [64 c, 64 c + 64), read with
  // readlane: a slice starts without waiting for a scalar load of its
  // scales (two per slice cost a memory round trip after every barrier).
  int ea_v = 0, eb_v = 0;
  auto load_scales = [&](int c) {
    const int k = 64 * c + lane;
    const int kk = k < nkb ? k : nkb - 1;
    ea_v = (int)((__float_as_uint(sa[(size_t)ablock * nkb + kk]) >> 23) & 0xFF);
    eb_v = (int)((__float_as_uint(sb[(size_t)bblock * nkb + kk]) >> 23) & 0xFF);
  };
  load_scales(0);
  auto slice = [&](const uint8_t* Ac, const uint8_t* Bc, uint8_t* An, uint8_t* Bn, int k) {
    if (k > 0 && (k & 63) == 0) load_scales(k >> 6);
    const int ea = __builtin_amdgcn_readlane(ea_v, k & 63);
    const int eb = __builtin_amdgcn_readlane(eb_v, k & 63);
    // B fragments for the whole slice; A fragments one ahead of their four
    // MFMAs (scheduling barriers keep the compiler from hoisting all of them,
    // which would not fit in the registers)
    i32x8 b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = rdB(Bc, j);
    i32x8 a = rdA(Ac, 0);
#pragma unroll
    for (int i = 0; i < WR; ++i) {
      const i32x8 an = i + 1 < WR ? rdA(Ac, i + 1) : a;
#pragma unroll
      for (int j = 0; j < 4; ++j)  // format 0 = e4m3 for both operands
        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b[j], acc[i][j], 0, 0,
                                                                     0, ea, 0, eb);
      if constexpr (SPREAD) {
#pragma unroll
        for (int p = 2 * i; p < 2 * i + 2; ++p)
          if (p < PA + PB) fill_piece(An, Bn, k + 1, p);
      }
      __builtin_amdgcn_sched_barrier(0);
      a = an;
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  // slice k from (Ac, Bc); the other buffer (An, Bn), which slice k - 1
  // read, is refilled with slice k + 1
  auto step = [&](const uint8_t* Ac, const uint8_t* Bc, uint8_t* An, uint8_t* Bn, int k) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's part of fill k has landed
    // every wave's part, and every wave is done with slice k - 1.  A plain
    // barrier (a release fence, as in __syncthreads, would add nothing after
    // the wait above); the memory clobber keeps the slice's LDS reads below
    // it.
    asm volatile("s_barrier" ::: "memory");
    if constexpr (!SPREAD) fill(An, Bn, k + 1);
    __builtin_amdgcn_sched_barrier(0);
    slice(Ac, Bc, An, Bn, k);
  };

  fill(As0, Bs0, 0);
  int k = 0;
  for (; k + 2 <= nkb; k += 2) {
    step(As0, Bs0, As1, Bs1, k);
    step(As1, Bs1, As0, Bs0, k + 1);
  }
  if (k < nkb) step(As0, Bs0, As1, Bs1, k);
  __builtin_amdgcn_s_waitcnt(0);  // no fill may still be writing LDS at the end

  // Non-temporal stores of C: back to back, 8192^3 products took 0.418 ms
  // against 0.458 ms with plain stores (the write-back of one product's C
  // then no longer competes with the next one's operand reads); after the
  // quantiser the gain is about 1%
  const int cc = lane & 15, cr = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < WR; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + cc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + arow0 + i * 16 + cr + r;
        if (row < M && col < N) __builtin_nontemporal_store(acc[i][j][r], C + (size_t)row * ldc + col);
      }
    }
}

__global__ __launch_bounds__(256) void gemm_f32_kernel(const float* __restrict__ A,
                                                       const float* __restrict__ B,
                                                       float* __restrict__ C, int M, int N,
                                                       int K, int ta, int tb, int ldc)
{
  __shared__ float As[FK][BM + 4];
  __shared__ float Bs[FK][BN + 4];
  int tm, tn;
  tile_coords(tm, tn);
  f32_tile(A, B, C, M, N, K, ta, tb, ldc, tm * BM, tn * BN, As, Bs);
}

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

int st(hipError_t e) { return e == hipSuccess ? SKL_OK : SKL_ERR_HIP; }

}  // namespace

// Compute units of the current device (256 if the query fails).
static int device_cus()
{
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus;
}

// Workspace sizes (bytes) of the MFMA forms for an M x N x K product.
extern "C" size_t skl_internal_matmul_workspace(size_t M, size_t N, size_t K, int precision)
{
  const size_t Np = round_up(N, 128);
  if (precision == SKL_PREC_BF16X3) {
    const size_t Kp = round_up(K, 64);
    return 3 * (round_up(M, 256) + Np) * Kp * 2 + 256;  // A rows in pairs of blocks
  }
  if (precision == SKL_PREC_FP8) {
    const size_t Kp = round_up(K, 128), Ma = round_up(M, 256), Na = round_up(N, 256);
    return (Ma + Na) * Kp + ((Ma + Na) / 128) * (Kp / 128) * 4 + 512;
  }
  return 0;
}

// tile: the MFMA work-group tile, 0 = by the size rule below.  BF16X3: 1 =
// 128 x 128, 2 or 3 = 256 x 128.  FP8: 1 = 128 x 128, 2 = 256 x 128, 3 =
// 256 x 256.
extern "C" int skl_internal_matmul(float* C, const float* A, const float* B, int M, int N,
                                   int K, int ta, int tb, int precision, int tile, void* ws,
                                   hipStream_t s)
{
  const int Np = (int)round_up(N, 128);
  if (precision == SKL_PREC_BF16X3) {
    const int Kp = (int)round_up(K, 64);
    const int Ma = (int)round_up(M, 256);  // A rows: whole pairs of 128-row blocks
    uint16_t* Ap = static_cast<uint16_t*>(ws);
    uint16_t* Bp = Ap + 3 * (size_t)Ma * Kp;
    // tiled planes of A (M x K, or K x M when ta)
    if (!ta) {
      const size_t tot = (size_t)Ma * (Kp / 8);
      const unsigned g = (unsigned)std::min<size_t>((tot + 255) / 256, 65536);
      const int vec = (uintptr_t)A % 16 == 0 && K % 4 == 0;
      split3_kernel<<<g, 256, 0, s>>>(A, M, K, vec, Ap, Ma, Kp);
    } else {
      split3_transpose_kernel<<<dim3(Ma / 64, Kp / 32), 256, 0, s>>>(A, M, K, Ap, Ma, Kp);
    }
    // tiled planes of B^T (B is K x N, or N x K when tb)
    if (tb) {
      const size_t tot = (size_t)Np * (Kp / 8);
      const unsigned g = (unsigned)std::min<size_t>((tot + 255) / 256, 65536);
      const int vec = (uintptr_t)B % 16 == 0 && K % 4 == 0;
      split3_kernel<<<g, 256, 0, s>>>(B, N, K, vec, Bp, Np, Kp);
    } else {
      split3_transpose_kernel<<<dim3(Np / 64, Kp / 32), 256, 0, s>>>(B, N, K, Bp, Np, Kp);
    }
    // 256 x 128 tiles (8 waves, twice the MFMA work per fill) when there are
    // enough of them for every CU; 128 x 128 otherwise
    if (tile == 0) tile = (long long)(Ma / 256) * (Np / BN) >= device_cus() ? 2 : 1;
    if (tile >= 2) {
      gemm_bf16x3_kernel<2><<<dim3(Np / BN, Ma / 256), 512, 0, s>>>(Ap, Bp, C, M, N, Kp, N);
    } else {
      gemm_bf16x3_kernel<1><<<dim3(Np / BN, Ma / 128), 256, 0, s>>>(Ap, Bp, C, M, N, Kp, N);
    }
    return st(hipGetLastError());
  }
  if (precision == SKL_PREC_FP8) {
    const int Kp = (int)round_up(K, 128);
    const int Ma = (int)round_up(M, 256), Na = (int)round_up(N, 256);  // pairs of blocks
    uint8_t* Aq = static_cast<uint8_t*>(ws);
    uint8_t* Bq = Aq + (size_t)Ma * Kp;
    float* sa = reinterpret_cast<float*>(Bq + (size_t)Na * Kp);
    float* sb = sa + (size_t)(Ma / 128) * (Kp / 128);
    const int va = ((uintptr_t)A % 16 == 0) && ((ta ? M : K) % 4 == 0);
    const int vb = ((uintptr_t)B % 16 == 0) && ((tb ? K : N) % 4 == 0);
    // A, and B^T: rows n, columns k; element (n, k) = B[k][n] (or B[n][k]
    // when tb); one launch for both
    const QuantJob qa{A, M, K, ta, va, Aq, sa, Ma, Kp};
    const QuantJob qb{B, N, K, tb ? 0 : 1, vb, Bq, sb, Na, Kp};
    quant_fp8_kernel<<<dim3(Kp / 128, std::max(Ma, Na) / 128, 2), 256, 0, s>>>(qa, qb);
    // 256 x 256 tiles (8 waves of 128 x 64: per MFMA 3/4 of the LDS reads
    // and 2/3 of the fill bytes of 256 x 128) when every CU gets one; at
    // 8192^3 465 us against 590.  Then 256 x 128, then 128 x 128.
    if (tile == 0)
      tile = (long long)(Ma / 256) * (Na / 256) >= device_cus()   ? 3
             : (long long)(Ma / 256) * (Na / 128) >= device_cus() ? 2
                                                                   : 1;
    if (tile == 3) {
      gemm_fp8_kernel<2, 2, 8><<<dim3(Na / 256, Ma / 256), 512, 0, s>>>(Aq, Bq, sa, sb, C, M, N,
                                                                         Kp, N);
    } else if (tile == 2) {
      gemm_fp8_kernel<2, 1, 4><<<dim3(Na / 128, Ma / 256), 512, 0, s>>>(Aq, Bq, sa, sb, C, M, N,
                                                                         Kp, N);
    } else {
      gemm_fp8_kernel<1, 1, 4><<<dim3(Na / 128, Ma / 128), 256, 0, s>>>(Aq, Bq, sa, sb, C, M, N,
                                                                        Kp, N);
    }
    return st(hipGetLastError());
  }
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  gemm_f32_kernel<<<grid, 256, 0, s>>>(A, B, C, M, N, K, ta, tb, N);
  return st(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Batched products (many small AllPairs(+,*) on float) in one launch after a
// one-work-group prologue.  Work item = one 128x128 tile of one product;
// tiles are independent, so the work-groups take them round-robin (no
// ticket: a shared counter would serialise one RMW per tile).
//   FP32    the single product's tile body (f32_tile).
//   BF16X3  the FP32 operand slices are split into three bf16 terms with the
//           hardware round-to-nearest conversion while they are staged in
//           LDS (no pre-pass; the next slice is loaded during the MFMAs),
//           then the six MFMA products of gemm_bf16x3_kernel.  Products
//           with K and N multiples of 4 go to a variant that stages with
//           16-byte loads and is held to 256 VGPRs (two work-groups per
//           CU); the others to one with element loads.  The prologue numbers
//           the tiles of each class separately.
namespace {

struct BatchCtl {
  unsigned long long total;      // tiles of all products (FP32 kernel)
  unsigned long long total_vec;  // tiles of products with K % 4 == 0 and N % 4 == 0
  unsigned long long total_gen;  // tiles of the other products
};

// tstart[0 .. count]: first tile of each product over all products;
// tstart[count + 1 + p] / tstart[2 (count + 1) + p]: the same numbering
// over the products of one class (the others count no tiles).
__global__ __launch_bounds__(1024) void batch_prologue(const skl_gemm_desc* __restrict__ d,
                                                       int count, long long* __restrict__ tstart,
                                                       BatchCtl* ctl)
{
  __shared__ long long part[3][1024];
  __shared__ long long carry[3];
  if (threadIdx.x < 3) carry[threadIdx.x] = 0;
  __syncthreads();
  for (int base = 0; base < count; base += 1024) {
    const int p = base + threadIdx.x;
    long long t = 0;
    bool v = false;
    if (p < count && d[p].M > 0 && d[p].N > 0) {
      t = (long long)((d[p].M + BM - 1) / BM) * ((d[p].N + BN - 1) / BN);
      v = batch_vec(d[p]);
    }
    const long long tc[3] = {t, v ? t : 0, v ? 0 : t};
#pragma unroll
    for (int c = 0; c < 3; ++c) part[c][threadIdx.x] = tc[c];
    __syncthreads();
    for (int s = 1; s < 1024; s <<= 1) {
      long long w[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) w[c] = threadIdx.x >= (unsigned)s ? part[c][threadIdx.x - s] : 0;
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 3; ++c) part[c][threadIdx.x] += w[c];
      __syncthreads();
    }
    if (p < count)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        tstart[c * (count + 1) + p] = carry[c] + part[c][threadIdx.x] - tc[c];
    __syncthreads();
    if (threadIdx.x < 3) carry[threadIdx.x] += part[threadIdx.x][1023];
    __syncthreads();
  }
  if (threadIdx.x < 3) tstart[threadIdx.x * (count + 1) + count] = carry[threadIdx.x];
  if (threadIdx.x == 0) {
    ctl->total = (unsigned long long)carry[0];
    ctl->total_vec = (unsigned long long)carry[1];
    ctl->total_gen = (unsigned long long)carry[2];
  }
}

__global__ __launch_bounds__(256) void batch_f32_kernel(const skl_gemm_desc* __restrict__ d,
                                                        int count,
                                                        const long long* __restrict__ tstart,
                                                        const BatchCtl* __restrict__ ctl)
{
  __shared__ float As[FK][BM + 4];
  __shared__ float Bs[FK][BN + 4];
  const long long total = (long long)ctl->total;
  for (long long t = blockIdx.x; t < total; t += gridDim.x) {
    const int p = find_item(tstart, count, t);
    const skl_gemm_desc g = d[p];
    const long long lt = t - tstart[p];
    const int ntn = (g.N + BN - 1) / BN;
    f32_tile(g.A, g.B, g.C, g.M, g.N, g.K, 0, 0, g.N, (int)(lt / ntn) * BM, (int)(lt % ntn) * BN,
             As, Bs);
  }
}

// VEC: products with K % 4 == 0 and N % 4 == 0 only (16-byte staging loads,
// two waves per SIMD); otherwise the others (element loads)
template <bool VEC>
__device__ __forceinline__ void batch_bf16x3_body(const skl_gemm_desc* __restrict__ d, int count,
                                                  const long long* __restrict__ tstart,
                                                  const BatchCtl* __restrict__ ctl)
{
  __shared__ __attribute__((aligned(16))) __bf16 As[3][BM][BLD];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[3][BN][BLD];
  const long long total = (long long)(VEC ? ctl->total_vec : ctl->total_gen);
  for (long long t = blockIdx.x; t < total; t += gridDim.x) {
    const int p = find_item(tstart, count, t);
    const skl_gemm_desc g = d[p];
    const long long lt = t - tstart[p];
    const int ntn = (g.N + BN - 1) / BN;
    bf16x3_tile<VEC ? 1 : 0>(g, (int)(lt / ntn) * BM, (int)(lt % ntn) * BN, As, Bs);
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
batch_bf16x3_vec_kernel(const skl_gemm_desc* __restrict__ d, int count,
                        const long long* __restrict__ tstart, const BatchCtl* __restrict__ ctl)
{
  batch_bf16x3_body<true>(d, count, tstart, ctl);
}

__global__ __launch_bounds__(256) void batch_bf16x3_gen_kernel(const skl_gemm_desc* __restrict__ d,
                                                               int count,
                                                               const long long* __restrict__ tstart,
                                                               const BatchCtl* __restrict__ ctl)
{
  batch_bf16x3_body<false>(d, count, tstart, ctl);
}

}  // namespace

extern "C" size_t skl_internal_batched_workspace(int count)
{
  return 256 + 3 * (size_t)(count + 1) * 8;
}

extern "C" int skl_internal_batched_gemm_ws(const skl_gemm_desc* descs, int count, int precision,
                                            void* ws, hipStream_t s)
{
  BatchCtl* ctl = static_cast<BatchCtl*>(ws);
  long long* tstart = reinterpret_cast<long long*>(static_cast<char*>(ws) + 256);
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      cus = v;
  }
  batch_prologue<<<1, 1024, 0, s>>>(descs, count, tstart, ctl);
  if (precision == SKL_PREC_BF16X3) {
    batch_bf16x3_vec_kernel<<<cus * 2, 256, 0, s>>>(descs, count, tstart + (count + 1), ctl);
    batch_bf16x3_gen_kernel<<<cus, 256, 0, s>>>(descs, count, tstart + 2 * (count + 1), ctl);
  } else
    batch_f32_kernel<<<cus * 3, 256, 0, s>>>(descs, count, tstart, ctl);  // 3 fit per CU
  return st(hipGetLastError());
}

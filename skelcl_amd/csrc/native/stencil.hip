// MapOverlap forward for user functions of the weighted-stencil form
//   float func(input_matrix_t img, __global float* w) {
//     float sum = 0.0f;
//     for (int dy = -R; dy <= R; ++dy)
//       for (int dx = -R; dx <= R; ++dx)
//         sum += getData(img, dx, dy) * w[(dy + R) * (2R+1) + (dx + R)];
//     return sum;
//   }
// which the skeleton layer recognises token for token.  Every output takes
// the same sequence of operations as that loop: dy ascending, dx ascending,
// one fused multiply-add per term from 0.0f (or, without contraction, a
// rounded product and a rounded sum: this file is built with
// -ffp-contract=off and fuses only through fmaf), so the results are
// bit-identical to the generic fused kernel's.
//
// The generic kernel runs the loop as written for several rows per thread;
// the compiler then keeps all (2R+1)^2 weights live in scalar registers,
// which do not hold 121 of them, and spends a third of its instructions
// moving spilled weights back (v_readlane) -- the kernel is bound by VALU
// issue.  Here a work-group of 64 x 4 threads computes a 64 x 4RB tile, RB
// rows of one column per thread, streaming the tile rows: at step t a thread
// reads tile row t of its window (2R+1 values) once and feeds it to every
// output whose window holds it (weight row t - i for output i), and weight
// row t enters at step t.  So only RB weight rows are live at a time, in
// scalar registers that the FMAs read directly.
//
// The tile is staged row by row: wave v loads tile rows v, v+4, ..., each
// lane one column (and the 2R halo columns past 64 in a second load), from
// clamped coordinates (NEAREST) with the neutral value outside (NEUTRAL),
// every load in flight before the LDS writes.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"

#ifndef SKL_STENCIL_LAG
#define SKL_STENCIL_LAG 1
#endif

namespace {

// Global memory at a wave-uniform row address plus a 32-bit lane offset.
// The row address passes through an empty asm, so that it stays a scalar
// base: the compiler would otherwise fold the lane offset into the row
// address once and add every row's offset in 64-bit vector arithmetic.
__device__ __forceinline__ float load_row(const void* row, unsigned off)
{
  unsigned long long r = (unsigned long long)row;
  asm volatile("" : "+s"(r));
  return *(const __attribute__((address_space(1))) float*)(r + off);
}

__device__ __forceinline__ void store_row(void* row, unsigned off, float v)
{
  unsigned long long r = (unsigned long long)row;
  asm volatile("" : "+s"(r));
  *(__attribute__((address_space(1))) float*)(r + off) = v;
}

template <int R, int RB, bool NEUTRAL, bool FUSED>
__global__ __launch_bounds__(256) void stencil_forward_kernel(const float* __restrict__ in,
                                                              float* __restrict__ out,
                                                              const float* __restrict__ w,
                                                              int rows, int cols, float neutral)
{
  constexpr int W = 2 * R + 1;
  constexpr int TW = 64, TH = 4 * RB, TWP = TW + 2 * R, THP = TH + 2 * R;
  constexpr int NR = (THP + 3) / 4;  // tile rows per wave
  __shared__ float tile[THP * TWP];
  const int lane = threadIdx.x & 63;
  // uniform per wave, so row addresses are scalar
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c0 = blockIdx.x * TW, r0 = blockIdx.y * TH;

  // ---- stage the tile.  Tiles whose window lies inside the matrix (all but the
  // border tiles) load without clamping or neutral fill.
  const unsigned l4 = (unsigned)lane * 4u;
  const bool halo = lane < 2 * R;  // lanes that also load a column past 64
  float a[NR], b[NR];
  const bool interior = c0 >= R && c0 + TW + R <= cols && r0 >= R && r0 + TH + R <= rows;
  if (interior) {
    const float* base = in + (size_t)(r0 - R) * (size_t)cols + (c0 - R);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int tr = wv + 4 * k;
      if (tr < THP) {
        const float* row = base + (size_t)tr * (size_t)cols;
        a[k] = load_row(row, l4);
        if (halo) b[k] = load_row(row, l4 + 256u);
      }
    }
  } else {
    const int gc0 = c0 - R + lane, gc1 = gc0 + 64;
    const int cc0 = gc0 < 0 ? 0 : (gc0 >= cols ? cols - 1 : gc0);
    const int cc1 = gc1 < 0 ? 0 : (gc1 >= cols ? cols - 1 : gc1);
    const unsigned o0 = (unsigned)cc0 * 4u, o1 = (unsigned)cc1 * 4u;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int tr = wv + 4 * k;
      if (tr < THP) {
        const int gr = r0 - R + tr;
        const int cr = gr < 0 ? 0 : (gr >= rows ? rows - 1 : gr);
        const float* row = in + (size_t)cr * (size_t)cols;
        a[k] = load_row(row, o0);
        if (halo) b[k] = load_row(row, o1);
        if (NEUTRAL) {
          if (gr != cr || gc0 != cc0) a[k] = neutral;
          if (gr != cr || gc1 != cc1) b[k] = neutral;
        }
      }
    }
  }
  float* trow = tile + wv * TWP + lane;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    if (wv + 4 * k < THP) {
      trow[4 * k * TWP] = a[k];
      if (halo) trow[4 * k * TWP + 64] = b[k];
    }
  }
  __syncthreads();

  // ---- RB outputs per thread, tile rows streamed
  float acc[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) acc[i] = 0.0f;
  float wr[W][W];
  const char* win = (const char*)(tile + (wv * RB) * TWP + lane);
  // The loads of step t (weight row t, tile row t) take byte offsets that are
  // opaque to the compiler and become available once every accumulator has
  // finished step t - LAG.  Without that the scheduler runs one output's
  // chain far ahead of the others (they are independent), and the weight
  // rows the lagging outputs still need overflow the scalar registers; and
  // an opaque tile offset costs one address add per step instead of one per
  // LDS read.  So the outputs advance together, about RB + LAG weight rows
  // are live, in scalar registers that the FMAs read directly.  (Constant
  // address space: the weight loads stay scalar.)
  constexpr int LAG = SKL_STENCIL_LAG;
  unsigned zt[LAG + 1], zw[LAG + 1];
#pragma unroll
  for (int q = 0; q < LAG; ++q) {
    zt[q] = q * TWP * 4;
    zw[q] = q * W * 4;
    asm volatile("" : "+s"(zt[q]), "+s"(zw[q]));
  }
#pragma unroll
  for (int t = 0; t < RB + 2 * R; ++t) {
    const int q = t % (LAG + 1);
    if (t < W) {
      const __attribute__((address_space(4))) float* wrow =
          (const __attribute__((address_space(4))) float*)((unsigned long long)w + zw[q]);
#pragma unroll
      for (int j = 0; j < W; ++j) wr[t][j] = wrow[j];
    }
    float v[W];
    const float* wt = (const float*)(win + zt[q]);
#pragma unroll
    for (int j = 0; j < W; ++j) v[j] = wt[j];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int k = t - i;  // weight row (dy + R) of output i at this step
      if (k < 0 || k >= W) continue;
#pragma unroll
      for (int j = 0; j < W; ++j)
        acc[i] = FUSED ? __builtin_fmaf(v[j], wr[k][j], acc[i]) : acc[i] + v[j] * wr[k][j];
    }
    const int n = (t + LAG) % (LAG + 1);
    zt[n] = (t + LAG) * TWP * 4;
    zw[n] = (t + LAG) * W * 4;
#pragma unroll
    for (int i = 0; i < RB; ++i) asm volatile("" : "+s"(zt[n]), "+s"(zw[n]) : "v"(acc[i]));
  }
  if (c0 + lane < cols) {
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int r = r0 + wv * RB + i;
      if (r < rows) store_row(out + (size_t)r * (size_t)cols + c0, l4, acc[i]);
    }
  }
}

// The same schedule with two adjacent columns per thread: a 128 x 4RB tile,
// outputs (r, 2l) and (r, 2l + 1) for lane l.  Each tile row is read as
// R + 1 aligned 8-byte pairs (2R + 2 values) that feed both columns'
// windows, so every LDS byte feeds twice the fused multiply-adds of the
// one-column kernel and every weight row loaded feeds two columns.  The
// weights enter by scalar loads and are moved to vector registers.  Every
// output still takes its terms in the loop's order.
template <int R, int RB, bool NEUTRAL, bool FUSED, bool ADJ = false>
__global__ __launch_bounds__(256) void stencil_forward2_kernel(const float* __restrict__ in,
                                                               float* __restrict__ out,
                                                               const float* __restrict__ w,
                                                               int rows, int cols, float neutral)
{
  constexpr int W = 2 * R + 1;
  constexpr int TW = 128, TH = 4 * RB, TWP = TW + 2 * R, THP = TH + 2 * R;
  constexpr int NR = (THP + 3) / 4;  // tile rows per wave
  static_assert(TWP % 2 == 0, "tile rows must stay 8-byte aligned");
  __shared__ __align__(16) float tile[THP * TWP];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c0 = blockIdx.x * TW, r0 = blockIdx.y * TH;

  // ---- stage the tile: lane l loads columns l, l + 64 and, for l < 2R, the
  // halo column l + 128 of each of its wave's tile rows.
  const unsigned l4 = (unsigned)lane * 4u;
  const bool halo = lane < 2 * R;
  float a[NR], b[NR], h[NR];
  const bool interior = c0 >= R && c0 + TW + R <= cols && r0 >= R && r0 + TH + R <= rows;
  if (interior) {
    const float* base = in + (size_t)(r0 - R) * (size_t)cols + (c0 - R);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int tr = wv + 4 * k;
      if (tr < THP) {
        const float* row = base + (size_t)tr * (size_t)cols;
        a[k] = load_row(row, l4);
        b[k] = load_row(row, l4 + 256u);
        if (halo) h[k] = load_row(row, l4 + 512u);
      }
    }
  } else {
    const int gc0 = c0 - R + lane, gc1 = gc0 + 64, gc2 = gc0 + 128;
    const int cc0 = gc0 < 0 ? 0 : (gc0 >= cols ? cols - 1 : gc0);
    const int cc1 = gc1 < 0 ? 0 : (gc1 >= cols ? cols - 1 : gc1);
    const int cc2 = gc2 < 0 ? 0 : (gc2 >= cols ? cols - 1 : gc2);
    const unsigned o0 = (unsigned)cc0 * 4u, o1 = (unsigned)cc1 * 4u, o2 = (unsigned)cc2 * 4u;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int tr = wv + 4 * k;
      if (tr < THP) {
        const int gr = r0 - R + tr;
        const int cr = gr < 0 ? 0 : (gr >= rows ? rows - 1 : gr);
        const float* row = in + (size_t)cr * (size_t)cols;
        a[k] = load_row(row, o0);
        b[k] = load_row(row, o1);
        if (halo) h[k] = load_row(row, o2);
        if (NEUTRAL) {
          if (gr != cr || gc0 != cc0) a[k] = neutral;
          if (gr != cr || gc1 != cc1) b[k] = neutral;
          if (gr != cr || gc2 != cc2) h[k] = neutral;
        }
      }
    }
  }
  float* trow = tile + wv * TWP + lane;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    if (wv + 4 * k < THP) {
      trow[4 * k * TWP] = a[k];
      trow[4 * k * TWP + 64] = b[k];
      if (halo) trow[4 * k * TWP + 128] = h[k];
    }
  }
  __syncthreads();

  // ---- RB rows x 2 columns per thread, tile rows streamed (the loads of
  // step t wait, through opaque offsets, for every accumulator of step
  // t - LAG, as in the one-column kernel)
  float acc[RB][2];
#pragma unroll
  for (int i = 0; i < RB; ++i) acc[i][0] = acc[i][1] = 0.0f;
  float wr[W][W];
  const char* win = (const char*)(tile + (wv * RB) * TWP + 2 * lane);
  constexpr int LAG = SKL_STENCIL_LAG;
  constexpr int NSTEP = RB + 2 * R;
  // the tile row read at step t: ascending, or descending for the adjoint
  auto trow_of = [](int t) { return ADJ ? (t < NSTEP ? NSTEP - 1 - t : 0) : t; };
  unsigned zt[LAG + 1], zw[LAG + 1];
#pragma unroll
  for (int q = 0; q < LAG; ++q) {
    zt[q] = trow_of(q) * TWP * 4;
    zw[q] = q * W * 4;
    asm volatile("" : "+s"(zt[q]), "+s"(zw[q]));
  }
#pragma unroll
  for (int t = 0; t < RB + 2 * R; ++t) {
    const int q = t % (LAG + 1);
    if (t < W) {
      const __attribute__((address_space(4))) float* wrow =
          (const __attribute__((address_space(4))) float*)((unsigned long long)w + zw[q]);
      // each weight moved to a vector register once: fused multiply-adds
      // with vector operands only issue faster than with a scalar one
      // (8192^2 R = 5: 0.30 -> 0.27-0.28 ms), and a weight row serves the
      // RB steps that use it
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const float ws = wrow[j];
        asm volatile("v_mov_b32 %0, %1" : "=v"(wr[t][j]) : "s"(ws));
      }
    }
    float v[W + 1];
    const float2* wt = (const float2*)(win + zt[q]);
#pragma unroll
    for (int j = 0; j < (W + 1) / 2; ++j) {
      const float2 p = wt[j];
      v[2 * j] = p.x;
      v[2 * j + 1] = p.y;
    }
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int k = t - i;  // weight row (dy + R) of row i at this step
      if (k < 0 || k >= W) continue;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        // forward: column c + j of the window; adjoint: c + 2R - j
        const float a0 = ADJ ? v[2 * R - j] : v[j];
        const float a1 = ADJ ? v[2 * R + 1 - j] : v[j + 1];
        acc[i][0] = FUSED ? __builtin_fmaf(a0, wr[k][j], acc[i][0]) : acc[i][0] + a0 * wr[k][j];
        acc[i][1] = FUSED ? __builtin_fmaf(a1, wr[k][j], acc[i][1]) : acc[i][1] + a1 * wr[k][j];
      }
    }
    const int n = (t + LAG) % (LAG + 1);
    zt[n] = trow_of(t + LAG) * TWP * 4;
    zw[n] = (t + LAG) * W * 4;
#pragma unroll
    for (int i = 0; i < RB; ++i)
      asm volatile("" : "+s"(zt[n]), "+s"(zw[n]) : "v"(acc[i][0]), "v"(acc[i][1]));
  }
  const int cl = c0 + 2 * lane;
  const unsigned o8 = (unsigned)lane * 8u;
  if ((cols & 1) == 0 && ((unsigned long long)out & 7u) == 0 && cl + 1 < cols) {
    // even column count and an 8-byte aligned output: every row starts
    // 8-byte aligned
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int r = r0 + wv * RB + (ADJ ? RB - 1 - i : i);
      if (r < rows) {
        unsigned long long rb = (unsigned long long)(out + (size_t)r * (size_t)cols + c0);
        asm volatile("" : "+s"(rb));
        typedef float f2 __attribute__((ext_vector_type(2)));
        *(__attribute__((address_space(1))) f2*)(rb + o8) = f2{acc[i][0], acc[i][1]};
      }
    }
  } else if (cl < cols) {
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int r = r0 + wv * RB + (ADJ ? RB - 1 - i : i);
      if (r < rows) {
        float* row = out + (size_t)r * (size_t)cols + c0;
        store_row(row, o8, acc[i][0]);
        if (cl + 1 < cols) store_row(row, o8 + 4u, acc[i][1]);
      }
    }
  }
}

template <int R, int RB>
hipError_t launch_rb2(float* out, const float* in, const float* w, int rows, int cols,
                      bool neutral_pad, float neutral, bool fused, hipStream_t s)
{
  dim3 grid((cols + 127) / 128, (rows + 4 * RB - 1) / (4 * RB)), block(256);
  if (neutral_pad) {
    if (fused)
      hipLaunchKernelGGL((stencil_forward2_kernel<R, RB, true, true>), grid, block, 0, s, in, out,
                         w, rows, cols, neutral);
    else
      hipLaunchKernelGGL((stencil_forward2_kernel<R, RB, true, false>), grid, block, 0, s, in,
                         out, w, rows, cols, neutral);
  } else {
    if (fused)
      hipLaunchKernelGGL((stencil_forward2_kernel<R, RB, false, true>), grid, block, 0, s, in,
                         out, w, rows, cols, neutral);
    else
      hipLaunchKernelGGL((stencil_forward2_kernel<R, RB, false, false>), grid, block, 0, s, in,
                         out, w, rows, cols, neutral);
  }
  return hipGetLastError();
}

template <int R, int RB>
hipError_t launch_rb(float* out, const float* in, const float* w, int rows, int cols,
                     bool neutral_pad, float neutral, bool fused, hipStream_t s)
{
  dim3 grid((cols + 63) / 64, (rows + 4 * RB - 1) / (4 * RB)), block(256);
  if (neutral_pad) {
    if (fused)
      hipLaunchKernelGGL((stencil_forward_kernel<R, RB, true, true>), grid, block, 0, s, in, out,
                         w, rows, cols, neutral);
    else
      hipLaunchKernelGGL((stencil_forward_kernel<R, RB, true, false>), grid, block, 0, s, in,
                         out, w, rows, cols, neutral);
  } else {
    if (fused)
      hipLaunchKernelGGL((stencil_forward_kernel<R, RB, false, true>), grid, block, 0, s, in,
                         out, w, rows, cols, neutral);
    else
      hipLaunchKernelGGL((stencil_forward_kernel<R, RB, false, false>), grid, block, 0, s, in,
                         out, w, rows, cols, neutral);
  }
  return hipGetLastError();
}

template <int R>
hipError_t launch_r(float* out, const float* in, const float* w, int rows, int cols,
                    bool neutral_pad, float neutral, bool fused, int rb, int cb, hipStream_t s)
{
  if (cb == 2) {
    switch (rb) {
      case 4: return launch_rb2<R, 4>(out, in, w, rows, cols, neutral_pad, neutral, fused, s);
      case 6: return launch_rb2<R, 6>(out, in, w, rows, cols, neutral_pad, neutral, fused, s);
      case 8: return launch_rb2<R, 8>(out, in, w, rows, cols, neutral_pad, neutral, fused, s);
    }
    return hipErrorInvalidValue;
  }
  if (cb != 1) return hipErrorInvalidValue;
  switch (rb) {
    case 4: return launch_rb<R, 4>(out, in, w, rows, cols, neutral_pad, neutral, fused, s);
    case 6: return launch_rb<R, 6>(out, in, w, rows, cols, neutral_pad, neutral, fused, s);
    case 8: return launch_rb<R, 8>(out, in, w, rows, cols, neutral_pad, neutral, fused, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

// The transposed stencil of g under zero padding for R in 1..5 (every
// pixel; a NEAREST caller rewrites the border afterwards): 8 rows x 2
// columns per thread.
extern "C" int skl_internal_stencil_adjoint_zero(float* din, const float* g, const float* w,
                                                 int rows, int cols, int R, hipStream_t s)
{
  if (rows <= 0 || cols <= 0) return rows < 0 || cols < 0 ? SKL_ERR_INVALID : SKL_OK;
  constexpr int RB = 8;
  dim3 grid((cols + 127) / 128, (rows + 4 * RB - 1) / (4 * RB)), block(256);
  if (grid.y > 65535) return SKL_ERR_INVALID;
  switch (R) {
    case 1: hipLaunchKernelGGL((stencil_forward2_kernel<1, RB, true, true, true>), grid, block, 0, s, g, din, w, rows, cols, 0.0f); break;
    case 2: hipLaunchKernelGGL((stencil_forward2_kernel<2, RB, true, true, true>), grid, block, 0, s, g, din, w, rows, cols, 0.0f); break;
    case 3: hipLaunchKernelGGL((stencil_forward2_kernel<3, RB, true, true, true>), grid, block, 0, s, g, din, w, rows, cols, 0.0f); break;
    case 4: hipLaunchKernelGGL((stencil_forward2_kernel<4, RB, true, true, true>), grid, block, 0, s, g, din, w, rows, cols, 0.0f); break;
    case 5: hipLaunchKernelGGL((stencil_forward2_kernel<5, RB, true, true, true>), grid, block, 0, s, g, din, w, rows, cols, 0.0f); break;
    default: return SKL_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

// Largest radius the stencil kernel is instantiated for.
extern "C" int skl_internal_stencil_max_radius() { return 5; }

// rb: rows per thread (4, 6 or 8); cb: columns per thread (1 or 2).
// Returns SKL_OK, or SKL_ERR_UNSUPPORTED for a radius, rb or cb without an
// instantiation.
extern "C" int skl_internal_stencil_forward(float* out, const float* in, const float* w,
                                            int rows, int cols, int R, int neutral_pad,
                                            float neutral, int fused, int rb, int cb,
                                            hipStream_t s)
{
  if (rows <= 0 || cols <= 0) return rows < 0 || cols < 0 ? SKL_ERR_INVALID : SKL_OK;
  hipError_t e = hipErrorInvalidValue;
  switch (R) {
    case 0: e = launch_r<0>(out, in, w, rows, cols, neutral_pad, neutral, fused, rb, cb, s); break;
    case 1: e = launch_r<1>(out, in, w, rows, cols, neutral_pad, neutral, fused, rb, cb, s); break;
    case 2: e = launch_r<2>(out, in, w, rows, cols, neutral_pad, neutral, fused, rb, cb, s); break;
    case 3: e = launch_r<3>(out, in, w, rows, cols, neutral_pad, neutral, fused, rb, cb, s); break;
    case 4: e = launch_r<4>(out, in, w, rows, cols, neutral_pad, neutral, fused, rb, cb, s); break;
    case 5: e = launch_r<5>(out, in, w, rows, cols, neutral_pad, neutral, fused, rb, cb, s); break;
    default: return SKL_ERR_UNSUPPORTED;
  }
  if (e == hipErrorInvalidValue) return SKL_ERR_UNSUPPORTED;
  return e == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

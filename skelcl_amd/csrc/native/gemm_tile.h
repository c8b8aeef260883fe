// Output tiles of the float products, shared by the product kernels
// (matmul.hip) and the one-launch batch kernel (batch.hip): the FP32 VALU
// tile and the BF16X3 MFMA tile of the batched small products, each one
// 128x128 block of C = A.B for 256 threads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_skeletons.h"
#include "ld_util.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128;

// ---------------------------------------------------------------------------
// FP32 VALU GEMM: 128x128 block tile, 8x8 per thread (rows ty*4+{0..3} and
// 64+ty*4+{0..3}, same split for columns), k-slices of 8.
constexpr int FK = 8;
// One 128x128 output tile at (m0, n0); shared by the single product and the
// batched products.  Ends with a barrier, so the LDS can be reused at once.
__device__ __forceinline__ void f32_tile(const float* __restrict__ A, const float* __restrict__ B,
                                         float* __restrict__ C, int M, int N, int K, int ta,
                                         int tb, int ldc, int m0, int n0,
                                         float (*__restrict__ As)[BM + 4],
                                         float (*__restrict__ Bs)[BN + 4])
{
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  float acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = 0.0f;
  // slice staging: A 128 rows x 8 k and B 8 k x 128 columns, 4 + 4 elements
  // per thread.  The next slice is loaded into registers (raw, from clamped
  // indices) while the current one is multiplied; masks apply at the store.
  auto a_at = [&](int j, int k0, int& r, int& kk, bool& ok) -> size_t {
    const int e = tid + 256 * j;
    if (ta) {  // A stored K x M: read along m
      kk = e >> 7;
      r = e & 127;
    } else {
      r = e >> 3;
      kk = e & 7;
    }
    const int gr = m0 + r, gk = k0 + kk;
    ok = gr < M && gk < K;
    return ta ? (size_t)gk * M + gr : (size_t)gr * K + gk;
  };
  auto b_at = [&](int j, int k0, int& c, int& kk, bool& ok) -> size_t {
    const int e = tid + 256 * j;
    if (tb) {  // B stored N x K: read along k
      c = e >> 3;
      kk = e & 7;
    } else {
      kk = e >> 7;
      c = e & 127;
    }
    const int gc = n0 + c, gk = k0 + kk;
    ok = gc < N && gk < K;
    return tb ? (size_t)gc * K + gk : (size_t)gk * N + gc;
  };
  float pa[4], pb[4];
  auto gl = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int r, kk, c;
      bool ok;
      const size_t ia = a_at(j, k0, r, kk, ok);
      pa[j] = as_global(A)[ok ? ia : 0];
      const size_t ib = b_at(j, k0, c, kk, ok);
      pb[j] = as_global(B)[ok ? ib : 0];
    }
  };
  auto st = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int r, kk, c;
      bool ok;
      a_at(j, k0, r, kk, ok);
      As[kk][r] = ok ? pa[j] : 0.0f;
      b_at(j, k0, c, kk, ok);
      Bs[kk][c] = ok ? pb[j] : 0.0f;
    }
  };
  gl(0);
  for (int k0 = 0; k0 < K; k0 += FK) {
    st(k0);
    __syncthreads();
    if (k0 + FK < K) gl(k0 + FK);
#pragma unroll
    for (int kk = 0; kk < FK; ++kk) {
      float a[8], b[8];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = As[kk][ty * 4 + i];
        a[4 + i] = As[kk][64 + ty * 4 + i];
        b[i] = Bs[kk][tx * 4 + i];
        b[4 + i] = Bs[kk][64 + tx * 4 + i];
      }
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int row = m0 + (i < 4 ? ty * 4 + i : 64 + ty * 4 + i - 4);
    if (row >= M) continue;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int col = n0 + (j < 4 ? tx * 4 + j : 64 + tx * 4 + j - 4);
      if (col < N) as_global(C)[(size_t)row * ldc + col] = acc[i][j];
    }
  }
}

__device__ __forceinline__ bool batch_vec(const skl_gemm_desc& g)
{
  return (g.K & 3) == 0 && (g.N & 3) == 0;
}

__device__ __forceinline__ int find_item(const long long* ts, int count, long long t)
{
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (ts[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// x = h0 + h1 + h2 (+ a remainder below 2^-24 |x|), each term the
// round-to-nearest bf16 of what is left (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ void split3_hw(float x, __bf16& h0, __bf16& h1, __bf16& h2)
{
  h0 = (__bf16)x;
  const float r1 = x - (float)h0;
  h1 = (__bf16)r1;
  h2 = (__bf16)(r1 - (float)h1);
}

constexpr int BKS = 32;               // k per slice
constexpr int BLD = BKS + 8;          // bf16 per LDS row (padding vs bank conflicts)

// One 128x128 tile (m0, n0) of product g in the BF16X3 form (A, B, C
// row-major with leading dimensions K, N, N): each 32-k slice of both
// operands is split into three bf16 planes while it is staged in LDS
// (As, Bs: three planes each), then per 16x16 block the six products with
// i + j <= 2 run on v_mfma_f32_16x16x32_bf16, x0y0 into one accumulator and
// the five corrections into another.  Staging, per thread and slice: MODE 0
// element loads in rows of A and columns of B; MODE 1 16-byte loads (K and N
// multiples of 4); MODE 2 the layout of MODE 1, with 16-byte loads when K
// and N are multiples of 4 and element loads otherwise.  Ends with a
// barrier, so the LDS can be reused at once.
template <int MODE>
__device__ __forceinline__ void bf16x3_tile(const skl_gemm_desc& g, int m0, int n0,
                                            __bf16 (*As)[BM][BLD], __bf16 (*Bs)[BN][BLD])
{
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f32x4 acc0[4][4], acc1[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc0[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      acc1[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  // Per thread and slice, A: 4 groups of 4 consecutive k (row r, k kq..kq+3).
  // B: when K and N are multiples of 4 (vec), a 4 x 4 block (k 4kb..4kb+3,
  // n 4nb..4nb+3) read as 4 float4 rows and transposed in registers;
  // otherwise 4 groups of 4 consecutive k of one column.  With vec every
  // 16-byte pack lies wholly inside or outside the operand.  The prefetch
  // loads go straight into ra / rb from clamped indices (a select right
  // after a load makes the compiler wait for it); out-of-range elements
  // are zeroed when the slice is written to LDS, each group's three bf16
  // terms as one 8-byte store per plane.
  typedef float f4a __attribute__((ext_vector_type(4), aligned(4)));
  const bool vec = MODE == 2 ? batch_vec(g) : MODE == 1;
  const int vkb = tid >> 5, vnb = tid & 31;
  float ra[4][4], rb[4][4];
  auto gload = [&](int k0) {
    if (MODE != 0) {
      // the 16-byte layout; MODE 2 loads element by element when K or N is
      // not a multiple of 4 (the packs may then straddle row ends)
      const bool wide = MODE == 1 || vec;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = tid + 256 * j;
        const int gr = m0 + (e >> 3), gk = k0 + (e & 7) * 4;
        if (wide) {
          const f4a v = *reinterpret_cast<const __attribute__((address_space(1))) f4a*>(
              as_global(g.A) + (gr < g.M && gk < g.K ? (size_t)gr * g.K + gk : 0));
#pragma unroll
          for (int i = 0; i < 4; ++i) ra[j][i] = v[i];
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            ra[j][i] = as_global(g.A)[gr < g.M && gk + i < g.K ? (size_t)gr * g.K + gk + i : 0];
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gk = k0 + 4 * vkb + i, gc = n0 + 4 * vnb;
        if (wide) {
          const f4a v = *reinterpret_cast<const __attribute__((address_space(1))) f4a*>(
              as_global(g.B) + (gk < g.K && gc < g.N ? (size_t)gk * g.N + gc : 0));
#pragma unroll
          for (int c = 0; c < 4; ++c) rb[i][c] = v[c];
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c)
            rb[i][c] = as_global(g.B)[gk < g.K && gc + c < g.N ? (size_t)gk * g.N + gc + c : 0];
        }
      }
    } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      const int r = e >> 3, kq = (e & 7) * 4;    // A: row r, k kq..kq+3
      const int n = e & 127, kb = (e >> 7) * 4;  // B: column n, k kb..kb+3
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gr = m0 + r, gk = k0 + kq + i;
        ra[j][i] = as_global(g.A)[gr < g.M && gk < g.K ? (size_t)gr * g.K + gk : 0];
        const int gkb = k0 + kb + i, gc = n0 + n;
        rb[j][i] = as_global(g.B)[gkb < g.K && gc < g.N ? (size_t)gkb * g.N + gc : 0];
      }
    }
    }
  };
  auto put4 = [&](__bf16 (*plane)[BLD], int row, int k, const float* v, int part) {
    __bf16 h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      __bf16 h0, h1, h2;
      split3_hw(v[i], h0, h1, h2);
      h[i] = part == 0 ? h0 : (part == 1 ? h1 : h2);
    }
    unsigned long long w;
    __builtin_memcpy(&w, h, 8);
    *reinterpret_cast<unsigned long long*>(&plane[row][k]) = w;
  };
  auto lstore = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      const int r = e >> 3, kq = (e & 7) * 4;
      float va[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) va[i] = m0 + r < g.M && k0 + kq + i < g.K ? ra[j][i] : 0.0f;
#pragma unroll
      for (int q = 0; q < 3; ++q) put4(As[q], r, kq, va, q);
    }
    if (MODE != 0) {
      // column 4nb + c of the block: k 4kb..4kb+3
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        float vb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          vb[i] = k0 + 4 * vkb + i < g.K && n0 + 4 * vnb + c < g.N ? rb[i][c] : 0.0f;
#pragma unroll
        for (int q = 0; q < 3; ++q) put4(Bs[q], 4 * vnb + c, 4 * vkb, vb, q);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = tid + 256 * j;
        const int n = e & 127, kb = (e >> 7) * 4;
        float vb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) vb[i] = k0 + kb + i < g.K && n0 + n < g.N ? rb[j][i] : 0.0f;
#pragma unroll
        for (int q = 0; q < 3; ++q) put4(Bs[q], n, kb, vb, q);
      }
    }
  };
  gload(0);
  for (int k0 = 0; k0 < g.K; k0 += BKS) {
    lstore(k0);
    __syncthreads();
    if (k0 + BKS < g.K) gload(k0 + BKS);  // next slice in flight during the MFMAs
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    // B fragments held for the slice, A fragments read per 16-row block:
    // per (i, j) the same order of the six products as SKL_BF16X3_STEP
    bf16x8 b[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        b[j][q] = *reinterpret_cast<const bf16x8*>(&Bs[q][wn * 64 + j * 16 + fr][fk]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      bf16x8 a[3];
#pragma unroll
      for (int q = 0; q < 3; ++q)
        a[q] = *reinterpret_cast<const bf16x8*>(&As[q][wm * 64 + i * 16 + fr][fk]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc0[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[j][0], acc0[i][j], 0, 0, 0);
      constexpr int pa[5] = {1, 0, 2, 0, 1}, pb[5] = {1, 2, 0, 1, 0};
#pragma unroll
      for (int t = 0; t < 5; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc1[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[pa[t]], b[j][pb[t]],
                                                               acc1[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  const int cc = lane & 15, cr = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + cc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + cr + r;
        if (row < g.M && col < g.N) as_global(g.C)[(size_t)row * g.N + col] = acc0[i][j][r] + acc1[i][j][r];
      }
    }
}

}  // namespace

// Adjoint kernels of the float skeletons that do not depend on user code
// (built with hipcc for gfx950):
//   broadcast         d in_i = d out  (Reduce(+))
//   stencil transpose d in = A^T d out for the weighted stencil
//                     out[r][c] = sum_{dy,dx} w[dy][dx] view(r+dy, c+dx)
//                     under NEUTRAL (zero) or NEAREST (clamped) padding.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"
#include "ld_util.h"

namespace {

template <typename T>
__global__ __launch_bounds__(256) void broadcast_kernel(T* __restrict__ dst,
                                                        const T* __restrict__ src, size_t n)
{
  const T v = src[0];
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    dst[i] = v;
}

// d in[p][q] = sum_{dy,dx} w[dy+R][dx+R] * sum_{r in S(p,dy)} sum_{c in S'(q,dx)} g[r][c]
// where S(p,dy) = {r in [0,rows): view row of r+dy is p}.  For NEUTRAL the
// view of an out-of-range index is the neutral zero (no contribution), so
// S = {p-dy} if in range; for NEAREST the clamped index equals p for
// r+dy <= 0 (p = 0) or r+dy >= rows-1 (p = rows-1) as well.
// A 64x4 work-group computes a 64x16 output tile; d out tile + halo in LDS.
constexpr int TW = 64, TH = 16;

__global__ __launch_bounds__(256) void stencil_transpose_kernel(
    float* __restrict__ din, const float* __restrict__ g, const float* __restrict__ w, int rows,
    int cols, int R, int nearest)
{
  extern __shared__ float sm[];
  const int TWP = TW + 2 * R, THP = TH + 2 * R;
  float* tile = sm;                   // THP x TWP of g (zero outside)
  float* ws = sm + THP * TWP;         // (2R+1)^2 weights
  const int W = 2 * R + 1;
  for (int i = threadIdx.x; i < W * W; i += 256) ws[i] = w[i];
  const int c0 = blockIdx.x * TW, r0 = blockIdx.y * TH;
  for (int idx = threadIdx.x; idx < THP * TWP; idx += 256) {
    const int tr = idx / TWP, tc = idx - tr * TWP;
    const int gr = r0 - R + tr, gc = c0 - R + tc;
    tile[idx] = ld_or(g, (size_t)gr * cols + gc, 0, gr >= 0 && gr < rows && gc >= 0 && gc < cols,
                      0.0f);
  }
  __syncthreads();
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int q = c0 + tx;
  for (int i = 0; i < TH / 4; ++i) {
    const int lr = ty * (TH / 4) + i;
    const int p = r0 + lr;
    if (p >= rows || q >= cols) continue;
    float acc = 0.0f;
    const bool edge = nearest && (p < R || p >= rows - R || q < R || q >= cols - R);
    if (!edge) {
      // interior: plain transposed stencil, d in[p][q] = sum w[dy][dx] g[p-dy][q-dx]
      for (int dy = -R; dy <= R; ++dy)
        for (int dx = -R; dx <= R; ++dx)
          acc = fmaf(ws[(dy + R) * W + dx + R], tile[(lr + R - dy) * TWP + (tx + R - dx)], acc);
    } else {
      for (int dy = -R; dy <= R; ++dy) {
        // rows r with clamp(r+dy) == p
        int ra = p - dy, rb = p - dy;
        if (p == 0) ra = 0;                 // r + dy <= 0
        if (p == rows - 1) rb = rows - 1;   // r + dy >= rows-1
        if (ra < 0) ra = 0;
        if (rb > rows - 1) rb = rows - 1;
        if (p != 0 && (p - dy < 0)) continue;
        if (p != rows - 1 && (p - dy > rows - 1)) continue;
        for (int dx = -R; dx <= R; ++dx) {
          int ca = q - dx, cb = q - dx;
          if (q == 0) ca = 0;
          if (q == cols - 1) cb = cols - 1;
          if (ca < 0) ca = 0;
          if (cb > cols - 1) cb = cols - 1;
          if (q != 0 && (q - dx < 0)) continue;
          if (q != cols - 1 && (q - dx > cols - 1)) continue;
          float s = 0.0f;
          for (int r = ra; r <= rb; ++r)
            for (int c = ca; c <= cb; ++c) s += g[(size_t)r * cols + c];
          acc = fmaf(ws[(dy + R) * W + dx + R], s, acc);
        }
      }
    }
    din[(size_t)p * cols + q] = acc;
  }
}

}  // namespace

extern "C" int skl_internal_broadcast(void* dst, const void* src, size_t elem, size_t count,
                                      hipStream_t s)
{
  size_t blocks = (count + 1023) / 1024;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) blocks = 1;
  switch (elem) {
    case 4:
      broadcast_kernel<uint32_t><<<(unsigned)blocks, 256, 0, s>>>((uint32_t*)dst,
                                                                 (const uint32_t*)src, count);
      break;
    case 8:
      broadcast_kernel<uint64_t><<<(unsigned)blocks, 256, 0, s>>>((uint64_t*)dst,
                                                                 (const uint64_t*)src, count);
      break;
    default:
      return SKL_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

extern "C" int skl_internal_stencil_transpose(float* din, const float* dout, const float* w,
                                              int rows, int cols, int R, int nearest,
                                              hipStream_t s)
{
  if (R < 0 || R > 32) return SKL_ERR_INVALID;
  const size_t lds = ((size_t)(TH + 2 * R) * (TW + 2 * R) + (2 * R + 1) * (2 * R + 1)) * 4;
  dim3 grid((cols + TW - 1) / TW, (rows + TH - 1) / TH);
  if (grid.y > 65535) return SKL_ERR_INVALID;
  stencil_transpose_kernel<<<grid, 256, lds, s>>>(din, dout, w, rows, cols, R, nearest);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

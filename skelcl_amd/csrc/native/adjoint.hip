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

extern "C" int skl_internal_stencil_adjoint_zero(float* din, const float* g, const float* w,
                                                 int rows, int cols, int R, hipStream_t s);

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

// d in[p][q] under NEAREST padding for a pixel within R of the border: the
// clamped views that land on row p (column q) from beyond the border add up
// (g read from memory; ws = the (2R+1)^2 weights).
__device__ float nearest_edge_value(const float* __restrict__ g, const float* ws, int p, int q,
                                    int rows, int cols, int R)
{
  const int W = 2 * R + 1;
  float acc = 0.0f;
  for (int dy = -R; dy <= R; ++dy) {
    // rows r in [0, rows) with clamp(r+dy) == p: exactly p - dy, extended
    // down to 0 when p is the first row (r + dy <= 0) and up to rows - 1
    // when it is the last (r + dy >= rows - 1); an empty range when ra > rb
    int ra = p == 0 ? 0 : p - dy;
    int rb = p == rows - 1 ? rows - 1 : p - dy;
    if (ra < 0) ra = 0;
    if (rb > rows - 1) rb = rows - 1;
    for (int dx = -R; dx <= R; ++dx) {
      int ca = q == 0 ? 0 : q - dx;
      int cb = q == cols - 1 ? cols - 1 : q - dx;
      if (ca < 0) ca = 0;
      if (cb > cols - 1) cb = cols - 1;
      if (ra > rb || ca > cb) continue;
      float s = 0.0f;
      for (int r = ra; r <= rb; ++r)
        for (int c = ca; c <= cb; ++c) s += g[(size_t)r * cols + c];
      acc = fmaf(ws[(dy + R) * W + dx + R], s, acc);
    }
  }
  return acc;
}

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
      acc = nearest_edge_value(g, ws, p, q, rows, cols, R);
    }
    din[(size_t)p * cols + q] = acc;
  }
}

// NEAREST padding: rewrites the pixels within R of the border (one thread
// each) after the interior formula has written every pixel.  Bands: top
// rows, bottom rows, then the left and right columns of the rows between.
__global__ __launch_bounds__(256) void stencil_transpose_border(float* __restrict__ din,
                                                                const float* __restrict__ g,
                                                                const float* __restrict__ w,
                                                                int rows, int cols, int R)
{
  const long long top = R < rows ? R : rows;
  const long long bstart = rows - R > top ? rows - R : top;  // first bottom-band row
  const long long bottom = rows - bstart, mid = bstart - top;
  const long long left = R < cols ? R : cols;
  const long long rstart = cols - R > left ? cols - R : left;  // first right-band column
  const long long side = left + (cols - rstart);
  const long long total = (top + bottom) * cols + mid * side;
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  int p, q;
  if (e < top * cols) {
    p = (int)(e / cols);
    q = (int)(e % cols);
  } else if (e < (top + bottom) * cols) {
    const long long e2 = e - top * cols;
    p = (int)(bstart + e2 / cols);
    q = (int)(e2 % cols);
  } else {
    const long long e3 = e - (top + bottom) * cols, k = e3 % side;
    p = (int)(top + e3 / side);
    q = (int)(k < left ? k : rstart + (k - left));
  }
  din[(size_t)p * cols + q] = nearest_edge_value(g, w, p, q, rows, cols, R);
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
  if (R >= 1 && R <= 5) {
    // every pixel by the zero-padding formula on the register-tiled stencil
    // kernel (native/stencil.hip, adjoint mode: the same terms in the same
    // order as the one-output kernel below); NEAREST then rewrites the border
    const int st = skl_internal_stencil_adjoint_zero(din, dout, w, rows, cols, R, s);
    if (st != SKL_OK) return st;
    if (nearest) {
      const long long top = R < rows ? R : rows;
      const long long bstart = rows - R > top ? rows - R : top;
      const long long left = R < cols ? R : cols;
      const long long rstart = cols - R > left ? cols - R : left;
      const long long total = (top + (rows - bstart)) * cols +
                              (bstart - top) * (left + (cols - rstart));
      if (total > 0)
        stencil_transpose_border<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(din, dout, w,
                                                                                rows, cols, R);
    }
    return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
  }
  const size_t lds = ((size_t)(TH + 2 * R) * (TW + 2 * R) + (2 * R + 1) * (2 * R + 1)) * 4;
  dim3 grid((cols + TW - 1) / TW, (rows + TH - 1) / TH);
  if (grid.y > 65535) return SKL_ERR_INVALID;
  stencil_transpose_kernel<<<grid, 256, lds, s>>>(din, dout, w, rows, cols, R, nearest);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

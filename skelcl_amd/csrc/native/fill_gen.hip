// Fill and synthetic-input kernels (built with hipcc for gfx950).
// The generators are counter-based (SplitMix64 finaliser applied to
// seed + golden*(i+1)), so element i has the same bits on any device or host.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t seed, uint64_t i)
{
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float unit_f32(uint64_t z)
{
  return (float)(uint32_t)(z >> 40) * 0x1p-24f;
}

constexpr int kBlock = 256;

unsigned grid_for(size_t n, size_t per_thread)
{
  size_t threads = (n + per_thread - 1) / per_thread;
  size_t blocks = (threads + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  return (unsigned)(blocks ? blocks : 1);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fill_kernel(T* __restrict__ dst,
                                                      T value, size_t count)
{
  size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < count;
       i += stride)
    dst[i] = value;
}

struct alignas(16) b128 {
  uint32_t w[4];
};

__global__ __launch_bounds__(kBlock) void gen_f32_sym_kernel(
    float* __restrict__ out, size_t n, uint64_t seed)
{
  size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    out[i] = 2.0f * unit_f32(splitmix64(seed, i)) - 1.0f;
}

__global__ __launch_bounds__(kBlock) void gen_f32_unit_kernel(
    float* __restrict__ out, size_t n, uint64_t seed, float scale)
{
  size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    out[i] = scale * unit_f32(splitmix64(seed, i));
}

__global__ __launch_bounds__(kBlock) void gen_i32_below_kernel(
    int32_t* __restrict__ out, size_t n, uint64_t seed, uint32_t m)
{
  size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += stride) {
    uint64_t hi = splitmix64(seed, i) >> 32;
    out[i] = (int32_t)((hi * (uint64_t)m) >> 32);
  }
}

int status(hipError_t e)
{
  return e == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

}  // namespace

extern "C" int skl_internal_fill(void* dst, const void* pattern, size_t elem,
                                 size_t count, hipStream_t s)
{
  unsigned g = grid_for(count, 4);
  switch (elem) {
    case 1:
      fill_kernel<uint8_t><<<g, kBlock, 0, s>>>(
          (uint8_t*)dst, *(const uint8_t*)pattern, count);
      break;
    case 2:
      fill_kernel<uint16_t><<<g, kBlock, 0, s>>>(
          (uint16_t*)dst, *(const uint16_t*)pattern, count);
      break;
    case 4:
      fill_kernel<uint32_t><<<g, kBlock, 0, s>>>(
          (uint32_t*)dst, *(const uint32_t*)pattern, count);
      break;
    case 8:
      fill_kernel<uint64_t><<<g, kBlock, 0, s>>>(
          (uint64_t*)dst, *(const uint64_t*)pattern, count);
      break;
    case 16: {
      b128 v;
      __builtin_memcpy(&v, pattern, 16);
      fill_kernel<b128><<<g, kBlock, 0, s>>>((b128*)dst, v, count);
      break;
    }
    default:
      return SKL_ERR_INVALID;
  }
  return status(hipGetLastError());
}

extern "C" int skl_internal_gen_f32_sym(float* out, size_t n, uint64_t seed,
                                        hipStream_t s)
{
  if (n == 0) return SKL_OK;
  gen_f32_sym_kernel<<<grid_for(n, 8), kBlock, 0, s>>>(out, n, seed);
  return status(hipGetLastError());
}

extern "C" int skl_internal_gen_f32_unit(float* out, size_t n, uint64_t seed,
                                         float scale, hipStream_t s)
{
  if (n == 0) return SKL_OK;
  gen_f32_unit_kernel<<<grid_for(n, 8), kBlock, 0, s>>>(out, n, seed, scale);
  return status(hipGetLastError());
}

extern "C" int skl_internal_gen_i32_below(int32_t* out, size_t n,
                                          uint64_t seed, uint32_t m,
                                          hipStream_t s)
{
  if (n == 0) return SKL_OK;
  gen_i32_below_kernel<<<grid_for(n, 8), kBlock, 0, s>>>(out, n, seed, m);
  return status(hipGetLastError());
}

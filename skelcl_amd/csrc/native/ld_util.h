// Branch-free guarded load for the native kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

// p[i] when ok, else `fill`, with the load outside any branch: the index
// falls back to `safe` (any valid index) and the load always issues.  A load
// under a branch is waited for before the next one is issued; these keep
// several in flight.
template <typename T>
__device__ __forceinline__ T ld_or(const T* p, size_t i, size_t safe, bool ok, T fill)
{
  const T v = p[ok ? i : safe];
  return ok ? v : fill;
}

// A pointer read from memory (a descriptor field) is generic to the
// compiler, which then emits flat loads/stores; flat loads also count on the
// LDS counter, so LDS waits start waiting for them too.  Device memory is
// global: say so.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p)
{
  return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p)
{
  return (__attribute__((address_space(1))) T*)p;
}

template <typename T>
__device__ __forceinline__ T ld_or_g(const T* p, size_t i, size_t safe, bool ok, T fill)
{
  const T v = as_global(p)[ok ? i : safe];
  return ok ? v : fill;
}

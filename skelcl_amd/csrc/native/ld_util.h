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

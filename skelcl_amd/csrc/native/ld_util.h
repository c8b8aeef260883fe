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

// Accumulators of the float (+) kernels: FP64, or a double-float pair (a
// compensated FP32 sum: Knuth's two-sum on the high words, the rounding
// errors and low words summed in the low word).  Both are 8 bytes; `acc_*`
// are overloaded on the two.
struct acc_df {
  float hi, lo;
};

template <typename A>
__device__ __forceinline__ A acc_zero();
template <>
__device__ __forceinline__ double acc_zero<double>() { return 0.0; }
template <>
__device__ __forceinline__ acc_df acc_zero<acc_df>() { return acc_df{0.0f, 0.0f}; }

__device__ __forceinline__ double acc_add(double a, double b) { return a + b; }
__device__ __forceinline__ double acc_add(double a, float x) { return a + (double)x; }
__device__ __forceinline__ acc_df acc_add(const acc_df& a, const acc_df& b)
{
  const float s = a.hi + b.hi;
  const float bb = s - a.hi;
  const float e = (a.hi - (s - bb)) + (b.hi - bb);
  return acc_df{s, (a.lo + b.lo) + e};
}
__device__ __forceinline__ acc_df acc_add(const acc_df& a, float x)
{
  const float s = a.hi + x;
  const float bb = s - a.hi;
  const float e = (a.hi - (s - bb)) + (x - bb);
  return acc_df{s, a.lo + e};
}
__device__ __forceinline__ double acc_to_double(double a) { return a; }
__device__ __forceinline__ double acc_to_double(const acc_df& a)
{
  return (double)a.hi + (double)a.lo;
}

// the two dwords of an accumulator, and back
__device__ __forceinline__ void acc_bits(double v, int* lo, int* hi)
{
  const long long b = __double_as_longlong(v);
  *lo = (int)b;
  *hi = (int)(b >> 32);
}
__device__ __forceinline__ void acc_bits(const acc_df& v, int* lo, int* hi)
{
  *lo = __float_as_int(v.hi);
  *hi = __float_as_int(v.lo);
}
template <typename A>
__device__ __forceinline__ A acc_from_bits(int lo, int hi);
template <>
__device__ __forceinline__ double acc_from_bits<double>(int lo, int hi)
{
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <>
__device__ __forceinline__ acc_df acc_from_bits<acc_df>(int lo, int hi)
{
  return acc_df{__int_as_float(lo), __int_as_float(hi)};
}

// Exact int32 (+) accumulator of the segmented forms: the element bits read
// as an integer, sums modulo 2^32 (as the reference's int Scan and Reduce
// wrap).
struct acc_u32 {
  unsigned v;
};
template <>
__device__ __forceinline__ acc_u32 acc_zero<acc_u32>() { return acc_u32{0u}; }
__device__ __forceinline__ acc_u32 acc_add(const acc_u32& a, const acc_u32& b)
{
  return acc_u32{a.v + b.v};
}
__device__ __forceinline__ acc_u32 acc_add(const acc_u32& a, float x)
{
  return acc_u32{a.v + __float_as_uint(x)};
}
__device__ __forceinline__ void acc_bits(const acc_u32& v, int* lo, int* hi)
{
  *lo = (int)v.v;
  *hi = 0;
}
template <>
__device__ __forceinline__ acc_u32 acc_from_bits<acc_u32>(int lo, int)
{
  return acc_u32{(unsigned)lo};
}


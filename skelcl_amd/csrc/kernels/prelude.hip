// ---------------------------------------------------------------------------
// Prelude of every fused skeleton kernel (compiled at run time by hiprtc for
// gfx950).  It gives user functions written for SkelCL the OpenCL C
// vocabulary they use -- address-space qualifiers, work-item functions,
// barrier, mad24, the vector types with .s0/.x members, T_matrix_t with
// get/set (MatrixDef.h:834-867), Index/IndexPoint (MapDef.h:419-423,
// 610-615), skelcl_get_device_id() (Program.cpp:229-250) -- as plain HIP.
// This is source compatibility for user strings, not an OpenCL runtime.
// ---------------------------------------------------------------------------
#undef __kernel
#undef __global
#undef __local
#undef __constant
#undef __private
#undef __read_only
#undef __write_only
#define __kernel extern "C" __global__
#define __global
#define __local
#define __constant
#define __private
#define __read_only
#define __write_only
#define CLK_LOCAL_MEM_FENCE 1
#define CLK_GLOBAL_MEM_FENCE 2
#define skelcl_get_device_id() 0

typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
typedef unsigned long ulong;

__device__ __forceinline__ size_t get_global_id(uint d)
{
  return d == 0 ? (size_t)blockIdx.x * blockDim.x + threadIdx.x
       : d == 1 ? (size_t)blockIdx.y * blockDim.y + threadIdx.y
                : (size_t)blockIdx.z * blockDim.z + threadIdx.z;
}
__device__ __forceinline__ size_t get_local_id(uint d)
{
  return d == 0 ? threadIdx.x : d == 1 ? threadIdx.y : threadIdx.z;
}
__device__ __forceinline__ size_t get_group_id(uint d)
{
  return d == 0 ? blockIdx.x : d == 1 ? blockIdx.y : blockIdx.z;
}
__device__ __forceinline__ size_t get_local_size(uint d)
{
  return d == 0 ? blockDim.x : d == 1 ? blockDim.y : blockDim.z;
}
__device__ __forceinline__ size_t get_num_groups(uint d)
{
  return d == 0 ? gridDim.x : d == 1 ? gridDim.y : gridDim.z;
}
__device__ __forceinline__ size_t get_global_size(uint d)
{
  return get_num_groups(d) * get_local_size(d);
}
__device__ __forceinline__ uint get_work_dim() { return 1; }
__device__ __forceinline__ void barrier(int) { __syncthreads(); }
__device__ __forceinline__ void mem_fence(int) { __threadfence(); }
__device__ __forceinline__ int mad24(int a, int b, int c) { return a * b + c; }

// p[i] when ok, else `fill`, with the load outside any branch: the index
// falls back to `safe` (any valid index) and the load always issues.  A load
// under a branch is waited for before the next one is issued; these keep
// several in flight.
template <typename T>
__device__ __forceinline__ T skl_ld_or(const T* p, size_t i, size_t safe, bool ok, T fill)
{
  const T v = p[ok ? i : safe];
  return ok ? v : fill;
}
__device__ __forceinline__ uint mad24(uint a, uint b, uint c) { return a * b + c; }
__device__ __forceinline__ int mul24(int a, int b) { return a * b; }
__device__ __forceinline__ uint mul24(uint a, uint b) { return a * b; }
__device__ __forceinline__ float mad(float a, float b, float c) { return a * b + c; }
__device__ __forceinline__ double mad(double a, double b, double c) { return a * b + c; }
template <typename T>
__device__ __forceinline__ T clamp(T x, T lo, T hi)
{
  return x < lo ? lo : (x > hi ? hi : x);
}
#define native_exp expf
#define native_log logf
#define native_sqrt sqrtf
#define native_recip(x) (1.0f / (x))
#define native_divide(a, b) ((a) / (b))
#define half_exp expf

// ---- OpenCL vector types (Types.h:43-152 host side: fields s0..sf) --------
template <typename T, int N> struct skl_vec;
template <typename T> struct alignas(2 * sizeof(T)) skl_vec<T, 2> {
  union { T v[2]; struct { T s0, s1; }; struct { T x, y; }; };
};
template <typename T> struct alignas(4 * sizeof(T)) skl_vec<T, 4> {
  union { T v[4]; struct { T s0, s1, s2, s3; }; struct { T x, y, z, w; }; };
};
template <typename T> struct alignas(8 * sizeof(T)) skl_vec<T, 8> {
  union { T v[8]; struct { T s0, s1, s2, s3, s4, s5, s6, s7; }; };
};
template <typename T> struct alignas(16 * sizeof(T) > 128 ? 128 : 16 * sizeof(T)) skl_vec<T, 16> {
  union { T v[16]; struct { T s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, sa, sb, sc, sd, se, sf; }; };
};
template <typename S> struct skl_is_scalar { static constexpr bool value = __is_arithmetic(S); };
template <bool B, typename T = void> struct skl_enable_if {};
template <typename T> struct skl_enable_if<true, T> { typedef T type; };

#define SKL_VEC_BINOP(OP)                                                      \
  template <typename T, int N>                                                 \
  __device__ __forceinline__ skl_vec<T, N> operator OP(const skl_vec<T, N>& a, \
                                                       const skl_vec<T, N>& b) \
  {                                                                            \
    skl_vec<T, N> r;                                                           \
    _Pragma("unroll") for (int i = 0; i < N; ++i) r.v[i] = a.v[i] OP b.v[i];   \
    return r;                                                                  \
  }                                                                            \
  template <typename T, int N, typename S>                                     \
  __device__ __forceinline__                                                   \
      typename skl_enable_if<skl_is_scalar<S>::value, skl_vec<T, N>>::type     \
      operator OP(const skl_vec<T, N>& a, S b)                                 \
  {                                                                            \
    skl_vec<T, N> r;                                                           \
    _Pragma("unroll") for (int i = 0; i < N; ++i) r.v[i] = a.v[i] OP(T) b;     \
    return r;                                                                  \
  }                                                                            \
  template <typename T, int N, typename S>                                     \
  __device__ __forceinline__                                                   \
      typename skl_enable_if<skl_is_scalar<S>::value, skl_vec<T, N>>::type     \
      operator OP(S a, const skl_vec<T, N>& b)                                 \
  {                                                                            \
    skl_vec<T, N> r;                                                           \
    _Pragma("unroll") for (int i = 0; i < N; ++i) r.v[i] = (T)a OP b.v[i];     \
    return r;                                                                  \
  }                                                                            \
  template <typename T, int N, typename S>                                     \
  __device__ __forceinline__ skl_vec<T, N>& operator OP##=(skl_vec<T, N>& a,   \
                                                           const S& b)         \
  {                                                                            \
    a = a OP b;                                                                \
    return a;                                                                  \
  }
SKL_VEC_BINOP(+)
SKL_VEC_BINOP(-)
SKL_VEC_BINOP(*)
SKL_VEC_BINOP(/)
#undef SKL_VEC_BINOP
template <typename T, int N>
__device__ __forceinline__ skl_vec<T, N> operator-(const skl_vec<T, N>& a)
{
  skl_vec<T, N> r;
#pragma unroll
  for (int i = 0; i < N; ++i) r.v[i] = -a.v[i];
  return r;
}

#define SKL_VEC_TYPES(T, NAME)                                                 \
  typedef skl_vec<T, 2> skl_##NAME##2;                                         \
  typedef skl_vec<T, 4> skl_##NAME##4;                                         \
  typedef skl_vec<T, 8> skl_##NAME##8;                                         \
  typedef skl_vec<T, 16> skl_##NAME##16;
SKL_VEC_TYPES(char, char)
SKL_VEC_TYPES(unsigned char, uchar)
SKL_VEC_TYPES(short, short)
SKL_VEC_TYPES(unsigned short, ushort)
SKL_VEC_TYPES(int, int)
SKL_VEC_TYPES(unsigned int, uint)
SKL_VEC_TYPES(long, long)
SKL_VEC_TYPES(unsigned long, ulong)
SKL_VEC_TYPES(float, float)
SKL_VEC_TYPES(double, double)
#undef SKL_VEC_TYPES
// the OpenCL names (3-component types are 4-component, Types.h:96-147)
#define char2 skl_char2
#define char3 skl_char4
#define char4 skl_char4
#define char8 skl_char8
#define char16 skl_char16
#define uchar2 skl_uchar2
#define uchar3 skl_uchar4
#define uchar4 skl_uchar4
#define uchar8 skl_uchar8
#define uchar16 skl_uchar16
#define short2 skl_short2
#define short3 skl_short4
#define short4 skl_short4
#define short8 skl_short8
#define short16 skl_short16
#define ushort2 skl_ushort2
#define ushort3 skl_ushort4
#define ushort4 skl_ushort4
#define ushort8 skl_ushort8
#define ushort16 skl_ushort16
#define int2 skl_int2
#define int3 skl_int4
#define int4 skl_int4
#define int8 skl_int8
#define int16 skl_int16
#define uint2 skl_uint2
#define uint3 skl_uint4
#define uint4 skl_uint4
#define uint8 skl_uint8
#define uint16 skl_uint16
#define long2 skl_long2
#define long3 skl_long4
#define long4 skl_long4
#define long8 skl_long8
#define long16 skl_long16
#define ulong2 skl_ulong2
#define ulong3 skl_ulong4
#define ulong4 skl_ulong4
#define ulong8 skl_ulong8
#define ulong16 skl_ulong16
#define float2 skl_float2
#define float3 skl_float4
#define float4 skl_float4
#define float8 skl_float8
#define float16 skl_float16
#define double2 skl_double2
#define double3 skl_double4
#define double4 skl_double4
#define double8 skl_double8
#define double16 skl_double16

// ---- Matrix arguments (MatrixDef.h:834-867): T_matrix_t + get/set --------
#define SKL_MATRIX_T(T, NAME)                                                  \
  typedef struct {                                                             \
    T* data;                                                                   \
    unsigned int col_count;                                                    \
  } NAME##_matrix_t;
SKL_MATRIX_T(char, char)
SKL_MATRIX_T(unsigned char, uchar)
SKL_MATRIX_T(short, short)
SKL_MATRIX_T(unsigned short, ushort)
SKL_MATRIX_T(int, int)
SKL_MATRIX_T(unsigned int, uint)
SKL_MATRIX_T(long, long)
SKL_MATRIX_T(unsigned long, ulong)
SKL_MATRIX_T(float, float)
SKL_MATRIX_T(double, double)
SKL_MATRIX_T(skl_float2, float2)
SKL_MATRIX_T(skl_float4, float4)
SKL_MATRIX_T(skl_int4, int4)
SKL_MATRIX_T(skl_double4, double4)
#undef SKL_MATRIX_T
// the names above are taken (Matrix<T>::deviceFunctions guards on these)
#define char_MATRIX_T
#define uchar_MATRIX_T
#define short_MATRIX_T
#define ushort_MATRIX_T
#define int_MATRIX_T
#define uint_MATRIX_T
#define long_MATRIX_T
#define ulong_MATRIX_T
#define float_MATRIX_T
#define double_MATRIX_T
#define float2_MATRIX_T
#define float4_MATRIX_T
#define int4_MATRIX_T
#define double4_MATRIX_T
#define get(m, y, x) (m).data[(int)((y) * (m).col_count + (x))]
#define set(m, y, x, v) (m).data[(int)((y) * (m).col_count + (x))] = (v)

// ---- index types ----------------------------------------------------------
typedef size_t Index;
typedef struct {
  size_t x;
  size_t y;
} IndexPoint;

// ---- skeleton-side helpers ------------------------------------------------
// Cross-lane moves of any trivially copyable type, 32 bits at a time.
template <typename T>
__device__ __forceinline__ T skl_shfl_down(const T& v, unsigned d)
{
  constexpr int W = (int)((sizeof(T) + 3) / 4);
  int w[W];
  T r;
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < W; ++i) w[i] = __shfl_down(w[i], d, 64);
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}
template <typename T>
__device__ __forceinline__ T skl_shfl_up(const T& v, unsigned d)
{
  constexpr int W = (int)((sizeof(T) + 3) / 4);
  int w[W];
  T r;
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < W; ++i) w[i] = __shfl_up(w[i], d, 64);
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}
template <typename T>
__device__ __forceinline__ T skl_shfl(const T& v, int src)
{
  constexpr int W = (int)((sizeof(T) + 3) / 4);
  int w[W];
  T r;
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < W; ++i) w[i] = __shfl(w[i], src, 64);
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// DPP moves (VALU, no LDS), 32 bits at a time: lane l takes the value of the
// lane that CTRL names inside the rows of ROW_MASK; other lanes keep theirs.
template <int CTRL, int ROW_MASK, typename T>
__device__ __forceinline__ T skl_dpp(const T& v)
{
  constexpr int W = (int)((sizeof(T) + 3) / 4);
  int w[W];
  T r;
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < W; ++i)
    w[i] = __builtin_amdgcn_update_dpp(w[i], w[i], CTRL, ROW_MASK, 0xf, false);
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}
template <typename T>
__device__ __forceinline__ T skl_readlane(const T& v, int src)
{
  constexpr int W = (int)((sizeof(T) + 3) / 4);
  int w[W];
  T r;
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < W; ++i) w[i] = __builtin_amdgcn_readlane(w[i], src);
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Inclusive wave64 scan in lane order (lane l ends with v_0 (+) .. (+) v_l;
// comb need not commute): row_shr 1, 2, 4, 8 inside each 16-lane row, then
// row_bcast 15 (lane 15 of rows 0 and 2 into rows 1 and 3) and row_bcast 31
// (lane 31 into rows 2 and 3).
template <typename T, typename F>
__device__ __forceinline__ T skl_wave_scan(T v, int lane, F comb)
{
  T o;
  o = skl_dpp<0x111, 0xf>(v);
  if ((lane & 15) >= 1) v = comb(o, v);
  o = skl_dpp<0x112, 0xf>(v);
  if ((lane & 15) >= 2) v = comb(o, v);
  o = skl_dpp<0x114, 0xf>(v);
  if ((lane & 15) >= 4) v = comb(o, v);
  o = skl_dpp<0x118, 0xf>(v);
  if ((lane & 15) >= 8) v = comb(o, v);
  o = skl_dpp<0x142, 0xa>(v);
  if (lane & 16) v = comb(o, v);
  o = skl_dpp<0x143, 0xc>(v);
  if (lane >= 32) v = comb(o, v);
  return v;
}
// value of lane - 1 (lane 0 keeps its own)
template <typename T>
__device__ __forceinline__ T skl_wave_shr1(const T& v) { return skl_dpp<0x138, 0xf>(v); }

// Packs of N elements moved with one wide load/store.
template <typename T, int N> struct alignas(N * sizeof(T)) skl_pack { T v[N]; };
template <typename T> struct skl_pack<T, 1> { T v[1]; };

// Streaming pack load / store.  16- and 8-byte packs go through the
// non-temporal forms (measured on MI355X at 2^28 elements: copy, saxpy, Scan
// and Reduce 4-7% faster); the build option -DSKL_NO_NT turns that off, e.g.
// for small inputs that are read again soon.
#ifndef SKL_NO_NT
#ifndef SKL_NT_LOAD
#define SKL_NT_LOAD 1
#endif
#ifndef SKL_NT_STORE
#define SKL_NT_STORE 1
#endif
#endif
template <typename P>
__device__ __forceinline__ P skl_load_pack(const P* src)
{
#ifdef SKL_NT_LOAD
  if constexpr (sizeof(P) == 16) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const u4 t = __builtin_nontemporal_load(reinterpret_cast<const u4*>(src));
    P v;
    __builtin_memcpy(&v, &t, 16);
    return v;
  } else if constexpr (sizeof(P) == 8) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    const u2 t = __builtin_nontemporal_load(reinterpret_cast<const u2*>(src));
    P v;
    __builtin_memcpy(&v, &t, 8);
    return v;
  }
#endif
  return *src;
}

template <typename P>
__device__ __forceinline__ void skl_store_pack(P* dst, const P& v)
{
#ifdef SKL_NT_STORE
  if constexpr (sizeof(P) == 16) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    u4 t;
    __builtin_memcpy(&t, &v, 16);
    __builtin_nontemporal_store(t, reinterpret_cast<u4*>(dst));
    return;
  } else if constexpr (sizeof(P) == 8) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    u2 t;
    __builtin_memcpy(&t, &v, 8);
    __builtin_nontemporal_store(t, reinterpret_cast<u2*>(dst));
    return;
  }
#endif
  *dst = v;
}

// Pack load / store at an element address that is only 4-byte aligned
// (sizeof(T) % 4 == 0): one dword-aligned vector access, which gfx9 memory
// instructions accept.
template <typename P, typename T>
__device__ __forceinline__ P skl_load_pack_u(const T* src)
{
  static_assert(sizeof(T) % 4 == 0, "dword-aligned packs need 4-byte elements");
  P v;
  if constexpr (sizeof(P) == 16) {
    typedef unsigned int u4a __attribute__((ext_vector_type(4), aligned(4)));
#ifdef SKL_NT_LOAD
    const u4a t = __builtin_nontemporal_load(reinterpret_cast<const u4a*>(src));
#else
    const u4a t = *reinterpret_cast<const u4a*>(src);
#endif
    __builtin_memcpy(&v, &t, 16);
  } else if constexpr (sizeof(P) == 8) {
    typedef unsigned int u2a __attribute__((ext_vector_type(2), aligned(4)));
#ifdef SKL_NT_LOAD
    const u2a t = __builtin_nontemporal_load(reinterpret_cast<const u2a*>(src));
#else
    const u2a t = *reinterpret_cast<const u2a*>(src);
#endif
    __builtin_memcpy(&v, &t, 8);
  } else {
    __builtin_memcpy(&v, src, sizeof(P));
  }
  return v;
}

template <typename P, typename T>
__device__ __forceinline__ void skl_store_pack_u(T* dst, const P& v)
{
  static_assert(sizeof(T) % 4 == 0, "dword-aligned packs need 4-byte elements");
  if constexpr (sizeof(P) == 16) {
    typedef unsigned int u4a __attribute__((ext_vector_type(4), aligned(4)));
    u4a t;
    __builtin_memcpy(&t, &v, 16);
#ifdef SKL_NT_STORE
    __builtin_nontemporal_store(t, reinterpret_cast<u4a*>(dst));
#else
    *reinterpret_cast<u4a*>(dst) = t;
#endif
  } else if constexpr (sizeof(P) == 8) {
    typedef unsigned int u2a __attribute__((ext_vector_type(2), aligned(4)));
    u2a t;
    __builtin_memcpy(&t, &v, 8);
#ifdef SKL_NT_STORE
    __builtin_nontemporal_store(t, reinterpret_cast<u2a*>(dst));
#else
    *reinterpret_cast<u2a*>(dst) = t;
#endif
  } else {
    __builtin_memcpy(dst, &v, sizeof(P));
  }
}

// Double-float pair used by the compensated float paths.
struct skl_df {
  float hi, lo;
};
__device__ __forceinline__ skl_df skl_df_add(const skl_df& a, const skl_df& b)
{
  // Knuth two-sum on the high parts; low parts and the rounding error are
  // accumulated in the low word.
  float s = a.hi + b.hi;
  float bb = s - a.hi;
  float e = (a.hi - (s - bb)) + (b.hi - bb);
  skl_df r;
  r.hi = s;
  r.lo = (a.lo + b.lo) + e;
  return r;
}

extern __shared__ __attribute__((aligned(16))) char skl_dyn_lds[];

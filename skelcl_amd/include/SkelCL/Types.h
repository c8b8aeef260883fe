// OpenCL-style vector element types (detail/Types.h:43-152): plain structs
// with members s0..sN, the same layout as the device-side float4 etc.
#pragma once

#include "detail/Runtime.h"

namespace skelcl {

#define SKELCL_DEF_VEC2(T, N)                                                  \
  struct N##2 {                                                                \
    T s0, s1;                                                                  \
  };
#define SKELCL_DEF_VEC4(T, N)                                                  \
  struct N##4 {                                                                \
    T s0, s1, s2, s3;                                                          \
  };
#define SKELCL_DEF_VEC8(T, N)                                                  \
  struct N##8 {                                                                \
    T s0, s1, s2, s3, s4, s5, s6, s7;                                          \
  };
#define SKELCL_DEF_VEC16(T, N)                                                 \
  struct N##16 {                                                               \
    T s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, sa, sb, sc, sd, se, sf;          \
  };
#define SKELCL_DEF_VECS(T, N)                                                  \
  SKELCL_DEF_VEC2(T, N)                                                        \
  SKELCL_DEF_VEC4(T, N)                                                        \
  typedef N##4 N##3;                                                           \
  SKELCL_DEF_VEC8(T, N)                                                        \
  SKELCL_DEF_VEC16(T, N)                                                       \
  SKELCL_TYPE_NAME(N##2, #N "2")                                               \
  SKELCL_TYPE_NAME(N##4, #N "4")                                               \
  SKELCL_TYPE_NAME(N##8, #N "8")                                               \
  SKELCL_TYPE_NAME(N##16, #N "16")

SKELCL_DEF_VECS(char, char)
SKELCL_DEF_VECS(unsigned char, uchar)
SKELCL_DEF_VECS(short, short)
SKELCL_DEF_VECS(unsigned short, ushort)
SKELCL_DEF_VECS(int, int)
SKELCL_DEF_VECS(unsigned int, uint)
SKELCL_DEF_VECS(long, long)
SKELCL_DEF_VECS(unsigned long, ulong)
SKELCL_DEF_VECS(float, float)
SKELCL_DEF_VECS(double, double)

#undef SKELCL_DEF_VEC2
#undef SKELCL_DEF_VEC4
#undef SKELCL_DEF_VEC8
#undef SKELCL_DEF_VEC16
#undef SKELCL_DEF_VECS

}  // namespace skelcl

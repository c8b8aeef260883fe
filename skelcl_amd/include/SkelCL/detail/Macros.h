// SKELCL_COMMON_DEFINITION / SKELCL_ADD_DEFINE (detail/Macros.h:55-61): a
// definition written once in host code that also reaches every skeleton
// kernel built afterwards.
//   SKELCL_COMMON_DEFINITION(typedef struct { uchar r, g, b; } Pixel;)
// declares Pixel in C++ and registers the same text, so Map<Pixel(...)>
// user functions can use it; SKELCL_ADD_DEFINE(WIDTH) registers
// "#define WIDTH <value of WIDTH>".
#pragma once

#include "../Source.h"

#define SKELCL_CONCAT_IMPL_(a, b) a##b
#define SKELCL_CONCAT_(a, b) SKELCL_CONCAT_IMPL_(a, b)
#ifdef __COUNTER__
#define SKELCL_UNIQUE_NAME_(base) SKELCL_CONCAT_(base, __COUNTER__)
#else
#define SKELCL_UNIQUE_NAME_(base) SKELCL_CONCAT_(base, __LINE__)
#endif

#define SKELCL_COMMON_DEFINITION(definition)                                  \
  definition static skelcl::detail::RegisterCommonDefinition                 \
      SKELCL_UNIQUE_NAME_(skelcl_common_definition_)(#definition);

#define SKELCL_ADD_DEFINE(define)                                             \
  static skelcl::detail::RegisterCommonMacroDefinition                       \
      SKELCL_UNIQUE_NAME_(skelcl_common_define_)(#define, define);

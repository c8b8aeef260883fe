// skelcl::detail::util (detail/Util.h:55-98, src/Util.cpp:149-230): the
// environment lookup, the SHA1 hash that names programs, the rounding
// helpers of the launch code, and the name a C++ type has in kernel source.
#pragma once

#include <cxxabi.h>

#include <cstddef>
#include <cstdlib>
#include <string>
#include <typeinfo>

#include "skelcl_rt.h"

namespace skelcl {
namespace detail {
namespace util {

// the variable's value, "" when it is not set
inline std::string envVarValue(const std::string& envVar)
{
  const char* v = std::getenv(envVar.c_str());
  return v ? std::string(v) : std::string();
}

// lower-case hex SHA1 of `text` (SHA1Tests.cpp:58-77)
inline std::string hash(const std::string& text)
{
  char out[41] = {0};
  skl_sha1_hex(text.data(), text.size(), out);
  return std::string(out);
}

// ceil(i / j)
inline size_t devideAndRoundUp(size_t i, size_t j) { return i / j + (i % j != 0 ? 1 : 0); }

// ceil(i / j) rounded up to a multiple of a
inline size_t devideAndAlign(size_t i, size_t j, size_t a)
{
  const size_t x = devideAndRoundUp(i, j);
  return x % a == 0 ? x : x + (a - x % a);
}

// i rounded up to a multiple of j; j for i == 0 (as the reference)
inline size_t ceilToMultipleOf(size_t i, size_t j)
{
  if (i == 0) return j;
  return i % j == 0 ? i : i + (j - i % j);
}

// true for 0 and the powers of two (as the reference's n & (n - 1) test)
inline bool isPowerOfTwo(size_t n) { return (n & (n - 1)) == 0; }

// the largest power of two <= n (n >= 1)
inline int floorPow2(int n)
{
  int p = 1;
  while (n >= 2 * p && p < (1 << 30)) p *= 2;
  return p;
}

// the smallest power of two > n, i.e. twice floorPow2(n) (n >= 1; the
// reference's frexp form also doubles an exact power of two)
inline int ceilPow2(int n) { return n < 1 ? 1 : 2 * floorPow2(n); }

// The type's name as kernel source spells it: the demangled C++ name with
// any namespace qualification removed, so skelcl::float4 is "float4" and a
// user struct Pixel is "Pixel".
template <typename T>
std::string typeToString()
{
  int status = 0;
  char* d = abi::__cxa_demangle(typeid(T).name(), nullptr, nullptr, &status);
  std::string name = (status == 0 && d) ? std::string(d) : std::string(typeid(T).name());
  std::free(d);
  const size_t colon = name.rfind(':');
  if (colon != std::string::npos) name.erase(0, colon + 1);
  return name;
}

}  // namespace util
}  // namespace detail
}  // namespace skelcl

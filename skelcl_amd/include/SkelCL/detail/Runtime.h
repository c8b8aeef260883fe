// Host-side helpers of the C++ API: error handling in the reference's
// ABORT_WITH_ERROR style (libraries/pvsutil/include/pvsutil/Logger.h:108-114),
// element type names for kernel source (detail/Util.h:81-98: the reference
// demangles C++ names; here a trait gives the name), and device memory
// ownership for the containers.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>

#include "skelcl_rt.h"
#include "skelcl_skeletons.h"
#include "Util.h"

namespace skelcl {
namespace detail {

// Errors end the program like the reference (log + std::exit(EXIT_FAILURE)),
// or throw std::runtime_error when SKELCL_THROW_ON_FAILURE is defined (the
// reference's THROW_ON_FAILURE option, CMakeLists.txt:38-39).
[[noreturn]] inline void abortWithError(const std::string& msg)
{
#ifdef SKELCL_THROW_ON_FAILURE
  throw std::runtime_error("skelcl: " + msg);
#else
  std::fprintf(stderr, "[skelcl] Error: %s\n", msg.c_str());
  std::exit(EXIT_FAILURE);
#endif
}

inline void check(int status, const char* what)
{
  if (status != SKL_OK)
    abortWithError(std::string(what) + ": " + (skl_last_error() ? skl_last_error() : ""));
}

// Device memory owned by a container (DeviceBuffer, src/DeviceBuffer.cpp).
class DeviceMemory {
public:
  DeviceMemory() = default;
  explicit DeviceMemory(size_t bytes) { allocate(bytes); }
  DeviceMemory(const DeviceMemory&) = delete;
  DeviceMemory& operator=(const DeviceMemory&) = delete;
  DeviceMemory(DeviceMemory&& o) noexcept : _ptr(o._ptr), _bytes(o._bytes)
  {
    o._ptr = nullptr;
    o._bytes = 0;
  }
  DeviceMemory& operator=(DeviceMemory&& o) noexcept
  {
    std::swap(_ptr, o._ptr);
    std::swap(_bytes, o._bytes);
    return *this;
  }
  ~DeviceMemory() { release(); }

  void allocate(size_t bytes)
  {
    release();
    if (bytes == 0) return;
    check(skl_malloc(&_ptr, bytes), "device allocation");
    _bytes = bytes;
  }
  void release()
  {
    if (_ptr) skl_free(_ptr);
    _ptr = nullptr;
    _bytes = 0;
  }
  void* get() const { return _ptr; }
  size_t bytes() const { return _bytes; }

private:
  void* _ptr = nullptr;
  size_t _bytes = 0;
};

}  // namespace detail

// ---- element type names as the kernel source spells them ------------------
// Built-in names are spelled out; any other type (skelcl::float4, a user
// struct from SKELCL_COMMON_DEFINITION) takes its demangled name without
// namespaces, as the reference's typeToString (detail/Util.h:81-98).
template <typename T> struct TypeName {
  static const char* get()
  {
    static const std::string name = detail::util::typeToString<T>();
    return name.c_str();
  }
};
#define SKELCL_TYPE_NAME(T, NAME)                                              \
  template <> struct TypeName<T> {                                             \
    static const char* get() { return NAME; }                                  \
  };
SKELCL_TYPE_NAME(float, "float")
SKELCL_TYPE_NAME(double, "double")
SKELCL_TYPE_NAME(int, "int")
SKELCL_TYPE_NAME(unsigned int, "unsigned int")
SKELCL_TYPE_NAME(long, "long")
SKELCL_TYPE_NAME(unsigned long, "unsigned long")
SKELCL_TYPE_NAME(short, "short")
SKELCL_TYPE_NAME(unsigned short, "unsigned short")
SKELCL_TYPE_NAME(char, "char")
SKELCL_TYPE_NAME(unsigned char, "unsigned char")
SKELCL_TYPE_NAME(signed char, "char")
SKELCL_TYPE_NAME(bool, "bool")

template <typename T> inline skl_type typeOf()
{
  return skl_type{TypeName<T>::get(), sizeof(T)};
}
template <> inline skl_type typeOf<void>() { return skl_type{"void", 0}; }

}  // namespace skelcl

// Internal helpers shared by the runtime and the skeleton layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "skelcl_rt.h"

namespace skl {

// Thread-local error message behind skl_last_error().
void set_error(const std::string& msg);
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// The library stream used when a caller passes NULL.
hipStream_t resolve_stream(skl_stream s);

// True once skl_init succeeded.
bool initialized();
int ensure_initialized();

// Device properties cached by skl_init.
const hipDeviceProp_t& device_props();
int device_id();

// Target architecture for hiprtc ("gfx950" unless SKELCL_ARCH is set).
std::string target_arch();

std::string sha1_hex(const std::string& data);

// Looks up the no-allocation scratch of `device` (noalloc.cpp); skl_init
// calls it with the device current.
int noalloc_setup(int device);

// Logging level from SKELCL_LOG (0 = errors .. 3 = debug), as
// pvsutil::Logger's severities (libraries/pvsutil/include/pvsutil/Logger.h).
int log_level();
void logmsg(int level, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace skl

#define SKL_HIP_TRY(expr)                                                     \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return ::skl::fail(SKL_ERR_HIP, "%s failed: %s (%s:%d)", #expr,         \
                         hipGetErrorString(_e), __FILE__, __LINE__);          \
  } while (0)

#define SKL_TRY(expr)                                                         \
  do {                                                                        \
    int _s = (expr);                                                          \
    if (_s != SKL_OK) return _s;                                              \
  } while (0)

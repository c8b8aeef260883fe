// skelcl_amd C++ API -- drop-in for SkelCL's public headers
// (include/SkelCL/SkelCL.h:137-207, Vector.h, Matrix.h, Map.h, Zip.h,
// Reduce.h, Scan.h, MapOverlap.h, AllPairs.h, Out.h, Local.h, Source.h,
// Index*.h, detail/Types.h, Distributions.h).  Host code written against
// SkelCL compiles unchanged; underneath, every call goes through the C-ABI
// of libskelcl_amd.so (include/skelcl_rt.h, include/skelcl_skeletons.h) to
// hand-written gfx950 kernels.  One GPU: device selection collapses to GPU 0
// and every distribution is the whole container on it.
#pragma once

#include <cstddef>
#include <string>

#include "detail/Runtime.h"
#include "Types.h"
#include "Source.h"
#include "Local.h"
#include "Out.h"
#include "Index.h"
#include "Vector.h"
#include "Matrix.h"
#include "IndexVector.h"
#include "IndexMatrix.h"
#include "Distributions.h"
#include "Map.h"
#include "Zip.h"
#include "Reduce.h"
#include "Scan.h"
#include "MapOverlap.h"
#include "AllPairs.h"

namespace skelcl {

namespace device_type {
enum DeviceType { ANY = 0, CPU, GPU, ACCELERATOR, DEFAULT };
}

// DeviceProperties (detail/DeviceProperties.h): which devices init takes.
class DeviceProperties {
public:
  DeviceProperties& nDevices(size_t n)
  {
    _n = n;
    return *this;
  }
  DeviceProperties& deviceType(int t)
  {
    _type = t;
    return *this;
  }
  size_t count() const { return _n; }
  int type() const { return _type; }

private:
  size_t _n = 0;  // 0 = all
  int _type = device_type::ANY;
};

inline DeviceProperties allDevices() { return DeviceProperties(); }
inline DeviceProperties nDevices(size_t n) { return DeviceProperties().nDevices(n); }

struct PlatformID {
  size_t id;
};
struct DeviceID {
  size_t id;
};
inline PlatformID platform(size_t id) { return PlatformID{id}; }
inline DeviceID device(size_t id) { return DeviceID{id}; }

// skelcl::init (src/SkelCL.cpp:51-54): one GPU, GPU 0.
inline void init(DeviceProperties p = allDevices())
{
  if (p.type() == device_type::CPU || p.type() == device_type::ACCELERATOR)
    detail::abortWithError("skelcl_amd runs on the MI355X GPU only");
  detail::check(skl_init((int)p.count(), SKL_DEVICE_ANY), "init");
}

inline void init(PlatformID p, DeviceID d)
{
  if (p.id != 0 || d.id != 0)
    detail::abortWithError("skelcl_amd exposes only platform 0 / device 0");
  detail::check(skl_init(1, SKL_DEVICE_ANY), "init");
}

inline void terminate() { skl_terminate(); }

// number of devices after init (DeviceList::size: one, the MI355X)
inline size_t deviceCount()
{
  int n = 0;
  skl_device_count(&n);
  return (size_t)n;
}

}  // namespace skelcl

// skelcl_amd C++ API -- drop-in for SkelCL's public headers
// (include/SkelCL/SkelCL.h:137-207, Vector.h, Matrix.h, Map.h, Zip.h,
// Reduce.h, Scan.h, MapOverlap.h, AllPairs.h, Out.h, Local.h, Source.h,
// Index*.h, detail/Types.h, Distributions.h, detail/Device*.h,
// detail/Util.h, detail/Macros.h).  Host code written against SkelCL
// compiles unchanged; underneath, every call goes through the C-ABI of
// libskelcl_amd.so (include/skelcl_rt.h, include/skelcl_skeletons.h) to
// hand-written gfx950 kernels.  One GPU: device selection collapses to GPU 0
// and every distribution is the whole container on it.
#pragma once

#include <cmath>
#include <cstddef>
#include <string>

#include "detail/Runtime.h"
#include "detail/Device.h"
#include "detail/DeviceList.h"
#include "detail/DeviceProperties.h"
#include "detail/Macros.h"
#include "detail/Util.h"
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
#include "Batch.h"

namespace skelcl {

// the public name of the device type (SkelCL.h:137)
typedef detail::Device::Type device_type;

inline detail::DeviceProperties allDevices() { return detail::DeviceProperties::allDevices(); }
inline detail::DeviceProperties nDevices(size_t n) { return detail::DeviceProperties::nDevices(n); }
inline detail::PlatformID platform(size_t pID) { return detail::PlatformID(pID); }
inline detail::DeviceID device(size_t dID) { return detail::DeviceID(dID); }

// skelcl::init (src/SkelCL.cpp:51-64): selects the GPU into globalDeviceList
inline void init(detail::DeviceProperties properties = allDevices())
{
  detail::globalDeviceList.init(properties);
}

inline void init(detail::PlatformID pID, detail::DeviceID dID)
{
  detail::globalDeviceList.init(pID, dID);
}

// skelcl::terminate (src/SkelCL.cpp:66-70): drops the device list; init may
// be called again
inline void terminate()
{
  detail::globalDeviceList.clear();
  skl_terminate();
}

// number of devices after init (globalDeviceList.size(): one, the MI355X)
inline size_t deviceCount() { return detail::globalDeviceList.size(); }

}  // namespace skelcl

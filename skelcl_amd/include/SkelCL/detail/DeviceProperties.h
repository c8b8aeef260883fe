// skelcl::detail::DeviceProperties (detail/DeviceProperties.h,
// src/DeviceProperties.cpp:53-117), DeviceID and PlatformID (DeviceID.h,
// PlatformID.h): which devices skelcl::init takes.  This library drives one
// GPU, so a request matches when it allows a GPU; the count is an upper
// bound that one device always meets.
#pragma once

#include <cstddef>

#include "Device.h"

namespace skelcl {
namespace detail {

class DeviceProperties {
public:
  static DeviceProperties nDevices(size_t n)
  {
    DeviceProperties p;
    p._count = n;
    p._takeAll = false;
    return p;
  }
  static DeviceProperties allDevices() { return DeviceProperties(); }

  DeviceProperties& deviceType(Device::Type value)
  {
    _deviceType = value;
    return *this;
  }
  // also accepts the integer form a command-line parser produces
  DeviceProperties& deviceType(size_t value) { return deviceType(static_cast<Device::Type>(value)); }

  // the device fits the requested type and the count is not yet used up
  bool match(const Device& d) const { return d.isType(_deviceType) && (_takeAll || _count > 0); }
  bool matchAndTake(const Device& d)
  {
    if (!match(d)) return false;
    if (!_takeAll) --_count;
    return true;
  }
  Device::Type type() const { return _deviceType; }
  size_t count() const { return _takeAll ? 0 : _count; }
  bool takeAll() const { return _takeAll; }

private:
  Device::Type _deviceType = Device::ANY;
  bool _takeAll = true;
  size_t _count = 0;
};

class DeviceID {
public:
  DeviceID(size_t id) : _id(id) {}
  size_t id() const { return _id; }

private:
  size_t _id;
};

class PlatformID {
public:
  PlatformID(size_t id) : _id(id) {}
  size_t id() const { return _id; }

private:
  size_t _id;
};

}  // namespace detail
}  // namespace skelcl

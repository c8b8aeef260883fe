// skelcl::detail::DeviceList and globalDeviceList (detail/DeviceList.h:62-122,
// src/DeviceList.cpp:65-173): the devices skelcl::init selected.  Here that
// is the library's one GPU, or nothing when the request excludes GPUs or
// asks for zero devices.
#pragma once

#include <initializer_list>
#include <memory>
#include <vector>

#include "Device.h"
#include "DeviceProperties.h"
#include "Runtime.h"

namespace skelcl {
namespace detail {

class DeviceList {
public:
  typedef std::vector<Device::ptr_type>::value_type value_type;
  typedef std::vector<Device::ptr_type>::size_type size_type;
  typedef std::vector<Device::ptr_type>::iterator iterator;
  typedef std::vector<Device::ptr_type>::const_iterator const_iterator;

  DeviceList() = default;
  DeviceList(std::initializer_list<Device::ptr_type> list) : _devices(list) {}

  bool operator==(const DeviceList& rhs) const { return _devices == rhs._devices; }
  bool operator!=(const DeviceList& rhs) const { return !(*this == rhs); }

  // DeviceList::init (DeviceList.cpp:82-123): takes the GPU if it matches
  void init(DeviceProperties properties = DeviceProperties::allDevices())
  {
    if (!_devices.empty()) return;  // initialised once
    check(skl_init(1, SKL_DEVICE_ANY), "init");
    auto gpu = std::make_shared<Device>(0);
    if (properties.matchAndTake(*gpu)) _devices.push_back(gpu);
    if (_devices.empty()) abortWithError("None OpenCL device was selected.");
  }
  // DeviceList::init(PlatformID, DeviceID) (DeviceList.cpp:125-157)
  void init(PlatformID pID, DeviceID dID)
  {
    if (!_devices.empty()) return;
    if (pID.id() != 0) abortWithError("Given PlatformID is invalid");
    if (dID.id() != 0) abortWithError("Given DeviceID is invalid");
    check(skl_init(1, SKL_DEVICE_ANY), "init");
    _devices.push_back(std::make_shared<Device>(0));
  }
  void clear() { _devices.clear(); }
  // DeviceList::barrier: every device finished its queued work
  void barrier() const
  {
    for (auto& d : _devices) d->wait();
  }

  iterator begin() { return _devices.begin(); }
  iterator end() { return _devices.end(); }
  const_iterator begin() const { return _devices.begin(); }
  const_iterator end() const { return _devices.end(); }
  size_type size() const { return _devices.size(); }
  bool empty() const { return _devices.empty(); }
  value_type& front() { return _devices.front(); }
  const value_type& front() const { return _devices.front(); }
  value_type& back() { return _devices.back(); }
  const value_type& back() const { return _devices.back(); }
  value_type& operator[](size_type i) { return _devices[i]; }
  const value_type& operator[](size_type i) const { return _devices[i]; }
  value_type& at(size_type i) { return _devices.at(i); }
  const value_type& at(size_type i) const { return _devices.at(i); }

private:
  std::vector<Device::ptr_type> _devices;
};

// one list per program for header-only code (a template's static member is
// shared by every translation unit)
template <typename Unused = void>
struct GlobalDeviceListHolder {
  static DeviceList list;
};
template <typename Unused>
DeviceList GlobalDeviceListHolder<Unused>::list;

__attribute__((unused)) static DeviceList& globalDeviceList = GlobalDeviceListHolder<>::list;

}  // namespace detail
}  // namespace skelcl

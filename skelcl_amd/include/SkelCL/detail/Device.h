// skelcl::detail::Device (detail/Device.h:69-366, src/Device.cpp:95-457):
// the device a skeleton runs on, here the one MI355X (GPU 0) behind the
// C-ABI.  Copies keep the reference's offset conventions: enqueueWrite /
// enqueueRead take element offsets and sizes (times the buffer's element
// size), enqueueCopy takes byte offsets (Device.h:142-245, Device.cpp:169-330).
// Every copy is enqueued on the library stream, which the skeletons also
// use, and returns an Event; wait() drains that stream.
#pragma once

#include <cstddef>
#include <istream>
#include <iterator>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>

#include "Event.h"
#include "Runtime.h"

namespace skelcl {
namespace detail {

class DeviceBuffer;

class Device {
public:
  typedef size_t id_type;
  typedef std::shared_ptr<Device> ptr_type;

  // OpenCL's device type bits (CL_DEVICE_TYPE_*), as the reference uses them
  enum Type : size_t {
    ALL = 0xFFFFFFFF,
    ANY = 0xFFFFFFFF,
    CPU = 1 << 1,
    GPU = 1 << 2,
    ACCELERATOR = 1 << 3,
    DEFAULT = 1 << 0
  };

  // the library's GPU; skl_init must have succeeded
  explicit Device(id_type id) : _id(id)
  {
    check(skl_device_get_info(0, &_info), "device info");
  }

  // host -> device, the whole buffer from hostPointer + hostOffset elements
  Event enqueueWrite(const DeviceBuffer& buffer, const void* hostPointer,
                     size_t hostOffset = 0) const;
  // size elements to element deviceOffset, from hostPointer + hostOffset
  Event enqueueWrite(const DeviceBuffer& buffer, void* const hostPointer, size_t size,
                     size_t deviceOffset, size_t hostOffset = 0) const;
  template <typename RandomAccessIterator>
  Event enqueueWrite(const DeviceBuffer& buffer, RandomAccessIterator iterator,
                     size_t hostOffset = 0) const
  {
    return enqueueWrite(buffer, static_cast<const void*>(&*iterator), hostOffset);
  }
  template <typename RandomAccessIterator>
  Event enqueueWrite(const DeviceBuffer& buffer, RandomAccessIterator iterator, size_t size,
                     size_t deviceOffset, size_t hostOffset = 0) const
  {
    return enqueueWrite(buffer, const_cast<void*>(static_cast<const void*>(&*iterator)), size,
                        deviceOffset, hostOffset);
  }

  // device -> host, the whole buffer to hostPointer + hostOffset elements
  Event enqueueRead(const DeviceBuffer& buffer, void* hostPointer, size_t hostOffset = 0) const;
  // size elements from element deviceOffset to hostPointer + hostOffset
  Event enqueueRead(const DeviceBuffer& buffer, void* const hostPointer, size_t size,
                    size_t deviceOffset, size_t hostOffset = 0) const;
  template <typename RandomAccessIterator>
  Event enqueueRead(const DeviceBuffer& buffer, RandomAccessIterator iterator,
                    size_t hostOffset = 0) const
  {
    return enqueueRead(buffer, static_cast<void*>(&*iterator), hostOffset);
  }
  template <typename RandomAccessIterator>
  Event enqueueRead(const DeviceBuffer& buffer, RandomAccessIterator iterator, size_t size,
                    size_t deviceOffset, size_t hostOffset = 0) const
  {
    return enqueueRead(buffer, static_cast<void*>(&*iterator), size, deviceOffset, hostOffset);
  }

  // device -> device: from.sizeInBytes() - fromOffset bytes, offsets in bytes
  Event enqueueCopy(const DeviceBuffer& from, const DeviceBuffer& to, size_t fromOffset = 0,
                    size_t toOffset = 0) const;

  // Device::wait (Device.cpp:332-341): everything enqueued has completed
  void wait() const { check(skl_stream_synchronize(nullptr), "device wait"); }

  id_type id() const { return _id; }
  bool isType(Type t) const { return (static_cast<size_t>(t) & GPU) != 0; }
  std::string typeAsString() const { return "GPU"; }
  std::string name() const { return _info.name; }
  std::string vendorName() const { return "Advanced Micro Devices, Inc."; }
  unsigned int maxClockFrequency() const { return (unsigned)(_info.clock_khz / 1000); }
  unsigned int maxComputeUnits() const { return (unsigned)_info.compute_units; }
  unsigned long maxMemAllocSize() const { return (unsigned long)_info.global_mem; }
  size_t maxWorkGroupSize() const { return (size_t)_info.max_work_group_size; }
  size_t maxWorkGroups() const { return 0xFFFFFFFFu / maxWorkGroupSize(); }
  unsigned long globalMemSize() const { return (unsigned long)_info.global_mem; }
  unsigned long localMemSize() const { return (unsigned long)_info.lds_per_block; }
  bool supportsDouble() const { return _info.supports_double != 0; }
  // the wave (wavefront) width, 64 on CDNA (no OpenCL counterpart)
  int waveSize() const { return _info.wavefront_size; }

private:
  id_type _id;
  skl_device_info _info;
};

// Device::Type from / to text (Device.cpp:427-453)
inline std::istream& operator>>(std::istream& stream, Device::Type& type)
{
  std::string s;
  stream >> s;
  if (s == "ALL") type = Device::ALL;
  else if (s == "ANY") type = Device::ANY;
  else if (s == "CPU") type = Device::CPU;
  else if (s == "GPU") type = Device::GPU;
  else if (s == "ACCELERATOR") type = Device::ACCELERATOR;
  else if (s == "DEFAULT") type = Device::DEFAULT;
  else throw std::invalid_argument("Could not parse (" + s + ") as Device::Type.");
  return stream;
}

inline std::ostream& operator<<(std::ostream& stream, const Device::Type& type)
{
  switch (type) {
    case Device::ANY: return stream << "ANY";
    case Device::CPU: return stream << "CPU";
    case Device::GPU: return stream << "GPU";
    case Device::ACCELERATOR: return stream << "ACCELERATOR";
    case Device::DEFAULT: return stream << "DEFAULT";
  }
  throw std::logic_error("This point should never be reached.");
}

}  // namespace detail
}  // namespace skelcl

#include "DeviceBuffer.h"

// skelcl::detail::DeviceBuffer (detail/DeviceBuffer.h:57-98,
// src/DeviceBuffer.cpp:84-174): `size` elements of `elemSize` bytes in the
// device's memory.  Copying a buffer copies its contents on the device
// (DeviceBuffer.cpp:98-111); moving hands the memory over.  devicePointer()
// takes the place of clBuffer().
#pragma once

#include <cstddef>
#include <memory>
#include <utility>

#include "Device.h"
#include "Runtime.h"

namespace skelcl {
namespace detail {

class DeviceBuffer {
public:
  typedef size_t size_type;

  DeviceBuffer() = default;
  DeviceBuffer(const std::shared_ptr<Device>& devicePtr, size_type size, size_type elemSize,
               unsigned int flags = 0)
      : _device(devicePtr), _size(size), _elemSize(elemSize), _flags(flags)
  {
    _mem.allocate(size * elemSize);
  }
  DeviceBuffer(const DeviceBuffer& rhs)
      : _device(rhs._device), _size(rhs._size), _elemSize(rhs._elemSize), _flags(rhs._flags)
  {
    copyFrom(rhs);
  }
  DeviceBuffer(DeviceBuffer&& rhs) noexcept
      : _device(std::move(rhs._device)), _size(rhs._size), _elemSize(rhs._elemSize),
        _flags(rhs._flags), _mem(std::move(rhs._mem))
  {
    rhs._size = 0;
  }
  DeviceBuffer& operator=(const DeviceBuffer& rhs)
  {
    if (this != &rhs) {
      _device = rhs._device;
      _size = rhs._size;
      _elemSize = rhs._elemSize;
      _flags = rhs._flags;
      copyFrom(rhs);
    }
    return *this;
  }
  DeviceBuffer& operator=(DeviceBuffer&& rhs) noexcept
  {
    _device = std::move(rhs._device);
    _size = rhs._size;
    _elemSize = rhs._elemSize;
    _flags = rhs._flags;
    _mem = std::move(rhs._mem);
    rhs._size = 0;
    return *this;
  }

  std::shared_ptr<Device> devicePtr() const { return _device; }
  size_type size() const { return _size; }
  size_type elemSize() const { return _elemSize; }
  size_type sizeInBytes() const { return _size * _elemSize; }
  unsigned int flags() const { return _flags; }
  void* devicePointer() const { return _mem.get(); }
  bool isValid() const { return _mem.get() != nullptr; }

private:
  void copyFrom(const DeviceBuffer& rhs)
  {
    _mem.allocate(rhs.sizeInBytes());
    if (rhs.sizeInBytes())
      check(skl_memcpy_d2d(_mem.get(), 0, rhs._mem.get(), 0, rhs.sizeInBytes(), nullptr),
            "device buffer copy");
  }

  std::shared_ptr<Device> _device;
  size_type _size = 0;
  size_type _elemSize = 0;
  unsigned int _flags = 0;
  DeviceMemory _mem;
};

// ---- Device's copies (they need the buffer's layout) ----------------------
inline Event Device::enqueueWrite(const DeviceBuffer& buffer, const void* hostPointer,
                                  size_t hostOffset) const
{
  const char* src = static_cast<const char*>(hostPointer) + hostOffset * buffer.elemSize();
  check(skl_memcpy_h2d(buffer.devicePointer(), 0, src, buffer.sizeInBytes(), nullptr, 0),
        "enqueueWrite");
  return Event::recordNow();
}

inline Event Device::enqueueWrite(const DeviceBuffer& buffer, void* const hostPointer,
                                  size_t size, size_t deviceOffset, size_t hostOffset) const
{
  const char* src = static_cast<const char*>(hostPointer) + hostOffset * buffer.elemSize();
  check(skl_memcpy_h2d(buffer.devicePointer(), deviceOffset * buffer.elemSize(), src,
                       size * buffer.elemSize(), nullptr, 0),
        "enqueueWrite");
  return Event::recordNow();
}

inline Event Device::enqueueRead(const DeviceBuffer& buffer, void* hostPointer,
                                 size_t hostOffset) const
{
  char* dst = static_cast<char*>(hostPointer) + hostOffset * buffer.elemSize();
  check(skl_memcpy_d2h(dst, buffer.devicePointer(), 0, buffer.sizeInBytes(), nullptr, 0),
        "enqueueRead");
  return Event::recordNow();
}

inline Event Device::enqueueRead(const DeviceBuffer& buffer, void* const hostPointer,
                                 size_t size, size_t deviceOffset, size_t hostOffset) const
{
  char* dst = static_cast<char*>(hostPointer) + hostOffset * buffer.elemSize();
  check(skl_memcpy_d2h(dst, buffer.devicePointer(), deviceOffset * buffer.elemSize(),
                       size * buffer.elemSize(), nullptr, 0),
        "enqueueRead");
  return Event::recordNow();
}

inline Event Device::enqueueCopy(const DeviceBuffer& from, const DeviceBuffer& to,
                                 size_t fromOffset, size_t toOffset) const
{
  if (fromOffset > from.sizeInBytes() || toOffset > to.sizeInBytes() ||
      from.sizeInBytes() - fromOffset > to.sizeInBytes() - toOffset)
    abortWithError("enqueueCopy: the source range does not fit the destination");
  check(skl_memcpy_d2d(to.devicePointer(), toOffset, from.devicePointer(), fromOffset,
                       from.sizeInBytes() - fromOffset, nullptr),
        "enqueueCopy");
  return Event::recordNow();
}

}  // namespace detail
}  // namespace skelcl

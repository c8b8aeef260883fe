// skelcl::Vector<T> (Vector.h, detail/VectorDef.h:74-695): a host
// std::vector plus one device buffer with the reference's lazy coherence:
// skeletons upload when the host copy is newer (startUpload,
// VectorDef.h:550-578), host access downloads when the device copy is newer
// (copyDataToHost, VectorDef.h:605-610), outputs are marked device-modified.
// As in the reference, writing through operator[] does not mark the device
// copy stale; call dataOnHostModified() after host writes to data that a
// skeleton already used.
#pragma once

#include <algorithm>
#include <initializer_list>
#include <vector>

#include "detail/Runtime.h"

namespace skelcl {

namespace detail {

template <typename T>
class HostDeviceData {
public:
  typedef std::vector<T> host_buffer_type;

  HostDeviceData() = default;
  explicit HostDeviceData(size_t n, const T& v = T()) : _host(n, v) {}
  explicit HostDeviceData(host_buffer_type&& h) : _host(std::move(h)) {}
  HostDeviceData(const HostDeviceData& o) : _host()
  {
    o.copyDataToHost();
    _host = o._host;
  }
  HostDeviceData& operator=(const HostDeviceData& o)
  {
    if (this != &o) {
      o.copyDataToHost();
      _host = o._host;
      _hostValid = true;
      _devValid = false;
    }
    return *this;
  }
  HostDeviceData(HostDeviceData&&) = default;
  HostDeviceData& operator=(HostDeviceData&&) = default;

  // ---- coherence (VectorDef.h:550-626)
  void copyDataToHost() const
  {
    if (!_hostValid) {
      _host.resize(elements());
      if (!_host.empty())
        check(skl_memcpy_d2h(_host.data(), _dev.get(), 0, _host.size() * sizeof(T), nullptr, 1),
              "download");
      _hostValid = true;
    }
  }
  void copyDataToDevice() const
  {
    ensureDevice();
    if (!_devValid) {
      if (!_host.empty())
        check(skl_memcpy_h2d(_dev.get(), 0, _host.data(), _host.size() * sizeof(T), nullptr, 1),
              "upload");
      _devValid = true;
    }
  }
  void dataOnDeviceModified() const
  {
    _hostValid = false;
    _devValid = true;
  }
  void dataOnHostModified() const
  {
    _hostValid = true;
    _devValid = false;
  }
  bool hostIsUpToDate() const { return _hostValid; }
  bool deviceIsUpToDate() const { return _devValid; }

  // device pointer with current data (uploads if needed)
  void* devicePointer() const
  {
    copyDataToDevice();
    return _dev.get();
  }
  // device storage a skeleton overwrites entirely (no upload)
  void* deviceOutputPointer() const
  {
    ensureDevice();
    return _dev.get();
  }
  // size an output to n elements; the host copy is refilled on download
  void prepareOutput(size_t n)
  {
    if (_elements != n) {
      host_buffer_type().swap(_host);
      _elements = n;
      _dev.release();
      _devValid = false;
      _hostValid = false;
    }
    ensureDevice();
  }

protected:
  size_t elements() const { return _elements; }
  void setElements(size_t n) { _elements = n; }
  void ensureDevice() const
  {
    size_t need = std::max<size_t>(1, _elements) * sizeof(T);
    if (_dev.bytes() < need) {
      if (_devValid && !_hostValid) copyDataToHost();
      _dev.allocate(need);
      _devValid = false;
    }
  }
  void dropDevice()
  {
    _dev.release();
    _devValid = false;
  }

  mutable host_buffer_type _host;
  mutable DeviceMemory _dev;
  mutable bool _hostValid = true;
  mutable bool _devValid = false;
  size_t _elements = 0;
};

}  // namespace detail

template <typename T>
class Vector : public detail::HostDeviceData<T> {
  typedef detail::HostDeviceData<T> base;

public:
  typedef std::vector<T> host_buffer_type;
  typedef T value_type;
  typedef size_t size_type;
  typedef T& reference;
  typedef const T& const_reference;
  typedef typename host_buffer_type::iterator iterator;
  typedef typename host_buffer_type::const_iterator const_iterator;

  Vector() { this->setElements(0); }
  explicit Vector(size_type n, const T& v = T()) : base(n, v) { this->setElements(n); }
  template <class It>
  Vector(It first, It last) : base(host_buffer_type(first, last))
  {
    this->setElements(this->_host.size());
  }
  Vector(std::initializer_list<T> l) : base(host_buffer_type(l))
  {
    this->setElements(this->_host.size());
  }
  Vector(const Vector& o) : base(o) { this->setElements(o.size()); }
  Vector& operator=(const Vector& o)
  {
    base::operator=(o);
    this->setElements(o.size());
    return *this;
  }
  Vector(Vector&&) = default;
  Vector& operator=(Vector&&) = default;

  size_type size() const { return this->elements(); }
  bool empty() const { return size() == 0; }

  // host access downloads first (VectorDef.h:206-232, 317-373)
  reference operator[](size_type i)
  {
    this->copyDataToHost();
    return this->_host[i];
  }
  const_reference operator[](size_type i) const
  {
    this->copyDataToHost();
    return this->_host[i];
  }
  reference at(size_type i)
  {
    this->copyDataToHost();
    return this->_host.at(i);
  }
  reference front() { return (*this)[0]; }
  const_reference front() const { return (*this)[0]; }
  reference back() { return (*this)[size() - 1]; }
  iterator begin()
  {
    this->copyDataToHost();
    return this->_host.begin();
  }
  iterator end()
  {
    this->copyDataToHost();
    return this->_host.end();
  }
  const_iterator begin() const
  {
    this->copyDataToHost();
    return this->_host.begin();
  }
  const_iterator end() const
  {
    this->copyDataToHost();
    return this->_host.end();
  }
  T* data()
  {
    this->copyDataToHost();
    return this->_host.data();
  }

  // resize drops the device copy (VectorDef.h:286-297)
  void resize(size_type n, T v = T())
  {
    this->copyDataToHost();
    this->_host.resize(n, v);
    this->setElements(n);
    this->dropDevice();
    this->dataOnHostModified();
  }
  void clear() { resize(0); }
  void push_back(const T& v)
  {
    this->copyDataToHost();
    this->_host.push_back(v);
    this->setElements(this->_host.size());
    this->dropDevice();
    this->dataOnHostModified();
  }
  host_buffer_type& hostBuffer() const
  {
    this->copyDataToHost();
    return this->_host;
  }
  // distribution hooks of the reference (one GPU: the whole vector)
  void createDeviceBuffers() const { this->ensureDevice(); }
  void startUpload() const { this->copyDataToDevice(); }
};

}  // namespace skelcl

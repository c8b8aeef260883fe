// skelcl::Vector<T> (Vector.h, detail/VectorDef.h:74-695): a host
// std::vector plus one device buffer with the reference's lazy coherence:
// skeletons upload when the host copy is newer (startUpload,
// VectorDef.h:550-578), host access downloads when the device copy is newer
// (copyDataToHost, VectorDef.h:605-610), outputs are marked device-modified.
// As in the reference, writing through operator[] does not mark the device
// copy stale; call dataOnHostModified() after host writes to data that a
// skeleton already used.  Each Vector records a distribution (not set until
// a skeleton or the caller sets one); on one GPU every distribution holds
// the whole vector.
#pragma once

#include <algorithm>
#include <initializer_list>
#include <limits>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "detail/DeviceBuffer.h"
#include "detail/DeviceList.h"
#include "detail/Event.h"
#include "detail/Runtime.h"

namespace skelcl {

template <typename T> class Vector;
template <typename T> class Matrix;

namespace detail {

// Host copy, device copy and the two up-to-date flags shared by Vector and
// Matrix (Vector.h:733-738).
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
        check(skl_memcpy_d2h(_host.data(), _dev.devicePointer(), 0, _host.size() * sizeof(T),
                             nullptr, 1),
              "download");
      _hostValid = true;
    }
  }
  void copyDataToDevices() const
  {
    ensureDevice();
    if (!_devValid) {
      if (!_host.empty())
        check(skl_memcpy_h2d(_dev.devicePointer(), 0, _host.data(), _host.size() * sizeof(T),
                             nullptr, 1),
              "upload");
      _devValid = true;
    }
  }
  void copyDataToDevice() const { copyDataToDevices(); }
  // startUpload / startDownload (VectorDef.h:550-604): the copy has
  // completed when the returned event has
  Event startUpload() const
  {
    copyDataToDevices();
    return Event::recordNow();
  }
  Event startDownload() const
  {
    copyDataToHost();
    return Event::recordNow();
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
  bool devicesAreUpToDate() const { return _devValid; }

  // createDeviceBuffers / forceCreateDeviceBuffers (VectorDef.h:515-548)
  void createDeviceBuffers() const { ensureDevice(); }
  void forceCreateDeviceBuffers() const
  {
    if (_devValid && !_hostValid) copyDataToHost();
    _dev = DeviceBuffer(firstDevice(), std::max<size_t>(1, elements()), sizeof(T));
    _devValid = false;
  }
  // the device buffer (one GPU: the same one for any device)
  const DeviceBuffer& deviceBuffer(const Device& device) const
  {
    (void)device;
    ensureDevice();
    return _dev;
  }
  DeviceBuffer& deviceBuffer(const Device& device)
  {
    (void)device;
    ensureDevice();
    return _dev;
  }

  // device pointer with current data (uploads if needed)
  void* devicePointer() const
  {
    copyDataToDevices();
    return _dev.devicePointer();
  }
  // device storage a skeleton overwrites entirely (no upload)
  void* deviceOutputPointer() const
  {
    ensureDevice();
    return _dev.devicePointer();
  }
  // size an output to n elements; the host copy is refilled on download
  void prepareOutput(size_t n)
  {
    if (_elements != n) {
      host_buffer_type().swap(_host);
      _elements = n;
      _dev = DeviceBuffer();
      _devValid = false;
      _hostValid = false;
    }
    ensureDevice();
  }

protected:
  size_t elements() const { return _elements; }
  void setElements(size_t n) { _elements = n; }
  static std::shared_ptr<Device> firstDevice()
  {
    return globalDeviceList.empty() ? std::shared_ptr<Device>() : globalDeviceList.front();
  }
  void ensureDevice() const
  {
    const size_t need = std::max<size_t>(1, _elements);
    if (!_dev.isValid() || _dev.size() < need) {
      if (_devValid && !_hostValid) copyDataToHost();
      _dev = DeviceBuffer(firstDevice(), need, sizeof(T));
      _devValid = false;
    }
  }
  void dropDevice()
  {
    _dev = DeviceBuffer();
    _devValid = false;
  }

  mutable host_buffer_type _host;
  mutable DeviceBuffer _dev;
  mutable bool _hostValid = true;
  mutable bool _devValid = false;
  size_t _elements = 0;
};

}  // namespace detail
}  // namespace skelcl

#include "detail/Distribution.h"

namespace skelcl {

template <typename T>
class Vector : public detail::HostDeviceData<T> {
  typedef detail::HostDeviceData<T> base;
  typedef detail::Distribution<Vector<T>> dist_type;

public:
  typedef std::vector<T> host_buffer_type;
  typedef typename host_buffer_type::value_type value_type;
  typedef typename host_buffer_type::pointer pointer;
  typedef typename host_buffer_type::const_pointer const_pointer;
  typedef typename host_buffer_type::reference reference;
  typedef typename host_buffer_type::const_reference const_reference;
  typedef typename host_buffer_type::iterator iterator;
  typedef typename host_buffer_type::const_iterator const_iterator;
  typedef typename host_buffer_type::reverse_iterator reverse_iterator;
  typedef typename host_buffer_type::const_reverse_iterator const_reverse_iterator;
  typedef typename host_buffer_type::size_type size_type;
  typedef typename host_buffer_type::difference_type difference_type;
  typedef typename host_buffer_type::allocator_type allocator_type;

  Vector() : _dist(new dist_type()) { this->setElements(0); }
  explicit Vector(size_type n, const T& v = T(), const dist_type& d = dist_type())
      : base(n, v), _dist(detail::cloneAndConvert<Vector<T>>(d))
  {
    this->setElements(n);
  }
  template <class It, class = typename std::iterator_traits<It>::iterator_category>
  Vector(It first, It last) : base(host_buffer_type(first, last)), _dist(new dist_type())
  {
    this->setElements(this->_host.size());
  }
  template <class It, class = typename std::iterator_traits<It>::iterator_category>
  Vector(It first, It last, const dist_type& d)
      : base(host_buffer_type(first, last)), _dist(detail::cloneAndConvert<Vector<T>>(d))
  {
    this->setElements(this->_host.size());
  }
  Vector(std::initializer_list<T> l) : base(host_buffer_type(l)), _dist(new dist_type())
  {
    this->setElements(this->_host.size());
  }
  Vector(const Vector& o) : base(o), _dist(detail::cloneAndConvert<Vector<T>>(*o._dist))
  {
    this->setElements(o.size());
  }
  Vector& operator=(const Vector& o)
  {
    if (this != &o) {
      base::operator=(o);
      _dist = detail::cloneAndConvert<Vector<T>>(*o._dist);
      this->setElements(o.size());
    }
    return *this;
  }
  Vector(Vector&&) = default;
  Vector& operator=(Vector&&) = default;

  // ---- capacity
  size_type size() const { return this->elements(); }
  bool empty() const { return size() == 0; }
  size_type max_size() const { return this->_host.max_size(); }
  size_type capacity() const { return this->_host.capacity(); }
  void reserve(size_type n)
  {
    this->copyDataToHost();
    this->_host.reserve(n);
  }

  // ---- element access: host access downloads first (VectorDef.h:206-232,
  // 317-373)
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
  const_reference at(size_type i) const
  {
    this->copyDataToHost();
    return this->_host.at(i);
  }
  reference front() { return (*this)[0]; }
  const_reference front() const { return (*this)[0]; }
  reference back() { return (*this)[size() - 1]; }
  const_reference back() const { return (*this)[size() - 1]; }
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

  // ---- modifiers: change the host copy and drop the device copy
  // (VectorDef.h:286-297 for resize)
  void resize(size_type n, T v = T())
  {
    modify([&](host_buffer_type& h) { h.resize(n, v); });
  }
  template <class It>
  void assign(It first, It last)
  {
    modify([&](host_buffer_type& h) { h.assign(first, last); });
  }
  void assign(size_type count, const T& value = T())
  {
    modify([&](host_buffer_type& h) { h.assign(count, value); });
  }
  void push_back(const T& v)
  {
    modify([&](host_buffer_type& h) { h.push_back(v); });
  }
  void pop_back()
  {
    modify([&](host_buffer_type& h) { h.pop_back(); });
  }
  iterator insert(iterator pos, const T& value)
  {
    iterator r;
    modify([&](host_buffer_type& h) { r = h.insert(pos, value); });
    return r;
  }
  iterator insert(iterator pos, size_type count, const T& value)
  {
    iterator r;
    modify([&](host_buffer_type& h) { r = h.insert(pos, count, value); });
    return r;
  }
  template <class It, class = typename std::iterator_traits<It>::iterator_category>
  void insert(iterator pos, It first, It last)
  {
    modify([&](host_buffer_type& h) { h.insert(pos, first, last); });
  }
  iterator erase(iterator pos)
  {
    iterator r;
    modify([&](host_buffer_type& h) { r = h.erase(pos); });
    return r;
  }
  iterator erase(iterator first, iterator last)
  {
    iterator r;
    modify([&](host_buffer_type& h) { r = h.erase(first, last); });
    return r;
  }
  void swap(Vector& rhs)
  {
    std::swap(static_cast<base&>(*this), static_cast<base&>(rhs));
    std::swap(_dist, rhs._dist);
  }
  void clear() { resize(0); }

  host_buffer_type& hostBuffer() const
  {
    this->copyDataToHost();
    return this->_host;
  }

  // ---- distribution (VectorDef.h:492-513)
  dist_type& distribution() const { return *_dist; }
  template <typename U>
  void setDistribution(const detail::Distribution<Vector<U>>& d) const
  {
    _dist = detail::cloneAndConvert<Vector<T>>(d);
  }
  template <typename U>
  void setDistribution(const std::unique_ptr<detail::Distribution<Vector<U>>>& d) const
  {
    setDistribution(*d);
  }
  void setDistribution(std::unique_ptr<dist_type>&& d) const
  {
    if (d) _dist = std::move(d);
  }

  // device-side helper text for this element type: none for Vector
  static std::string deviceFunctions() { return std::string(); }

private:
  template <typename F>
  void modify(F f)
  {
    this->copyDataToHost();
    f(this->_host);
    this->setElements(this->_host.size());
    this->dropDevice();
    this->dataOnHostModified();
  }

  mutable std::unique_ptr<dist_type> _dist;
};

}  // namespace skelcl

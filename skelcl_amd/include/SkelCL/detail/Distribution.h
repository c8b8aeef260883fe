// skelcl::detail::Distribution and its kinds (detail/Distribution.h,
// SingleDistribution.h, BlockDistribution.h, CopyDistribution.h,
// OLDistribution.h): how a container is laid out over the selected devices.
// With one GPU every valid distribution places the whole container on it,
// so the kinds differ only in what they record -- their kind, device list,
// combine function (Copy) and overlap settings (OL) -- which containers keep
// and skeletons set as the reference's prepareInput rules do
// (MapHelperDef.h:65-95, ZipDef.h:226-271, ReduceDef.h:129-157,
// ScanDef.h:299-324, MapOverlapDef.h:250-275, AllPairsDef.h:346-410).
// A default-constructed Distribution is "not set" (isValid() == false).
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <type_traits>
#include <typeinfo>

#include "DeviceList.h"

namespace skelcl {
namespace detail {

enum class Padding { NEUTRAL, NEAREST };  // Padding.h:48-51

template <typename> class Distribution;

// the neutral element carried over to another element type when it
// converts, else a value-initialised one
template <typename T, typename U>
typename std::enable_if<std::is_convertible<U, T>::value, T>::type convertNeutral(const U& u)
{
  return static_cast<T>(u);
}
template <typename T, typename U>
typename std::enable_if<!std::is_convertible<U, T>::value, T>::type convertNeutral(const U&)
{
  return T();
}

template <template <typename> class C, typename T>
class Distribution<C<T>> {
public:
  Distribution() = default;
  template <typename U>
  Distribution(const Distribution<C<U>>& rhs) : _devices(rhs.devices())
  {
  }
  virtual ~Distribution() = default;

  // not set: no devices
  virtual bool isValid() const { return false; }
  virtual bool isSingle() const { return false; }
  virtual bool isBlock() const { return false; }
  virtual bool isCopy() const { return false; }
  virtual bool isOverlap() const { return false; }
  virtual std::string kindName() const { return "none"; }

  bool operator==(const Distribution& rhs) const
  {
    return typeid(*this) == typeid(rhs) && _devices == rhs._devices && doCompare(rhs);
  }
  bool operator!=(const Distribution& rhs) const { return !(*this == rhs); }

  const DeviceList& devices() const { return _devices; }
  const DeviceList::value_type device(size_t i) const { return _devices[i]; }

  // elements of `c` held by `devicePtr`: all of them on the one GPU
  virtual size_t sizeForDevice(const C<T>& c, const std::shared_ptr<Device>& devicePtr) const
  {
    for (auto& d : _devices)
      if (d == devicePtr) return elementCount(c);
    return 0;
  }
  // whether changing to `other` moves data between devices: never here
  virtual bool dataExchangeOnDistributionChange(Distribution& other)
  {
    (void)other;
    return false;
  }

protected:
  explicit Distribution(const DeviceList& devices) : _devices(devices) {}
  virtual bool doCompare(const Distribution& rhs) const
  {
    (void)rhs;
    return true;
  }
  template <typename X>
  static size_t elementCount(const Vector<X>& v)
  {
    return v.size();
  }
  template <typename X>
  static size_t elementCount(const Matrix<X>& m)
  {
    return m.size().elemCount();
  }

  DeviceList _devices;
};

template <typename> class SingleDistribution;
template <template <typename> class C, typename T>
class SingleDistribution<C<T>> : public Distribution<C<T>> {
public:
  // the first selected device (SingleDistributionDef.h:79-118)
  SingleDistribution()
      : Distribution<C<T>>(globalDeviceList.empty() ? DeviceList()
                                                   : DeviceList{globalDeviceList.front()})
  {
  }
  explicit SingleDistribution(const std::shared_ptr<Device>& device)
      : Distribution<C<T>>(DeviceList{device})
  {
  }
  template <typename U>
  SingleDistribution(const SingleDistribution<C<U>>& rhs) : Distribution<C<T>>(rhs.devices())
  {
  }
  bool isValid() const override { return true; }
  bool isSingle() const override { return true; }
  std::string kindName() const override { return "single"; }
};

template <typename> class BlockDistribution;
template <template <typename> class C, typename T>
class BlockDistribution<C<T>> : public Distribution<C<T>> {
public:
  // one block per selected device (BlockDistributionDef.h:80-128): the whole
  // container on the one GPU
  explicit BlockDistribution(const DeviceList& devices = globalDeviceList)
      : Distribution<C<T>>(devices)
  {
  }
  template <typename U>
  BlockDistribution(const BlockDistribution<C<U>>& rhs) : Distribution<C<T>>(rhs.devices())
  {
  }
  bool isValid() const override { return true; }
  bool isBlock() const override { return true; }
  std::string kindName() const override { return "block"; }
};

template <typename> class CopyDistribution;
template <template <typename> class C, typename T>
class CopyDistribution<C<T>> : public Distribution<C<T>> {
public:
  typedef std::function<T(const T&, const T&)> combine_type;
  // a full copy per device, merged on download by combineFunc when set
  // (CopyDistributionDef.h:80-145); one GPU holds the only copy
  explicit CopyDistribution(const DeviceList& devices = globalDeviceList,
                            combine_type combineFunc = nullptr)
      : Distribution<C<T>>(devices), _combine(combineFunc)
  {
  }
  template <typename U>
  CopyDistribution(const CopyDistribution<C<U>>& rhs) : Distribution<C<T>>(rhs.devices())
  {
  }
  bool isValid() const override { return true; }
  bool isCopy() const override { return true; }
  std::string kindName() const override { return "copy"; }
  const combine_type& combineFunc() const { return _combine; }

private:
  combine_type _combine;
};

template <typename> class OLDistribution;
template <template <typename> class C, typename T>
class OLDistribution<C<T>> : public Distribution<C<T>> {
public:
  // row blocks with `overlapRadius` halo rows (OLDistributionDef.h:156-393);
  // on one GPU the block is the whole matrix and MapOverlap applies the
  // padding while staging tiles, so no halo rows are materialised
  explicit OLDistribution(unsigned int overlapRadius = 1, Padding padding = Padding::NEAREST,
                          T neutral = T(), const DeviceList& devices = globalDeviceList)
      : Distribution<C<T>>(devices), _radius(overlapRadius), _padding(padding),
        _neutral(neutral)
  {
  }
  template <typename U>
  OLDistribution(const OLDistribution<C<U>>& rhs)
      : Distribution<C<T>>(rhs.devices()), _radius(rhs.overlapRadius()),
        _padding(rhs.padding()), _neutral(convertNeutral<T>(rhs.neutralElement()))
  {
  }
  bool isValid() const override { return true; }
  bool isOverlap() const override { return true; }
  std::string kindName() const override { return "overlap"; }
  unsigned int overlapRadius() const { return _radius; }
  Padding padding() const { return _padding; }
  T neutralElement() const { return _neutral; }

protected:
  bool doCompare(const Distribution<C<T>>& rhs) const override
  {
    auto& o = static_cast<const OLDistribution&>(rhs);
    return _radius == o._radius && _padding == o._padding;
  }

private:
  unsigned int _radius;
  Padding _padding;
  T _neutral;
};

// a copy of `dist` of the same kind for another element type
// (Distributions.h:241-276)
template <typename OutT, typename InT>
std::unique_ptr<Distribution<OutT>> cloneAndConvert(const Distribution<InT>& dist)
{
  typedef std::unique_ptr<Distribution<OutT>> P;
  if (auto s = dynamic_cast<const SingleDistribution<InT>*>(&dist))
    return P(new SingleDistribution<OutT>(*s));
  if (auto b = dynamic_cast<const BlockDistribution<InT>*>(&dist))
    return P(new BlockDistribution<OutT>(*b));
  if (auto c = dynamic_cast<const CopyDistribution<InT>*>(&dist))
    return P(new CopyDistribution<OutT>(*c));
  if (auto o = dynamic_cast<const OLDistribution<InT>*>(&dist))
    return P(new OLDistribution<OutT>(*o));
  return P(new Distribution<OutT>(dist));
}

}  // namespace detail
}  // namespace skelcl

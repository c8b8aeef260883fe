// skelcl::distribution (Distributions.h:64-311): builders and setters of the
// distribution a container records (detail/Distribution.h).  The template
// form C<T> accepts Vector and Matrix alike; the element type is deduced
// from the container.
#pragma once

#include <functional>
#include <memory>

#include "Matrix.h"
#include "Vector.h"
#include "detail/Distribution.h"

namespace skelcl {
namespace distribution {

// blocks the deduction of T from a function argument
template <typename T>
struct identity {
  typedef T type;
};

template <template <typename> class C, typename T>
std::unique_ptr<detail::Distribution<C<T>>> Block(const C<T>&)
{
  return std::unique_ptr<detail::Distribution<C<T>>>(new detail::BlockDistribution<C<T>>());
}
template <template <typename> class C, typename T>
void setBlock(const C<T>& c)
{
  c.setDistribution(Block(c));
}

template <template <typename> class C, typename T>
std::unique_ptr<detail::Distribution<C<T>>> OL(const C<T>&)
{
  return std::unique_ptr<detail::Distribution<C<T>>>(new detail::OLDistribution<C<T>>());
}
template <template <typename> class C, typename T>
void setOL(const C<T>& c)
{
  c.setDistribution(OL(c));
}

template <template <typename> class C, typename T>
std::unique_ptr<detail::Distribution<C<T>>> Copy(
    const C<T>&,
    typename identity<std::function<T(const T&, const T&)>>::type combineFunc = nullptr)
{
  return std::unique_ptr<detail::Distribution<C<T>>>(
      new detail::CopyDistribution<C<T>>(detail::globalDeviceList, combineFunc));
}
template <template <typename> class C, typename T>
void setCopy(const C<T>& c, std::function<T(const T&, const T&)> combineFunc = nullptr)
{
  c.setDistribution(Copy(c, combineFunc));
}

template <template <typename> class C, typename T>
std::unique_ptr<detail::Distribution<C<T>>> Single(const C<T>&)
{
  return std::unique_ptr<detail::Distribution<C<T>>>(new detail::SingleDistribution<C<T>>());
}
template <template <typename> class C, typename T>
std::unique_ptr<detail::Distribution<C<T>>> Single(const C<T>&,
                                                   const std::shared_ptr<detail::Device>& device)
{
  return std::unique_ptr<detail::Distribution<C<T>>>(
      new detail::SingleDistribution<C<T>>(device));
}
template <template <typename> class C, typename T>
void setSingle(const C<T>& c)
{
  c.setDistribution(Single(c));
}
template <template <typename> class C, typename T>
void setSingle(const C<T>& c, const std::shared_ptr<detail::Device>& device)
{
  c.setDistribution(Single(c, device));
}

}  // namespace distribution
}  // namespace skelcl

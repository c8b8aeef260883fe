// Common part of the skeleton classes (detail/Skeleton.h:99-142,
// src/Skeleton.cpp:48,75-78) and the extra-argument marshalling of
// kernelUtil::setKernelArgs (detail/KernelUtil.h:134-245): a Vector becomes
// a device pointer, a Matrix a device pointer plus its column count, Local
// a dynamic __local buffer, out(c) a written container, anything else is
// passed by value.
#pragma once

#include <cstring>
#include <functional>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "../Local.h"
#include "../Matrix.h"
#include "../Out.h"
#include "../Source.h"
#include "../Vector.h"
#include "Distribution.h"
#include "Runtime.h"

namespace skelcl {
namespace detail {

class ArgList {
public:
  std::vector<skl_arg> args;
  std::vector<std::vector<char>> values;
  std::vector<std::function<void()>> after;

  const skl_arg* data() const { return args.empty() ? nullptr : args.data(); }
  int count() const { return (int)args.size(); }
  void finish() const
  {
    for (auto& f : after) f();
  }
  void fixValuePointers()
  {
    size_t v = 0;
    for (auto& a : args)
      if (a.kind == SKL_ARG_VALUE) a.value = values[v++].data();
  }
};

template <typename T>
inline void addArg(ArgList& L, const Vector<T>& v)
{
  L.args.push_back(skl_arg{SKL_ARG_BUFFER, v.devicePointer(), 0, 0});
}
template <typename T>
inline void addArg(ArgList& L, Vector<T>& v)
{
  addArg(L, static_cast<const Vector<T>&>(v));
}
template <typename T>
inline void addArg(ArgList& L, Out<Vector<T>> o)
{
  Vector<T>* c = &o.container();
  L.args.push_back(skl_arg{SKL_ARG_BUFFER, c->devicePointer(), 0, 0});
  L.after.push_back([c] { c->dataOnDeviceModified(); });
}
template <typename T>
inline void addArg(ArgList& L, const Matrix<T>& m)
{
  L.args.push_back(skl_arg{SKL_ARG_MATRIX, m.devicePointer(), 0, (unsigned)m.columnCount()});
}
template <typename T>
inline void addArg(ArgList& L, Matrix<T>& m)
{
  addArg(L, static_cast<const Matrix<T>&>(m));
}
template <typename T>
inline void addArg(ArgList& L, Out<Matrix<T>> o)
{
  Matrix<T>* c = &o.container();
  L.args.push_back(skl_arg{SKL_ARG_MATRIX, c->devicePointer(), 0, (unsigned)c->columnCount()});
  L.after.push_back([c] { c->dataOnDeviceModified(); });
}
inline void addArg(ArgList& L, const Local& l)
{
  L.args.push_back(skl_arg{SKL_ARG_LOCAL, nullptr, l.getSizeInBytes(), 0});
}
template <typename T>
inline typename std::enable_if<std::is_trivially_copyable<typename std::decay<T>::type>::value>::type
addArg(ArgList& L, const T& v)
{
  const char* p = reinterpret_cast<const char*>(&v);
  L.values.emplace_back(p, p + sizeof(T));
  L.args.push_back(skl_arg{SKL_ARG_VALUE, nullptr, sizeof(T), 0});
}

inline void marshal(ArgList&) {}
template <typename A, typename... Rest>
inline void marshal(ArgList& L, A&& a, Rest&&... rest)
{
  addArg(L, std::forward<A>(a));
  marshal(L, std::forward<Rest>(rest)...);
}

template <typename... Args>
inline ArgList makeArgs(Args&&... args)
{
  ArgList L;
  marshal(L, std::forward<Args>(args)...);
  L.fixValuePointers();
  return L;
}

template <typename T>
inline std::string valueString(const T& v)
{
  std::ostringstream o;
  o << std::setprecision(std::numeric_limits<double>::max_digits10) << v;
  return o.str();
}

// ---- the reference's distribution rules for skeleton inputs and outputs.
// On one GPU they change no data placement; containers record them so that
// host code reading distribution() sees what the reference would set.

// an input without a distribution gets kind D (MapHelperDef.h:70-71,
// ReduceDef.h:135-136, ScanDef.h:303-304)
template <template <typename> class D, typename C>
inline void defaultDistribution(const C& c)
{
  if (!c.distribution().isValid()) c.setDistribution(D<C>());
}

// the output takes the input's distribution (MapHelperDef.h:92,
// ZipDef.h:268, ReduceDef.h:157, ScanDef.h:324, MapOverlapDef.h:275)
template <typename O, typename I>
inline void adoptDistribution(const O& o, const I& in)
{
  o.setDistribution(in.distribution());
}

// Zip (ZipDef.h:232-243): both unset or unequal -> Block; one unset -> the
// other's
template <typename L, typename R>
inline void zipDistributions(const L& l, const R& r)
{
  const bool lv = l.distribution().isValid(), rv = r.distribution().isValid();
  if (!lv && !rv) {
    l.setDistribution(BlockDistribution<L>());
    r.setDistribution(BlockDistribution<R>());
  } else if (!lv) {
    l.setDistribution(r.distribution());
  } else if (!rv) {
    r.setDistribution(l.distribution());
  } else if (*cloneAndConvert<L>(r.distribution()) != l.distribution()) {
    l.setDistribution(BlockDistribution<L>());
    r.setDistribution(BlockDistribution<R>());
  }
}

// AllPairs (AllPairsDef.h:346-410): by default left Block, right Copy; a
// Single side makes the other Single, a Block right becomes Copy
template <typename L, typename R>
inline void allPairsDistributions(const L& l, const R& r)
{
  const auto& ld = l.distribution();
  const auto& rd = r.distribution();
  if (!ld.isValid() && !rd.isValid()) {
    l.setDistribution(BlockDistribution<L>());
    r.setDistribution(CopyDistribution<R>());
  } else if (!ld.isValid()) {
    if (rd.isSingle()) {
      l.setDistribution(SingleDistribution<L>());
    } else {
      l.setDistribution(BlockDistribution<L>());
      if (rd.isBlock()) r.setDistribution(CopyDistribution<R>());
    }
  } else if (!rd.isValid()) {
    if (ld.isSingle()) r.setDistribution(SingleDistribution<R>());
    else r.setDistribution(CopyDistribution<R>());
  } else if (ld.isCopy() && rd.isSingle()) {
    l.setDistribution(SingleDistribution<L>());
  } else if (ld.isSingle() && rd.isBlock()) {
    r.setDistribution(SingleDistribution<R>());
  } else if (rd.isSingle() || rd.isBlock()) {
    if (ld.isBlock() || ld.isCopy()) r.setDistribution(CopyDistribution<R>());
  }
}

class SkeletonBase {
public:
  SkeletonBase() = default;
  SkeletonBase(const SkeletonBase&) = delete;
  SkeletonBase& operator=(const SkeletonBase&) = delete;
  SkeletonBase(SkeletonBase&& o) noexcept : _h(o._h) { o._h = nullptr; }
  ~SkeletonBase()
  {
    if (_h) skl_skeleton_release(_h);
  }
  // Skeleton::setWorkGroupSize (src/Skeleton.cpp:75-78)
  size_t workGroupSize() const
  {
    int wg = 0;
    check(skl_skeleton_work_group_size(_h, &wg), "workGroupSize");
    return (size_t)wg;
  }
  void setWorkGroupSize(unsigned int wg)
  {
    check(skl_skeleton_set_work_group_size(_h, (int)wg), "setWorkGroupSize");
  }
  // precision form of the float paths (skl_precision)
  void setPrecision(int p) { check(skl_skeleton_set_precision(_h, p), "setPrecision"); }
  skl_skeleton handle() const { return _h; }

protected:
  void create(skl_skeleton_desc& d) { check(skl_skeleton_create(&d, &_h), "skeleton construction"); }
  static skl_skeleton_desc blankDesc(int kind)
  {
    skl_skeleton_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = kind;
    d.padding = SKL_PADDING_NEAREST;
    return d;
  }
  skl_skeleton _h = nullptr;
};

}  // namespace detail
}  // namespace skelcl

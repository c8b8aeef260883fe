// skelcl::Reduce<T(T)> (Reduce.h; detail/ReduceDef.h:79-293): folds a
// Vector into a Vector of size 1.  The identity is stored (and handed to
// AllPairs) but, as in the reference (ReduceDef.h:247-275), not used by the
// reduction itself.
#pragma once

#include <string>

#include "Zip.h"
#include "detail/Skeleton.h"

namespace skelcl {

template <typename> class Reduce;

template <typename T>
class Reduce<T(T)> : public detail::SkeletonBase {
public:
  Reduce(const Source& source, const std::string& id = "0", const std::string& funcName = "func")
      : _src(source), _id(id), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_REDUCE);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.identity = _id.c_str();
    d.types[0] = typeOf<T>();
    d.ntypes = 1;
    create(d);
  }

  template <typename... Args>
  Vector<T> operator()(const Vector<T>& in, Args&&... args)
  {
    Vector<T> output;
    (*this)(out(output), in, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Vector<T>& operator()(Out<Vector<T>> output, const Vector<T>& in, Args&&... args)
  {
    detail::defaultDistribution<detail::SingleDistribution>(in);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* src = in.devicePointer();
    Vector<T>& o = output.container();
    void* dst = o.size() >= 1 ? o.devicePointer() : (o.prepareOutput(1), o.deviceOutputPointer());
    detail::check(skl_reduce_run(_h, dst, src, in.size(), L.data(), L.count(), nullptr),
                  "Reduce");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, in);
    L.finish();
    return o;
  }

  // extension: Zip -> Reduce in one kernel (the zip output never reaches HBM)
  template <typename Tl, typename Tr>
  Vector<T> fused(Zip<T(Tl, Tr)>& zip, const Vector<Tl>& l, const Vector<Tr>& r)
  {
    if (l.size() != r.size()) detail::abortWithError("Zip: inputs differ in size");
    Vector<T> o;
    o.prepareOutput(1);
    detail::check(skl_zip_reduce_run(zip.handle(), _h, o.deviceOutputPointer(),
                                     l.devicePointer(), r.devicePointer(), l.size(), nullptr),
                  "Zip->Reduce");
    o.dataOnDeviceModified();
    return o;
  }

  // extension: adjoint of Reduce(+), d in_i = d out[0] for n inputs
  Vector<T> backward(const Vector<T>& dOut, size_t n)
  {
    Vector<T> d;
    d.prepareOutput(n);
    detail::check(skl_reduce_backward(_h, d.deviceOutputPointer(), dOut.devicePointer(), n,
                                      nullptr),
                  "Reduce backward");
    d.dataOnDeviceModified();
    return d;
  }

  // extension: batched segmented Reduce(+) -- out[s] folds
  // in[offsets[s] .. offsets[s+1]) -- all segments in one main launch after
  // small prologue launches (skelcl::batch does the sums and the scan in one)
  Vector<T> segmented(const Vector<T>& in, const Vector<long>& offsets)
  {
    Vector<T> o;
    int nseg = (int)offsets.size() - 1;
    o.prepareOutput((size_t)(nseg > 0 ? nseg : 0));
    detail::check(skl_reduce_segmented_run(_h, o.deviceOutputPointer(), in.devicePointer(),
                                           in.size(),
                                           static_cast<const int64_t*>(offsets.devicePointer()),
                                           nseg, nullptr),
                  "segmented Reduce");
    o.dataOnDeviceModified();
    return o;
  }

  std::string source() const { return _src; }
  std::string func() const { return _func; }
  std::string id() const { return _id; }

private:
  std::string _src, _id, _func;
};

}  // namespace skelcl

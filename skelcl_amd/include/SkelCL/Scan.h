// skelcl::Scan<T(T)> (Scan.h; detail/ScanDef.h:70-327): the EXCLUSIVE
// prefix scan under the user function with the given identity, as the
// reference computes it (ScanKernel.cl:142-153).  inclusive() is an
// extension.  Extra arguments reach the user function (the reference
// leaves them unset, ScanDef.h:206).
#pragma once

#include <string>

#include "detail/Skeleton.h"

namespace skelcl {

template <typename> class Scan;

template <typename T>
class Scan<T(T)> : public detail::SkeletonBase {
public:
  Scan(const Source& source, const std::string& id = "0", const std::string& funcName = "func")
      : _src(source), _id(id), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_SCAN);
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
    return run(output.container(), in, 0, std::forward<Args>(args)...);
  }
  template <typename... Args>
  Vector<T> inclusive(const Vector<T>& in, Args&&... args)
  {
    Vector<T> output;
    run(output, in, 1, std::forward<Args>(args)...);
    return output;
  }
  // extension: adjoint of Scan(+), the scan of dOut read from its end
  Vector<T> backward(const Vector<T>& dOut, bool inclusiveForward = false)
  {
    Vector<T> d;
    d.prepareOutput(dOut.size());
    detail::check(skl_scan_backward(_h, d.deviceOutputPointer(), dOut.devicePointer(),
                                    dOut.size(), inclusiveForward ? 1 : 0, nullptr),
                  "Scan backward");
    d.dataOnDeviceModified();
    return d;
  }

  // extension: batched segmented Scan(+), every segment on its own, one
  // main launch after a small prologue launch
  Vector<T> segmented(const Vector<T>& in, const Vector<long>& offsets, bool inclusive = false)
  {
    Vector<T> o;
    o.prepareOutput(in.size());
    int nseg = (int)offsets.size() - 1;
    detail::check(skl_scan_segmented_run(_h, o.deviceOutputPointer(), in.devicePointer(),
                                         in.size(),
                                         static_cast<const int64_t*>(offsets.devicePointer()),
                                         nseg, inclusive ? 1 : 0, nullptr),
                  "segmented Scan");
    o.dataOnDeviceModified();
    return o;
  }

  std::string source() const { return _src; }
  std::string func() const { return _func; }
  std::string id() const { return _id; }

private:
  template <typename... Args>
  Vector<T>& run(Vector<T>& o, const Vector<T>& in, int incl, Args&&... args)
  {
    detail::defaultDistribution<detail::SingleDistribution>(in);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* src = in.devicePointer();
    void* dst = (static_cast<const void*>(&o) == static_cast<const void*>(&in))
                    ? o.devicePointer()
                    : (o.prepareOutput(in.size()), o.deviceOutputPointer());
    detail::check(skl_scan_run(_h, dst, src, in.size(), incl, L.data(), L.count(), nullptr),
                  "Scan");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, in);
    L.finish();
    return o;
  }
  std::string _src, _id, _func;
};

}  // namespace skelcl

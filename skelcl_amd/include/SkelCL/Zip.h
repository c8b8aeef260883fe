// skelcl::Zip (Zip.h; detail/ZipDef.h:70-442): Zip<Tout(Tl, Tr)> and
// Zip<void(Tl, Tr)>.  Inputs must have equal sizes (ZipDef.h:101,126).
#pragma once

#include <string>

#include "detail/Skeleton.h"

namespace skelcl {

template <typename> class Zip;

template <typename Tout, typename Tl, typename Tr>
class Zip<Tout(Tl, Tr)> : public detail::SkeletonBase {
public:
  Zip(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_ZIP);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<Tout>();
    d.types[1] = typeOf<Tl>();
    d.types[2] = typeOf<Tr>();
    d.ntypes = 3;
    create(d);
  }
  // extension: with df/dl and df/dr (l, r, extras...) for backward()
  Zip(const Source& source, const std::string& funcName, const Source& derivSource,
      const std::string& dLeft, const std::string& dRight)
      : _src(source), _func(funcName), _dsrc(derivSource), _dl(dLeft), _dr(dRight)
  {
    skl_skeleton_desc d = blankDesc(SKL_ZIP);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.deriv_source = _dsrc.c_str();
    d.deriv_func = _dl.c_str();
    d.deriv_func2 = _dr.c_str();
    d.types[0] = typeOf<Tout>();
    d.types[1] = typeOf<Tl>();
    d.types[2] = typeOf<Tr>();
    d.ntypes = 3;
    create(d);
  }

  // extension: d_l = d_out * df/dl, d_r = d_out * df/dr (either may be null)
  template <typename... Args>
  void backward(const Vector<Tl>& l, const Vector<Tr>& r, const Vector<Tout>& dOut,
                Vector<Tl>* dL, Vector<Tr>* dR, Args&&... args)
  {
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    void* dl = nullptr;
    void* dr = nullptr;
    if (dL) {
      dL->prepareOutput(l.size());
      dl = dL->deviceOutputPointer();
    }
    if (dR) {
      dR->prepareOutput(r.size());
      dr = dR->deviceOutputPointer();
    }
    detail::check(skl_zip_backward(_h, dl, dr, l.devicePointer(), r.devicePointer(),
                                   dOut.devicePointer(), l.size(), L.data(), L.count(), nullptr),
                  "Zip backward");
    if (dL) dL->dataOnDeviceModified();
    if (dR) dR->dataOnDeviceModified();
  }

  template <typename... Args>
  Vector<Tout> operator()(const Vector<Tl>& l, const Vector<Tr>& r, Args&&... args)
  {
    Vector<Tout> output;
    (*this)(out(output), l, r, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Vector<Tout>& operator()(Out<Vector<Tout>> output, const Vector<Tl>& l, const Vector<Tr>& r,
                           Args&&... args)
  {
    if (l.size() != r.size()) detail::abortWithError("Zip: inputs differ in size");
    detail::zipDistributions(l, r);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* lp = l.devicePointer();
    const void* rp = r.devicePointer();
    Vector<Tout>& o = output.container();
    const void* op = &o;
    void* dst = (op == static_cast<const void*>(&l) || op == static_cast<const void*>(&r) ||
                 o.size() > l.size())
                    ? o.devicePointer()
                    : (o.prepareOutput(l.size()), o.deviceOutputPointer());
    detail::check(skl_zip_run(_h, dst, lp, rp, l.size(), L.data(), L.count(), nullptr), "Zip");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, l);
    L.finish();
    return o;
  }
  template <typename... Args>
  Matrix<Tout> operator()(const Matrix<Tl>& l, const Matrix<Tr>& r, Args&&... args)
  {
    Matrix<Tout> output;
    (*this)(out(output), l, r, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Matrix<Tout>& operator()(Out<Matrix<Tout>> output, const Matrix<Tl>& l, const Matrix<Tr>& r,
                           Args&&... args)
  {
    if (!(l.size() == r.size())) detail::abortWithError("Zip: inputs differ in size");
    detail::zipDistributions(l, r);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* lp = l.devicePointer();
    const void* rp = r.devicePointer();
    Matrix<Tout>& o = output.container();
    const void* op = &o;
    void* dst;
    if (op == static_cast<const void*>(&l) || op == static_cast<const void*>(&r)) {
      dst = o.devicePointer();
    } else {
      o.prepareOutputMatrix(l.size());
      dst = o.deviceOutputPointer();
    }
    detail::check(
        skl_zip_run(_h, dst, lp, rp, l.size().elemCount(), L.data(), L.count(), nullptr), "Zip");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, l);
    L.finish();
    return o;
  }
  std::string source() const { return _src; }
  std::string func() const { return _func; }

private:
  std::string _src, _func, _dsrc, _dl, _dr;
};

template <typename Tl, typename Tr>
class Zip<void(Tl, Tr)> : public detail::SkeletonBase {
public:
  Zip(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_ZIP);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<void>();
    d.types[1] = typeOf<Tl>();
    d.types[2] = typeOf<Tr>();
    d.ntypes = 3;
    create(d);
  }
  template <typename... Args>
  void operator()(const Vector<Tl>& l, const Vector<Tr>& r, Args&&... args)
  {
    if (l.size() != r.size()) detail::abortWithError("Zip: inputs differ in size");
    detail::zipDistributions(l, r);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    detail::check(skl_zip_run(_h, nullptr, l.devicePointer(), r.devicePointer(), l.size(),
                              L.data(), L.count(), nullptr),
                  "Zip");
    L.finish();
  }

private:
  std::string _src, _func;
};

}  // namespace skelcl

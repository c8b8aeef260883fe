// skelcl::Map (Map.h; detail/MapDef.h:75-752; src/Map.cpp:53-159):
// Map<Tout(Tin)>, Map<void(Tin)>, Map<Tout(Index)>, Map<void(Index)>,
// Map<Tout(IndexPoint)>, Map<void(IndexPoint)>.
#pragma once

#include <string>

#include "IndexMatrix.h"
#include "IndexVector.h"
#include "detail/Skeleton.h"

namespace skelcl {

template <typename> class Map;

template <typename Tout, typename Tin>
class Map<Tout(Tin)> : public detail::SkeletonBase {
public:
  Map(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAP);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<Tout>();
    d.types[1] = typeOf<Tin>();
    d.ntypes = 2;
    create(d);
  }
  // extension: with the derivative f'(x, extras...) for backward()
  Map(const Source& source, const std::string& funcName, const Source& derivSource,
      const std::string& derivFunc)
      : _src(source), _func(funcName), _dsrc(derivSource), _dfunc(derivFunc)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAP);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.deriv_source = _dsrc.c_str();
    d.deriv_func = _dfunc.c_str();
    d.types[0] = typeOf<Tout>();
    d.types[1] = typeOf<Tin>();
    d.ntypes = 2;
    create(d);
  }

  // extension: d_in = d_out * f'(in, extras...)
  template <typename... Args>
  Vector<Tin> backward(const Vector<Tin>& in, const Vector<Tout>& dOut, Args&&... args)
  {
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    Vector<Tin> d;
    d.prepareOutput(in.size());
    detail::check(skl_map_backward(_h, d.deviceOutputPointer(), in.devicePointer(),
                                   dOut.devicePointer(), in.size(), L.data(), L.count(), nullptr),
                  "Map backward");
    d.dataOnDeviceModified();
    return d;
  }

  template <typename... Args>
  Vector<Tout> operator()(const Vector<Tin>& in, Args&&... args)
  {
    Vector<Tout> output;
    (*this)(out(output), in, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Vector<Tout>& operator()(Out<Vector<Tout>> output, const Vector<Tin>& in, Args&&... args)
  {
    detail::defaultDistribution<detail::BlockDistribution>(in);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* src = in.devicePointer();
    Vector<Tout>& o = output.container();
    // an output at least as long keeps its size and tail (MapHelperDef.h:65-95)
    void* dst = (static_cast<const void*>(&o) == static_cast<const void*>(&in) ||
                 o.size() > in.size())
                    ? o.devicePointer()
                    : (o.prepareOutput(in.size()), o.deviceOutputPointer());
    detail::check(skl_map_run(_h, dst, src, in.size(), L.data(), L.count(), nullptr), "Map");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, in);
    L.finish();
    return o;
  }
  template <typename... Args>
  Matrix<Tout> operator()(const Matrix<Tin>& in, Args&&... args)
  {
    Matrix<Tout> output;
    (*this)(out(output), in, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Matrix<Tout>& operator()(Out<Matrix<Tout>> output, const Matrix<Tin>& in, Args&&... args)
  {
    detail::defaultDistribution<detail::BlockDistribution>(in);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* src = in.devicePointer();
    Matrix<Tout>& o = output.container();
    o.prepareOutputMatrix(in.size());
    detail::check(skl_map_run(_h, o.deviceOutputPointer(), src, in.size().elemCount(), L.data(),
                              L.count(), nullptr),
                  "Map");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, in);
    L.finish();
    return o;
  }
  std::string source() const { return _src; }
  std::string func() const { return _func; }

private:
  std::string _src, _func, _dsrc, _dfunc;
};

template <typename Tin>
class Map<void(Tin)> : public detail::SkeletonBase {
public:
  Map(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAP);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<void>();
    d.types[1] = typeOf<Tin>();
    d.ntypes = 2;
    create(d);
  }
  template <typename C, typename... Args>
  void operator()(const C& in, Args&&... args)
  {
    detail::defaultDistribution<detail::BlockDistribution>(in);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    size_t n = elementCount(in);
    detail::check(skl_map_run(_h, nullptr, in.devicePointer(), n, L.data(), L.count(), nullptr),
                  "Map");
    L.finish();
  }

private:
  template <typename T> static size_t elementCount(const Vector<T>& v) { return v.size(); }
  template <typename T> static size_t elementCount(const Matrix<T>& m)
  {
    return m.size().elemCount();
  }
  std::string _src, _func;
};

template <typename Tout>
class Map<Tout(Index)> : public detail::SkeletonBase {
public:
  Map(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAP_INDEX);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<Tout>();
    d.ntypes = 1;
    create(d);
  }
  template <typename... Args>
  Vector<Tout> operator()(IndexVector& in, Args&&... args)
  {
    Vector<Tout> output;
    (*this)(out(output), in, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Vector<Tout>& operator()(Out<Vector<Tout>> output, IndexVector& in, Args&&... args)
  {
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    Vector<Tout>& o = output.container();
    o.prepareOutput(in.size());
    detail::check(skl_map_index_run(_h, o.deviceOutputPointer(), in.size(), 0, L.data(),
                                    L.count(), nullptr),
                  "Map<Index>");
    o.dataOnDeviceModified();
    L.finish();
    return o;
  }

private:
  std::string _src, _func;
};

template <>
class Map<void(Index)> : public detail::SkeletonBase {
public:
  Map(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAP_INDEX);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<void>();
    d.ntypes = 1;
    create(d);
  }
  template <typename... Args>
  void operator()(IndexVector& in, Args&&... args)
  {
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    detail::check(skl_map_index_run(_h, nullptr, in.size(), 0, L.data(), L.count(), nullptr),
                  "Map<void(Index)>");
    L.finish();
  }

private:
  std::string _src, _func;
};

template <typename Tout>
class Map<Tout(IndexPoint)> : public detail::SkeletonBase {
public:
  Map(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAP_INDEXPOINT);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<Tout>();
    d.ntypes = 1;
    create(d);
  }
  template <typename... Args>
  Matrix<Tout> operator()(IndexMatrix& in, Args&&... args)
  {
    Matrix<Tout> output;
    (*this)(out(output), in, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Matrix<Tout>& operator()(Out<Matrix<Tout>> output, IndexMatrix& in, Args&&... args)
  {
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    Matrix<Tout>& o = output.container();
    o.prepareOutputMatrix(in.size());
    detail::check(skl_map_indexpoint_run(_h, o.deviceOutputPointer(), in.rowCount(),
                                         in.columnCount(), 0, L.data(), L.count(), nullptr),
                  "Map<IndexPoint>");
    o.dataOnDeviceModified();
    L.finish();
    return o;
  }

private:
  std::string _src, _func;
};

template <>
class Map<void(IndexPoint)> : public detail::SkeletonBase {
public:
  Map(const Source& source, const std::string& funcName = "func") : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAP_INDEXPOINT);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<void>();
    d.ntypes = 1;
    create(d);
  }
  template <typename... Args>
  void operator()(IndexMatrix& in, Args&&... args)
  {
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    detail::check(skl_map_indexpoint_run(_h, nullptr, in.rowCount(), in.columnCount(), 0,
                                         L.data(), L.count(), nullptr),
                  "Map<void(IndexPoint)>");
    L.finish();
  }

private:
  std::string _src, _func;
};

}  // namespace skelcl

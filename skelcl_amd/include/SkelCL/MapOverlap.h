// skelcl::MapOverlap<Tout(Tin)> (MapOverlap.h:119-122 defaults: range 1,
// NEAREST, neutral 0; detail/MapOverlapDef.h:71-279): the user function
// reads its neighbourhood through getData(m, x, y) (x columns right, y rows
// down of the centre); reads outside the matrix give the neutral value or
// the nearest edge element.
#pragma once

#include <string>

#include "Distributions.h"
#include "detail/Skeleton.h"

namespace skelcl {

template <typename> class MapOverlap;

template <typename Tout, typename Tin>
class MapOverlap<Tout(Tin)> : public detail::SkeletonBase {
public:
  MapOverlap(const Source& source, unsigned int overlap_range = 1,
             detail::Padding padding = detail::Padding::NEAREST, Tin neutral_element = Tin(),
             const std::string& funcName = "func")
      : _src(source), _func(funcName), _range(overlap_range), _padding(padding),
        _neutral(detail::valueString(neutral_element)), _neutralValue(neutral_element)
  {
    skl_skeleton_desc d = blankDesc(SKL_MAPOVERLAP);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.types[0] = typeOf<Tout>();
    d.types[1] = typeOf<Tin>();
    d.ntypes = 2;
    d.overlap_range = (int)overlap_range;
    d.padding = padding == detail::Padding::NEUTRAL ? SKL_PADDING_NEUTRAL : SKL_PADDING_NEAREST;
    d.neutral = _neutral.c_str();
    create(d);
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
    if (in.rowCount() == 0 || in.columnCount() == 0)
      detail::abortWithError("MapOverlap: empty matrix");
    // MapOverlapDef.h:250-262: the input always gets this skeleton's overlap
    in.setDistribution(detail::OLDistribution<Matrix<Tin>>(_range, _padding, _neutralValue));
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* src = in.devicePointer();
    Matrix<Tout>& o = output.container();
    o.prepareOutputMatrix(in.size());
    detail::check(skl_mapoverlap_run(_h, o.deviceOutputPointer(), src, in.rowCount(),
                                     in.columnCount(), L.data(), L.count(), nullptr),
                  "MapOverlap");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, in);
    L.finish();
    return o;
  }
  unsigned int overlapRange() const { return _range; }

  // extension: adjoint of a weighted-stencil user function
  // sum_{dy,dx} w[(dy+R)(2R+1)+dx+R] * getData(m, dx, dy): the transposed
  // stencil of dOut under the same padding
  Matrix<Tin> stencilBackward(const Matrix<Tout>& dOut, const Vector<float>& weights)
  {
    Matrix<Tin> d;
    d.prepareOutputMatrix(dOut.size());
    detail::check(skl_mapoverlap_stencil_backward(
                      _h, d.deviceOutputPointer(), dOut.devicePointer(),
                      static_cast<const float*>(weights.devicePointer()), dOut.rowCount(),
                      dOut.columnCount(), nullptr),
                  "MapOverlap backward");
    d.dataOnDeviceModified();
    return d;
  }

private:
  std::string _src, _func;
  unsigned int _range;
  detail::Padding _padding;
  std::string _neutral;
  Tin _neutralValue;
};

}  // namespace skelcl

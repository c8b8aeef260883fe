// skelcl::AllPairs<Tout(Tl, Tr)> (AllPairs.h; detail/AllPairsDef.h:77-426):
//   AllPairs(reduce, zip): P[i][j] = fold_k reduce(acc, zip(M[i][k], N[k][j]))
//                          from the reduce's identity, k in order;
//   AllPairs(source, func): P[i][j] = f(&row_i, &col_j, dim).
// left is rows x dim, right dim x cols (row-major).  Extra arguments follow
// the reduce's, then the zip's (AllPairsDef.h:264-269).  setPrecision
// selects the MFMA forms of the (+,*) contraction on float.
#pragma once

#include <string>

#include "Reduce.h"
#include "Zip.h"
#include "detail/Skeleton.h"

namespace skelcl {

template <typename> class AllPairs;

template <typename Tout, typename Tl, typename Tr>
class AllPairs<Tout(Tl, Tr)> : public detail::SkeletonBase {
public:
  AllPairs(const Reduce<Tout(Tout)>& reduce, const Zip<Tout(Tl, Tr)>& zip)
      : _src(reduce.source()), _func(reduce.func()), _src2(zip.source()), _func2(zip.func()),
        _id(reduce.id())
  {
    skl_skeleton_desc d = blankDesc(SKL_ALLPAIRS);
    d.source = _src.c_str();
    d.func = _func.c_str();
    d.source2 = _src2.c_str();
    d.func2 = _func2.c_str();
    d.identity = _id.c_str();
    setTypes(d);
    create(d);
  }
  AllPairs(const Source& source, const std::string& funcName = "func")
      : _src(source), _func(funcName)
  {
    skl_skeleton_desc d = blankDesc(SKL_ALLPAIRS_GENERAL);
    d.source = _src.c_str();
    d.func = _func.c_str();
    setTypes(d);
    create(d);
  }

  template <typename... Args>
  Matrix<Tout> operator()(const Matrix<Tl>& left, const Matrix<Tr>& right, Args&&... args)
  {
    Matrix<Tout> output;
    (*this)(out(output), left, right, std::forward<Args>(args)...);
    return output;
  }
  template <typename... Args>
  Matrix<Tout>& operator()(Out<Matrix<Tout>> output, const Matrix<Tl>& left,
                           const Matrix<Tr>& right, Args&&... args)
  {
    if (left.columnCount() != right.rowCount())  // AllPairsDef.h:127-129
      detail::abortWithError("AllPairs: left.columnCount() != right.rowCount()");
    detail::allPairsDistributions(left, right);
    auto L = detail::makeArgs(std::forward<Args>(args)...);
    const void* lp = left.devicePointer();
    const void* rp = right.devicePointer();
    Matrix<Tout>& o = output.container();
    o.prepareOutputMatrix(MatrixSize(left.rowCount(), right.columnCount()));
    detail::check(skl_allpairs_run(_h, o.deviceOutputPointer(), lp, rp, left.rowCount(),
                                   right.columnCount(), left.columnCount(), L.data(), L.count(),
                                   nullptr),
                  "AllPairs");
    o.dataOnDeviceModified();
    detail::adoptDistribution(o, left);
    L.finish();
    return o;
  }

  // extension: adjoint of the (+,*) contraction, dA = dC.B^T, dB = A^T.dC
  void backward(const Matrix<Tl>& left, const Matrix<Tr>& right, const Matrix<Tout>& dC,
                Matrix<Tl>* dA, Matrix<Tr>* dB)
  {
    void* da = nullptr;
    void* db = nullptr;
    if (dA) {
      dA->prepareOutputMatrix(left.size());
      da = dA->deviceOutputPointer();
    }
    if (dB) {
      dB->prepareOutputMatrix(right.size());
      db = dB->deviceOutputPointer();
    }
    detail::check(skl_allpairs_backward(_h, da, db, left.devicePointer(), right.devicePointer(),
                                        dC.devicePointer(), left.rowCount(),
                                        right.columnCount(), left.columnCount(), nullptr),
                  "AllPairs backward");
    if (dA) dA->dataOnDeviceModified();
    if (dB) dB->dataOnDeviceModified();
  }

private:
  static void setTypes(skl_skeleton_desc& d)
  {
    d.types[0] = typeOf<Tout>();
    d.types[1] = typeOf<Tl>();
    d.types[2] = typeOf<Tr>();
    d.ntypes = 3;
  }
  std::string _src, _func, _src2, _func2, _id;
};

}  // namespace skelcl

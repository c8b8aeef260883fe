// Forward/backward Function pairs over the float skeletons (extension, no
// counterpart in the reference), shaped like torch.autograd.Function: a
// static forward(ctx, ...) saves what backward(ctx, grad) needs in the
// Context; backward returns the vector-Jacobian products.
//   MapFn          y = Map(x)            dx = dy * f'(x)        (x saved)
//   ZipFn          y = Zip(a, b)         da = dy*df/da, db = dy*df/db
//   ReduceSumFn    y = Reduce(+)(x)      dx_i = dy               (n saved)
//   ScanSumFn      y = Scan(+)(x)        dx = reversed scan of dy
//   StencilFn      y = MapOverlap(x, w)  dx = transposed stencil of dy; the
//                                        forward is not stored
//   MatMulFn       C = AllPairs(+,*)(A,B) dA = dC.B^T, dB = A^T.dC
#pragma once

#include <memory>
#include <vector>

#include "SkelCL.h"

namespace skelcl {
namespace autodiff {

class Context {
public:
  template <typename T>
  void save(const T& v)
  {
    _saved.push_back(std::make_shared<Holder<T>>(v));
  }
  template <typename T>
  const T& saved(size_t i) const
  {
    return static_cast<const Holder<T>*>(_saved.at(i).get())->v;
  }
  size_t savedCount() const { return _saved.size(); }
  size_t n = 0;  // a size backward may need
  bool flag = false;

private:
  struct Base {
    virtual ~Base() = default;
  };
  template <typename T> struct Holder : Base {
    explicit Holder(const T& x) : v(x) {}
    T v;
  };
  std::vector<std::shared_ptr<Base>> _saved;
};

template <typename T>
struct MapFn {
  static Vector<T> forward(Context& ctx, Map<T(T)>& m, const Vector<T>& x)
  {
    ctx.save(x);
    return m(x);
  }
  static Vector<T> backward(Context& ctx, Map<T(T)>& m, const Vector<T>& dy)
  {
    return m.backward(ctx.saved<Vector<T>>(0), dy);
  }
};

template <typename T>
struct ZipFn {
  static Vector<T> forward(Context& ctx, Zip<T(T, T)>& z, const Vector<T>& a, const Vector<T>& b)
  {
    ctx.save(a);
    ctx.save(b);
    return z(a, b);
  }
  static std::pair<Vector<T>, Vector<T>> backward(Context& ctx, Zip<T(T, T)>& z,
                                                  const Vector<T>& dy)
  {
    Vector<T> da, db;
    z.backward(ctx.saved<Vector<T>>(0), ctx.saved<Vector<T>>(1), dy, &da, &db);
    return {std::move(da), std::move(db)};
  }
};

template <typename T>
struct ReduceSumFn {
  static Vector<T> forward(Context& ctx, Reduce<T(T)>& r, const Vector<T>& x)
  {
    ctx.n = x.size();
    return r(x);
  }
  static Vector<T> backward(Context& ctx, Reduce<T(T)>& r, const Vector<T>& dy)
  {
    return r.backward(dy, ctx.n);
  }
};

template <typename T>
struct ScanSumFn {
  static Vector<T> forward(Context& ctx, Scan<T(T)>& s, const Vector<T>& x, bool inclusive)
  {
    ctx.flag = inclusive;
    return inclusive ? s.inclusive(x) : s(x);
  }
  static Vector<T> backward(Context& ctx, Scan<T(T)>& s, const Vector<T>& dy)
  {
    return s.backward(dy, ctx.flag);
  }
};

struct StencilFn {
  static Matrix<float> forward(Context& ctx, MapOverlap<float(float)>& m, const Matrix<float>& x,
                               const Vector<float>& w)
  {
    ctx.save(w);
    return m(x, w);
  }
  static Matrix<float> backward(Context& ctx, MapOverlap<float(float)>& m,
                                const Matrix<float>& dy)
  {
    return m.stencilBackward(dy, ctx.saved<Vector<float>>(0));
  }
};

struct MatMulFn {
  static Matrix<float> forward(Context& ctx, AllPairs<float(float, float)>& ap,
                               const Matrix<float>& A, const Matrix<float>& B)
  {
    ctx.save(A);
    ctx.save(B);
    return ap(A, B);
  }
  static std::pair<Matrix<float>, Matrix<float>> backward(Context& ctx,
                                                          AllPairs<float(float, float)>& ap,
                                                          const Matrix<float>& dC)
  {
    Matrix<float> dA, dB;
    ap.backward(ctx.saved<Matrix<float>>(0), ctx.saved<Matrix<float>>(1), dC, &dA, &dB);
    return {std::move(dA), std::move(dB)};
  }
};

}  // namespace autodiff
}  // namespace skelcl

// skelcl::IndexVector = Vector<Index> (IndexVector.h:54-230): the indices
// 0..n-1 without storage; input of Map<Tout(Index)>.
#pragma once

#include "Index.h"
#include "Vector.h"

namespace skelcl {

template <>
class Vector<Index> {
public:
  typedef Index value_type;
  typedef size_t size_type;
  explicit Vector(size_type n = 0) : _n(n) {}
  size_type size() const { return _n; }
  Index operator[](size_type i) const { return Index(i); }
  Index front() const { return Index(0); }
  Index back() const { return Index(_n - 1); }

private:
  size_type _n;
};

typedef const Vector<Index> IndexVector;

}  // namespace skelcl

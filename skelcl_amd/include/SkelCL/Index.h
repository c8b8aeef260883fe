// skelcl::Index / IndexPoint (Index.h): the element types of the index
// containers; on the device Index is a size_t and IndexPoint {x, y}.
#pragma once

#include <cstddef>

#include "detail/Runtime.h"

namespace skelcl {

class Index {
public:
  typedef size_t index_type;
  Index() : _i(0) {}
  Index(index_type i) : _i(i) {}
  operator index_type() const { return _i; }
  bool operator==(const Index& r) const { return _i == r._i; }
  bool operator!=(const Index& r) const { return _i != r._i; }

private:
  index_type _i;
};

struct IndexPoint {
  size_t x, y;
  IndexPoint() : x(0), y(0) {}
  IndexPoint(size_t x_, size_t y_) : x(x_), y(y_) {}
  bool operator==(const IndexPoint& r) const { return x == r.x && y == r.y; }
};

SKELCL_TYPE_NAME(Index, "Index")
SKELCL_TYPE_NAME(IndexPoint, "IndexPoint")

}  // namespace skelcl

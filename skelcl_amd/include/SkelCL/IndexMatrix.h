// skelcl::IndexMatrix = Matrix<IndexPoint> (IndexMatrix.h): the points
// (x = column, y = row) without storage; input of Map<Tout(IndexPoint)>.
#pragma once

#include "Index.h"
#include "Matrix.h"

namespace skelcl {

template <>
class Matrix<IndexPoint> {
public:
  typedef IndexPoint value_type;
  typedef MatrixSize size_type;
  explicit Matrix(const size_type size = size_type()) : _size(size) {}
  size_type size() const { return _size; }
  size_t rowCount() const { return _size.rowCount(); }
  size_t columnCount() const { return _size.columnCount(); }
  IndexPoint operator()(size_t row, size_t col) const { return IndexPoint(col, row); }

private:
  size_type _size;
};

typedef const Matrix<IndexPoint> IndexMatrix;

}  // namespace skelcl

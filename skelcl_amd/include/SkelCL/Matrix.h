// skelcl::Matrix<T> (Matrix.h, detail/MatrixDef.h): row-major 2-D container
// (MatrixDef.h:292,307,421) with the same lazy coherence as Vector.
#pragma once

#include <vector>

#include "Vector.h"

namespace skelcl {

class MatrixSize {
public:
  typedef size_t size_type;
  MatrixSize() : _r(0), _c(0) {}
  MatrixSize(size_type rows, size_type cols) : _r(rows), _c(cols) {}
  size_type elemCount() const { return _r * _c; }
  size_type rowCount() const { return _r; }
  size_type columnCount() const { return _c; }
  bool operator==(const MatrixSize& o) const { return _r == o._r && _c == o._c; }
  bool operator!=(const MatrixSize& o) const { return !(*this == o); }

private:
  size_type _r, _c;
};

template <typename T>
class Matrix : public detail::HostDeviceData<T> {
  typedef detail::HostDeviceData<T> base;

public:
  typedef MatrixSize size_type;
  typedef T value_type;
  typedef T& reference;
  typedef const T& const_reference;
  typedef typename std::vector<T>::iterator iterator;
  typedef typename std::vector<T>::const_iterator const_iterator;
  struct coordinate {
    typedef size_t index_type;
    index_type rowIndex, columnIndex;
  };

  Matrix() : _size() { this->setElements(0); }
  explicit Matrix(const size_type size, const T& v = T())
      : base(size.elemCount(), v), _size(size)
  {
    this->setElements(size.elemCount());
  }
  // Matrix(vector, columnCount): rows = ceil(n / columns), zero-filled tail
  Matrix(const std::vector<T>& v, size_t columnCount)
      : base(std::vector<T>(v)), _size((v.size() + columnCount - 1) / columnCount, columnCount)
  {
    this->_host.resize(_size.elemCount());
    this->setElements(_size.elemCount());
  }
  Matrix(const std::vector<T>& v, const size_type size) : base(std::vector<T>(v)), _size(size)
  {
    this->_host.resize(size.elemCount());
    this->setElements(size.elemCount());
  }
  Matrix(const Matrix& o) : base(o), _size(o._size) { this->setElements(_size.elemCount()); }
  Matrix& operator=(const Matrix& o)
  {
    base::operator=(o);
    _size = o._size;
    this->setElements(_size.elemCount());
    return *this;
  }
  Matrix(Matrix&&) = default;
  Matrix& operator=(Matrix&&) = default;

  size_type size() const { return _size; }
  typename size_type::size_type rowCount() const { return _size.rowCount(); }
  typename size_type::size_type columnCount() const { return _size.columnCount(); }

  // m[row][col] (MatrixDef.h:421: row-major)
  iterator operator[](size_t row)
  {
    this->copyDataToHost();
    return this->_host.begin() + row * _size.columnCount();
  }
  const_iterator operator[](size_t row) const
  {
    this->copyDataToHost();
    return this->_host.cbegin() + row * _size.columnCount();
  }
  reference operator()(coordinate c)
  {
    this->copyDataToHost();
    return this->_host[c.rowIndex * _size.columnCount() + c.columnIndex];
  }
  const_reference operator()(coordinate c) const
  {
    this->copyDataToHost();
    return this->_host[c.rowIndex * _size.columnCount() + c.columnIndex];
  }
  iterator begin()
  {
    this->copyDataToHost();
    return this->_host.begin();
  }
  iterator end()
  {
    this->copyDataToHost();
    return this->_host.end();
  }
  void resize(const size_type& s, T v = T())
  {
    this->copyDataToHost();
    std::vector<T> h(s.elemCount(), v);
    for (size_t r = 0; r < std::min(s.rowCount(), _size.rowCount()); ++r)
      for (size_t c = 0; c < std::min(s.columnCount(), _size.columnCount()); ++c)
        h[r * s.columnCount() + c] = this->_host[r * _size.columnCount() + c];
    this->_host.swap(h);
    _size = s;
    this->setElements(s.elemCount());
    this->dropDevice();
    this->dataOnHostModified();
  }
  void prepareOutputMatrix(const size_type& s)
  {
    if (!(s == _size)) _size = s;
    this->prepareOutput(s.elemCount());
  }
  void createDeviceBuffers() const { this->ensureDevice(); }
  void startUpload() const { this->copyDataToDevice(); }

private:
  size_type _size;
};

}  // namespace skelcl

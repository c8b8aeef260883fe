// skelcl::Matrix<T> (Matrix.h, detail/MatrixDef.h): row-major 2-D container
// (MatrixDef.h:292,307,421) with the same lazy coherence as Vector, the
// reference's constructors (MatrixDef.h:77-215: size + value, vector +
// column count or size, iterator ranges, from2DVector), its row operations
// (MatrixDef.h:289-651) and a recorded distribution.  Using Matrix<T> also
// registers T_matrix_t for T (MatrixDef.h:66-74, 834-867) as a generated
// common definition, so user functions can take a Matrix of any element
// type as an extra argument.
#pragma once

#include <algorithm>
#include <cctype>
#include <stdexcept>
#include <initializer_list>
#include <iterator>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "Source.h"
#include "Vector.h"
#include "detail/Util.h"

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
  // ordered by element count, as the reference (src/MatrixSize.cpp)
  bool operator<(const MatrixSize& o) const { return elemCount() < o.elemCount(); }
  bool operator>(const MatrixSize& o) const { return o < *this; }
  bool operator<=(const MatrixSize& o) const { return !(o < *this); }
  bool operator>=(const MatrixSize& o) const { return !(*this < o); }

private:
  size_type _r, _c;
};

inline std::ostream& operator<<(std::ostream& o, const MatrixSize& s)
{
  return o << "{" << s.rowCount() << ", " << s.columnCount() << "}";
}

template <typename T>
class RegisterMatrixDeviceFunctions {
public:
  RegisterMatrixDeviceFunctions()
  {
    detail::CommonDefinitions::append(Matrix<T>::deviceFunctions(),
                                      detail::CommonDefinitions::GENERATED_DEFINITION);
  }
};

template <typename T>
class Matrix : public detail::HostDeviceData<T> {
  typedef detail::HostDeviceData<T> base;
  typedef detail::Distribution<Matrix<T>> dist_type;

public:
  typedef std::vector<T> host_buffer_type;
  typedef typename host_buffer_type::value_type value_type;
  typedef typename host_buffer_type::pointer pointer;
  typedef typename host_buffer_type::const_pointer const_pointer;
  typedef typename host_buffer_type::reference reference;
  typedef typename host_buffer_type::const_reference const_reference;
  typedef typename host_buffer_type::iterator iterator;
  typedef typename host_buffer_type::const_iterator const_iterator;
  typedef MatrixSize size_type;
  typedef typename host_buffer_type::difference_type difference_type;
  typedef typename host_buffer_type::allocator_type allocator_type;
  struct coordinate {
    typedef size_type::size_type index_type;
    index_type rowIndex, columnIndex;
  };

  Matrix() : _size(), _dist(new dist_type())
  {
    (void)registerMatrixDeviceFunctions;
    this->setElements(0);
  }
  explicit Matrix(const size_type size, const T& v = T(), const dist_type& d = dist_type())
      : base(size.elemCount(), v), _size(size), _dist(detail::cloneAndConvert<Matrix<T>>(d))
  {
    (void)registerMatrixDeviceFunctions;
    this->setElements(size.elemCount());
  }
  // rows = ceil(n / columnCount); the last row is filled with T()
  Matrix(const std::vector<T>& v, typename size_type::size_type columnCount,
         const dist_type& d = dist_type())
      : base(host_buffer_type(v)),
        _size((v.size() + columnCount - 1) / columnCount, columnCount),
        _dist(detail::cloneAndConvert<Matrix<T>>(d))
  {
    (void)registerMatrixDeviceFunctions;
    this->_host.resize(_size.elemCount());
    this->setElements(_size.elemCount());
  }
  Matrix(const std::vector<T>& v, const size_type size, const dist_type& d = dist_type())
      : base(host_buffer_type(v)), _size(size), _dist(detail::cloneAndConvert<Matrix<T>>(d))
  {
    (void)registerMatrixDeviceFunctions;
    this->_host.resize(size.elemCount());
    this->setElements(size.elemCount());
  }
  template <class It, class = typename std::iterator_traits<It>::iterator_category>
  Matrix(It first, It last, typename size_type::size_type columnCount,
         const dist_type& d = dist_type())
      : base(host_buffer_type(first, last)), _dist(detail::cloneAndConvert<Matrix<T>>(d))
  {
    (void)registerMatrixDeviceFunctions;
    const size_t n = this->_host.size();
    _size = size_type((n + columnCount - 1) / columnCount, columnCount);
    this->_host.resize(_size.elemCount());
    this->setElements(_size.elemCount());
  }
  template <class It, class = typename std::iterator_traits<It>::iterator_category>
  Matrix(It first, It last, const size_type size, const dist_type& d = dist_type())
      : base(host_buffer_type(first, last)), _size(size),
        _dist(detail::cloneAndConvert<Matrix<T>>(d))
  {
    (void)registerMatrixDeviceFunctions;
    this->_host.resize(size.elemCount());
    this->setElements(size.elemCount());
  }
  // one row per inner vector; the column count is that of the first row
  // (MatrixDef.h:147-171)
  static Matrix from2DVector(const std::vector<std::vector<T>>& rows,
                             const dist_type& d = dist_type())
  {
    const size_t cols = rows.empty() ? 0 : rows[0].size();
    Matrix m(size_type(rows.size(), cols), T(), d);
    auto it = m._host.begin();
    for (auto& r : rows) {
      std::copy(r.begin(), r.begin() + std::min(r.size(), cols), it);
      it += cols;
    }
    return m;
  }
  Matrix(const Matrix& o)
      : base(o), _size(o._size), _dist(detail::cloneAndConvert<Matrix<T>>(*o._dist))
  {
    this->setElements(_size.elemCount());
  }
  Matrix& operator=(const Matrix& o)
  {
    if (this != &o) {
      base::operator=(o);
      _size = o._size;
      _dist = detail::cloneAndConvert<Matrix<T>>(*o._dist);
      this->setElements(_size.elemCount());
    }
    return *this;
  }
  Matrix(Matrix&&) = default;
  Matrix& operator=(Matrix&&) = default;

  // ---- size
  size_type size() const { return _size; }
  typename size_type::size_type rowCount() const { return _size.rowCount(); }
  typename size_type::size_type columnCount() const { return _size.columnCount(); }
  typename size_type::size_type max_size() const { return this->_host.max_size(); }
  typename size_type::size_type capacity() const { return this->_host.capacity(); }
  bool empty() const { return _size.elemCount() == 0; }
  void reserve(typename size_type::size_type n)
  {
    this->copyDataToHost();
    this->_host.reserve(n);
  }

  // ---- element access (row-major, MatrixDef.h:289-343, 421)
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
  const_iterator begin() const
  {
    this->copyDataToHost();
    return this->_host.cbegin();
  }
  const_iterator end() const
  {
    this->copyDataToHost();
    return this->_host.cend();
  }
  // the first element of row r, end() past the last row
  iterator row_begin(typename coordinate::index_type r)
  {
    this->copyDataToHost();
    const size_t n = r * _size.columnCount();
    return n < _size.elemCount() ? this->_host.begin() + n : this->_host.end();
  }
  const_iterator row_begin(typename coordinate::index_type r) const
  {
    this->copyDataToHost();
    const size_t n = r * _size.columnCount();
    return n < _size.elemCount() ? this->_host.cbegin() + n : this->_host.cend();
  }
  iterator row_end(typename coordinate::index_type r) { return row_begin(r + 1); }
  const_iterator row_end(typename coordinate::index_type r) const { return row_begin(r + 1); }
  iterator operator[](typename coordinate::index_type r) { return row_begin(r); }
  const_iterator operator[](typename coordinate::index_type r) const { return row_begin(r); }
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
  reference at(coordinate c)
  {
    if (c.rowIndex >= rowCount() || c.columnIndex >= columnCount())
      throw std::out_of_range("Matrix::at");
    return (*this)(c);
  }
  const_reference at(coordinate c) const
  {
    if (c.rowIndex >= rowCount() || c.columnIndex >= columnCount())
      throw std::out_of_range("Matrix::at");
    return (*this)(c);
  }
  reference front() { return *begin(); }
  const_reference front() const { return *begin(); }
  reference back() { return *(end() - 1); }
  const_reference back() const { return *(end() - 1); }

  // ---- modifiers: change the host copy and drop the device copy
  // resize keeps the row-major prefix (MatrixDef.h:370-395)
  void resize(const size_type& s, T v = T())
  {
    modify([&](host_buffer_type& h) { h.resize(s.elemCount(), v); }, s);
  }
  template <class It>
  void assign(It first, It last)
  {
    modify(
        [&](host_buffer_type& h) {
          h.assign(first, last);
          h.resize(_size.elemCount());
        },
        _size);
  }
  void assign(size_type s, const T& v)
  {
    modify([&](host_buffer_type& h) { h.assign(s.elemCount(), v); }, s);
  }
  // a row from a range: cut to the column count, filled up with T()
  template <class It, class = typename std::iterator_traits<It>::iterator_category>
  void push_back_row(It first, It last)
  {
    insert_row(rowCount(), first, last);
  }
  void push_back_row(std::initializer_list<T> list) { push_back_row(list.begin(), list.end()); }
  void pop_back_row()
  {
    if (rowCount() == 0) return;
    resize(size_type(rowCount() - 1, columnCount()));
  }
  template <class It, class = typename std::iterator_traits<It>::iterator_category>
  void insert_row(typename coordinate::index_type r, It first, It last)
  {
    const size_t cols = columnCount();
    std::vector<T> row(first, last);
    row.resize(cols);
    const size_t at = std::min<size_t>(r, rowCount()) * cols;
    modify([&](host_buffer_type& h) { h.insert(h.begin() + at, row.begin(), row.end()); },
           size_type(rowCount() + 1, cols));
  }
  void insert_row(typename coordinate::index_type r, std::initializer_list<T> list)
  {
    insert_row(r, list.begin(), list.end());
  }
  iterator erase_row(typename coordinate::index_type r)
  {
    const size_t cols = columnCount();
    if (r >= rowCount()) return end();
    modify([&](host_buffer_type& h) { h.erase(h.begin() + r * cols, h.begin() + (r + 1) * cols); },
           size_type(rowCount() - 1, cols));
    return this->_host.begin() + r * cols;
  }
  void clear() { resize(size_type(0, 0)); }

  host_buffer_type& hostBuffer() const
  {
    this->copyDataToHost();
    return this->_host;
  }
  void prepareOutputMatrix(const size_type& s)
  {
    if (!(s == _size)) _size = s;
    this->prepareOutput(s.elemCount());
  }

  // ---- distribution (MatrixDef.h:653-700)
  dist_type& distribution() const { return *_dist; }
  template <typename U>
  void setDistribution(const detail::Distribution<Matrix<U>>& d) const
  {
    _dist = detail::cloneAndConvert<Matrix<T>>(d);
  }
  template <typename U>
  void setDistribution(const std::unique_ptr<detail::Distribution<Matrix<U>>>& d) const
  {
    setDistribution(*d);
  }
  void setDistribution(std::unique_ptr<dist_type>&& d) const
  {
    if (d) _dist = std::move(d);
  }

  // T_matrix_t for this element type (MatrixDef.h:834-867); the kernel
  // prelude defines the built-in element types' own, so the guard skips them
  static std::string deviceFunctions()
  {
    const std::string t = detail::util::typeToString<T>();
    for (char c : t)  // "unsigned int" and templates have no such name
      if (!(std::isalnum((unsigned char)c) || c == '_')) return std::string();
    std::ostringstream s;
    s << "#ifndef " << t << "_MATRIX_T\n"
      << "typedef struct {\n  " << t << "* data;\n  unsigned int col_count;\n} " << t
      << "_matrix_t;\n#define " << t << "_MATRIX_T\n#endif\n";
    return s.str();
  }

private:
  template <typename F>
  void modify(F f, const size_type& newSize)
  {
    this->copyDataToHost();
    f(this->_host);
    _size = newSize;
    this->setElements(_size.elemCount());
    this->dropDevice();
    this->dataOnHostModified();
  }

  static RegisterMatrixDeviceFunctions<T> registerMatrixDeviceFunctions;

  size_type _size;
  mutable std::unique_ptr<dist_type> _dist;
};

template <typename T>
RegisterMatrixDeviceFunctions<T> Matrix<T>::registerMatrixDeviceFunctions;

}  // namespace skelcl

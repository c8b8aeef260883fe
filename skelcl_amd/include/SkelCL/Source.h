// skelcl::Source (Source.h, src/Source.cpp:51-100): user source text from a
// string or a stream; an empty source is an error.
#pragma once

#include <istream>
#include <iterator>
#include <string>

#include "detail/Runtime.h"

namespace skelcl {

class Source {
public:
  Source() = default;
  Source(const char* s) : _s(s ? s : "") { checkNotEmpty(); }
  Source(const std::string& s) : _s(s) { checkNotEmpty(); }
  Source(std::istream& is)
      : _s(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>())
  {
    checkNotEmpty();
  }
  Source(std::istream&& is)
      : _s(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>())
  {
    checkNotEmpty();
  }
  operator std::string() const { return _s; }
  const std::string& str() const { return _s; }
  void append(const std::string& more) { _s.append("\n" + more); }

private:
  void checkNotEmpty() const
  {
    if (_s.empty())
      detail::abortWithError("Tried to create source object with empty user source.");
  }
  std::string _s;
};

}  // namespace skelcl

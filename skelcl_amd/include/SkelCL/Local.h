// skelcl::Local and skelcl::local (Local.h:58-87, src/Local.cpp): the size
// of a dynamic __local buffer passed to a user function parameter declared
// `__local T* name`.
#pragma once

#include <cstddef>

namespace skelcl {

class Local {
public:
  Local(size_t sizeInBytes) : _bytes(sizeInBytes) {}
  size_t getSizeInBytes() const { return _bytes; }

private:
  size_t _bytes;
};

inline Local local(size_t sizeInBytes) { return Local(sizeInBytes); }

}  // namespace skelcl

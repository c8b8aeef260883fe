// skelcl::Local (Local.h:58-87): size of a dynamic __local buffer passed to a
// user function parameter declared `__local T* name`.
#pragma once

#include <cstddef>

namespace skelcl {

class Local {
public:
  explicit Local(size_t bytes) : _bytes(bytes) {}
  size_t getSizeInBytes() const { return _bytes; }

private:
  size_t _bytes;
};

}  // namespace skelcl

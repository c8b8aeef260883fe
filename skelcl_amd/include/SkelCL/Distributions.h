// skelcl::distribution (Distributions.h:64-311, 365-400).  On one GPU every
// distribution -- Single, Block, Copy, overlap -- is the whole container on
// GPU 0, so these calls only record the request; the multi-GPU layouts are
// not part of this library.
#pragma once

namespace skelcl {
namespace detail {
enum class Padding { NEUTRAL, NEAREST };  // Padding.h:48-51
}

namespace distribution {
template <typename C> void setSingle(const C&) {}
template <typename C, typename D> void setSingle(const C&, const D&) {}
template <typename C> void setBlock(const C&) {}
template <typename C> void setCopy(const C&) {}
template <typename C, typename F> void setCopy(const C&, F) {}
template <typename C> bool isSingle(const C&) { return true; }
template <typename C> bool isBlock(const C&) { return true; }
template <typename C> bool isCopy(const C&) { return true; }
}  // namespace distribution

}  // namespace skelcl

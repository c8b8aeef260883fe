// skelcl::out (Out.h:65-97): marks a container passed to a skeleton as
// written, so its host copy is refreshed on the next host access.
#pragma once

namespace skelcl {

template <typename C>
class Out {
public:
  explicit Out(C& c) : _c(&c) {}
  C& container() const { return *_c; }

private:
  C* _c;
};

template <typename C>
Out<C> out(C& c)
{
  return Out<C>(c);
}

}  // namespace skelcl

// Extension (no counterpart in the reference): the segment sums and the
// segmented scan of one Vector over the same segments from one kernel
// launch (skl_batch_run), optionally with small float products.  float
// segments sum in the Reduce's / Scan's precision form; int segments sum
// exactly modulo 2^32.
#pragma once

#include "Reduce.h"
#include "Scan.h"
#include "Vector.h"

namespace skelcl {

template <typename T>
struct BatchResult {
  Vector<T> sums;  // sums[s] = Reduce(+) of in[offsets[s] .. offsets[s+1])
  Vector<T> scan;  // every segment scanned on its own (outside them: zero)
};

// offsets: nseg + 1 non-decreasing positions inside [0, in.size()]; they are
// checked on the host before the launch.  products / count: device
// descriptors of small float products (skl_gemm_desc, at most 2048) run in
// the same launch at `precision` (SKL_PREC_FP32 or SKL_PREC_BF16X3).
template <typename T>
BatchResult<T> batch(Reduce<T(T)>& reduce, Scan<T(T)>& scan, const Vector<T>& in,
                     const Vector<long>& offsets, bool inclusive = false,
                     const skl_gemm_desc* products = nullptr, int count = 0,
                     int precision = SKL_PREC_FP32)
{
  const int nseg = (int)offsets.size() - 1;
  long prev = 0;
  for (size_t i = 0; i < offsets.size(); ++i) {
    const long o = offsets[i];
    if (o < prev || (size_t)o > in.size())
      detail::check(SKL_ERR_INVALID, "batch: offsets must be non-decreasing inside [0, n]");
    prev = o;
  }
  BatchResult<T> r;
  r.sums.prepareOutput((size_t)(nseg > 0 ? nseg : 0));
  r.scan = Vector<T>(in.size(), T());  // elements outside the segments stay zero
  r.scan.devicePointer();              // uploaded, then overwritten inside the segments
  detail::check(skl_batch_run(reduce.handle(), nseg > 0 ? r.sums.deviceOutputPointer() : nullptr,
                              scan.handle(), r.scan.deviceOutputPointer(), inclusive ? 1 : 0,
                              in.devicePointer(), in.size(),
                              static_cast<const int64_t*>(offsets.devicePointer()), nseg,
                              products, count, precision, 0, nullptr),
                "batch");
  r.sums.dataOnDeviceModified();
  r.scan.dataOnDeviceModified();
  return r;
}

}  // namespace skelcl

// Extensions without a reference counterpart that are not yet available on
// this build report SKL_ERR_UNSUPPORTED.
#include "common.h"
#include "skelcl_skeletons.h"

using namespace skl;

extern "C" int skl_reduce_segmented_run(skl_skeleton, void*, const void*, const int64_t*, int,
                                        skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "segmented Reduce is not available");
}
extern "C" int skl_scan_segmented_run(skl_skeleton, void*, const void*, const int64_t*, int, int,
                                      skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "segmented Scan is not available");
}
extern "C" int skl_map_backward(skl_skeleton, void*, const void*, const void*, size_t,
                                const skl_arg*, int, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "Map backward is not available");
}
extern "C" int skl_zip_backward(skl_skeleton, void*, void*, const void*, const void*,
                                const void*, size_t, const skl_arg*, int, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "Zip backward is not available");
}
extern "C" int skl_reduce_backward(skl_skeleton, void*, const void*, size_t, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "Reduce backward is not available");
}
extern "C" int skl_scan_backward(skl_skeleton, void*, const void*, size_t, int, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "Scan backward is not available");
}
extern "C" int skl_mapoverlap_stencil_backward(skl_skeleton, void*, const void*, const float*,
                                               size_t, size_t, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "MapOverlap backward is not available");
}
extern "C" int skl_allpairs_backward(skl_skeleton, void*, void*, const void*, const void*,
                                     const void*, size_t, size_t, size_t, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "AllPairs backward is not available");
}
extern "C" int skl_matmul_f32(float*, const float*, const float*, size_t, size_t, size_t, int,
                              int, int, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "matmul precision forms are not available");
}
extern "C" int skl_matmul_batched_f32(const skl_gemm_desc*, int, int, skl_stream)
{
  return fail(SKL_ERR_UNSUPPORTED, "batched matmul is not available");
}
extern "C" int skl_autotune(skl_skeleton, size_t, int)
{
  return fail(SKL_ERR_UNSUPPORTED, "autotuning is not available");
}

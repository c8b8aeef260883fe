// Extensions of the skeleton layer that have no reference counterpart:
// precision forms of the dense AllPairs contraction, adjoints, batched
// segmented Reduce/Scan and batched small products.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "common.h"
#include "skeleton.h"
#include "skelcl_skeletons.h"

using namespace skl;

extern "C" size_t skl_internal_matmul_workspace(size_t M, size_t N, size_t K, int precision);
extern "C" int skl_internal_matmul(float* C, const float* A, const float* B, int M, int N,
                                   int K, int ta, int tb, int precision, int tile, void* ws,
                                   hipStream_t s);
extern "C" int skl_internal_stencil_transpose(float* din, const float* dout, const float* w,
                                              int rows, int cols, int R, int nearest,
                                              hipStream_t s);
extern "C" int skl_internal_segmented_sum(float* out, const float* in, unsigned long long n,
                                          const int64_t* offsets, int nseg, int form, void* ws,
                                          hipStream_t s);
extern "C" int skl_internal_segmented_scan(float* out, const float* in, unsigned long long n,
                                           const int64_t* offsets, int nseg, int inclusive,
                                           int form, void* ws, hipStream_t s);
extern "C" size_t skl_internal_segmented_workspace(unsigned long long n, int nseg);
extern "C" int skl_internal_broadcast(void* dst, const void* src, size_t elem, size_t count,
                                      hipStream_t s);
extern "C" size_t skl_internal_batched_workspace(int count);
extern "C" size_t skl_internal_batch_workspace(unsigned long long n, int nseg);
extern "C" int skl_internal_batch_max_products();
extern "C" int skl_internal_batch_status(const void* ws, int* bad);
extern "C" int skl_internal_batch_run(const float* in, float* scan_out, float* sums,
                                      unsigned long long n, const int64_t* offsets, int nseg,
                                      int inclusive, int form, const skl_gemm_desc* descs,
                                      int count, int precision, int workers, void* ws,
                                      hipStream_t s);
extern "C" int skl_internal_batched_gemm_ws(const skl_gemm_desc* descs, int count, int precision,
                                            void* ws, hipStream_t s);

// skeleton helpers implemented in skeletons.cpp
int skl_scan_run_reverse(skl_skeleton s, void* out, const void* in, size_t n, int inclusive,
                         skl_stream stream);
int skl_run_backward_map(skl_skeleton s, void* d_in, const void* in, const void* d_out,
                         size_t n, const skl_arg* args, int nargs, skl_stream stream);
int skl_run_backward_zip(skl_skeleton s, void* d_left, void* d_right, const void* left,
                         const void* right, const void* d_out, size_t n, const skl_arg* args,
                         int nargs, skl_stream stream);

namespace {
// Workspace of the MFMA forms: grows, never shrinks; calls on different
// streams at the same time race on it.  An outgrown buffer is kept, not
// freed: a captured graph may still use it.
std::mutex g_ws_mutex;
void* g_ws = nullptr;
size_t g_ws_bytes = 0;
std::vector<void*> g_ws_retired;

int workspace(size_t bytes, void** out)
{
  std::lock_guard<std::mutex> lk(g_ws_mutex);
  if (g_ws_bytes < bytes) {
    if (g_ws) {
      g_ws_retired.push_back(g_ws);
      g_ws = nullptr;
      g_ws_bytes = 0;
    }
    SKL_HIP_TRY(hipMalloc(&g_ws, bytes));
    g_ws_bytes = bytes;
  }
  *out = g_ws;
  return SKL_OK;
}

// Per-skeleton scratch of the segmented forms, zeroed when allocated (the
// look-back words are tagged with an epoch kept in it).
int seg_scratch(skl_skeleton_s* s, size_t bytes)
{
  if (s->seg_scratch_bytes >= bytes) return SKL_OK;
  if (s->seg_scratch) {  // kept until the skeleton is released (graphs may hold it)
    s->retired.push_back(s->seg_scratch);
    s->seg_scratch = nullptr;
    s->seg_scratch_bytes = 0;
  }
  SKL_HIP_TRY(hipMalloc(&s->seg_scratch, bytes));
  SKL_HIP_TRY(hipMemset(s->seg_scratch, 0, bytes));
  SKL_HIP_TRY(hipDeviceSynchronize());
  s->seg_scratch_bytes = bytes;
  return SKL_OK;
}

// The same for one-launch batches (their own scratch: the ticket and
// error words of the batch kernel must not share words with the
// prologue-based segmented forms).
int batch_scratch(skl_skeleton_s* s, size_t bytes)
{
  if (s->batch_scratch_bytes >= bytes) return SKL_OK;
  if (s->batch_scratch) {
    s->retired.push_back(s->batch_scratch);
    s->batch_scratch = nullptr;
    s->batch_scratch_bytes = 0;
  }
  SKL_HIP_TRY(hipMalloc(&s->batch_scratch, bytes));
  SKL_HIP_TRY(hipMemset(s->batch_scratch, 0, bytes));
  SKL_HIP_TRY(hipDeviceSynchronize());
  s->batch_scratch_bytes = bytes;
  return SKL_OK;
}

// Reduce(+) / Scan(+) without extra arguments on float (1) or on 32-bit
// integers (2, exact modulo 2^32); 0 otherwise
int segment_add_type(const skl_skeleton_s* s, int kind)
{
  if (!s || s->kind != kind || s->op != UserOp::Add || !s->extras.empty()) return 0;
  const std::string& t = s->types[0].name;
  if (t == "float") return 1;
  if (t == "int" || t == "unsigned int" || t == "unsigned" || t == "uint") return 2;
  return 0;
}

bool is_float_addmul(const skl_skeleton_s* s)
{
  return s->kind == SKL_ALLPAIRS && s->op == UserOp::Add && s->op2 == UserOp::Mul &&
         s->extras.empty() && s->types[0].name == "float" && s->types[1].name == "float" &&
         s->types[2].name == "float";
}
}  // namespace

int skl::matmul_run(float* C, const float* A, const float* B, size_t M, size_t N, size_t K,
                    int trans_a, int trans_b, int precision, int tile, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (M == 0 || N == 0) return SKL_OK;
  if (!C || !A || !B || K == 0) return fail(SKL_ERR_INVALID, "matmul: NULL buffer or K == 0");
  if (M > 0x7FFFFF80 || N > 0x7FFFFF80 || K > 0x7FFFFF80)
    return fail(SKL_ERR_INVALID, "matmul: dimension too large");
  if (precision < SKL_PREC_DEFAULT || precision > SKL_PREC_FP8)
    return fail(SKL_ERR_INVALID, "matmul: unknown precision %d", precision);
  if (precision == SKL_PREC_DEFAULT) precision = SKL_PREC_FP32;
  void* ws = nullptr;
  size_t wsb = skl_internal_matmul_workspace(M, N, K, precision);
  if (wsb) SKL_TRY(workspace(wsb, &ws));
  int st = skl_internal_matmul(C, A, B, (int)M, (int)N, (int)K, trans_a ? 1 : 0,
                               trans_b ? 1 : 0, precision, tile, ws, resolve_stream(stream));
  if (st != SKL_OK) return fail(st, "matmul launch failed: %s", hipGetErrorString(hipGetLastError()));
  return SKL_OK;
}

extern "C" int skl_matmul_f32(float* C, const float* A, const float* B, size_t M, size_t N,
                              size_t K, int trans_a, int trans_b, int precision,
                              skl_stream stream)
{
  return skl::matmul_run(C, A, B, M, N, K, trans_a, trans_b, precision, 0, stream);
}

extern "C" int skl_allpairs_backward(skl_skeleton s, void* d_left, void* d_right,
                                     const void* left, const void* right, const void* d_out,
                                     size_t rows, size_t cols, size_t dim, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!s) return fail(SKL_ERR_INVALID, "skeleton is NULL");
  if (!is_float_addmul(s))
    return fail(SKL_ERR_UNSUPPORTED,
                "AllPairs backward needs the (+,*) contraction on float (a matmul)");
  if (!d_out) return fail(SKL_ERR_INVALID, "d_out is NULL");
  // FP64 (a Reduce/Scan form) runs the products as DEFAULT: plain FP32
  int prec = s->precision == SKL_PREC_DEFAULT || s->precision == SKL_PREC_FP64 ? SKL_PREC_FP32
                                                                                : s->precision;
  // C = A.B with A rows x dim, B dim x cols; the tuned tile is the forward
  // product's, so the adjoint products take the size rule
  if (d_left)  // dA = dC . B^T : (rows x cols) . (cols x dim)
    SKL_TRY(skl_matmul_f32((float*)d_left, (const float*)d_out, (const float*)right, rows, dim,
                           cols, 0, 1, prec, stream));
  if (d_right)  // dB = A^T . dC : (dim x rows) . (rows x cols)
    SKL_TRY(skl_matmul_f32((float*)d_right, (const float*)left, (const float*)d_out, dim, cols,
                           rows, 1, 0, prec, stream));
  return SKL_OK;
}

extern "C" int skl_reduce_backward(skl_skeleton s, void* d_in, const void* d_out, size_t n,
                                   skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!s || s->kind != SKL_REDUCE) return fail(SKL_ERR_INVALID, "not a Reduce skeleton");
  if (s->op != UserOp::Add || !s->extras.empty())
    return fail(SKL_ERR_UNSUPPORTED, "Reduce backward is defined for Reduce(+)");
  if (n == 0) return SKL_OK;
  if (!d_in || !d_out) return fail(SKL_ERR_INVALID, "NULL buffer");
  // d in_i = d out (broadcast): copy the scalar into place 0 then fill by a
  // kernel that reads it on the device (no host round trip)
  return skl_internal_broadcast(d_in, d_out, s->types[0].size, n, resolve_stream(stream));
}

extern "C" int skl_scan_backward(skl_skeleton s, void* d_in, const void* d_out, size_t n,
                                 int inclusive, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!s || s->kind != SKL_SCAN) return fail(SKL_ERR_INVALID, "not a Scan skeleton");
  if (s->op != UserOp::Add || !s->extras.empty())
    return fail(SKL_ERR_UNSUPPORTED, "Scan backward is defined for Scan(+)");
  // y = scan(x): dx_i = sum_{j>i} dy_j (exclusive) / sum_{j>=i} dy_j
  // (inclusive): the scan of dy read from the end
  return skl_scan_run_reverse(s, d_in, d_out, n, inclusive, stream);
}

extern "C" int skl_map_backward(skl_skeleton s, void* d_in, const void* in, const void* d_out,
                                size_t n, const skl_arg* args, int nargs, skl_stream stream)
{
  return skl_run_backward_map(s, d_in, in, d_out, n, args, nargs, stream);
}

extern "C" int skl_zip_backward(skl_skeleton s, void* d_left, void* d_right, const void* left,
                                const void* right, const void* d_out, size_t n,
                                const skl_arg* args, int nargs, skl_stream stream)
{
  return skl_run_backward_zip(s, d_left, d_right, left, right, d_out, n, args, nargs, stream);
}

// The transposed stencil is the adjoint only when the forward user function
// is out = sum_{dy,dx} w[(dy+R)(2R+1) + dx+R] * getData(img, dx, dy) over the
// weight table it takes as its one extra argument.  That is checked once per
// skeleton by running the forward on a small seeded image (border pixels
// included, rows != cols) with seeded asymmetric weights and comparing with
// that sum on the host; a transposed weight index, a missing term or a
// non-linear term fails the check.  The probe allocates and synchronises,
// so a capture of the adjoint needs one eager call first.
namespace {
int check_weighted_stencil(skl_skeleton_s* s, skl_stream stream)
{
  if (s->stencil_checked == 1) return SKL_OK;
  auto reject = [&](const char* why) {
    s->stencil_checked = -1;
    return fail(SKL_ERR_UNSUPPORTED,
                "stencil backward: the MapOverlap user function is not the weighted stencil "
                "sum w[(dy+R)*(2R+1)+dx+R] * getData(img, dx, dy) (%s)",
                why);
  };
  if (s->stencil_checked == -1) return reject("checked before");
  if (s->extras.size() != 1 || s->extras[0].kind != ParamKind::Buffer)
    return reject("it must take the weight table as its only extra argument");
  const int R = s->overlap, W = 2 * R + 1;
  const int rows = 2 * R + 3, cols = 2 * R + 6;
  std::vector<float> img((size_t)rows * cols), w((size_t)W * W), out((size_t)rows * cols);
  uint32_t z = 0x9E3779B9u;
  auto next = [&]() {
    z = z * 1664525u + 1013904223u;
    return (float)((z >> 8) & 0xFFFF) / 65536.0f;
  };
  for (auto& v : img) v = 4.0f * next() - 2.0f;
  for (auto& v : w) v = 0.25f + next();
  void *dimg = nullptr, *dw = nullptr, *dout = nullptr;
  int st = SKL_OK;
  if (hipMalloc(&dimg, img.size() * 4) != hipSuccess || hipMalloc(&dw, w.size() * 4) != hipSuccess ||
      hipMalloc(&dout, out.size() * 4) != hipSuccess) {
    st = fail(SKL_ERR_HIP, "stencil backward: probe allocation failed");
  }
  hipStream_t hs = resolve_stream(stream);
  if (st == SKL_OK &&
      (hipMemcpyAsync(dimg, img.data(), img.size() * 4, hipMemcpyHostToDevice, hs) != hipSuccess ||
       hipMemcpyAsync(dw, w.data(), w.size() * 4, hipMemcpyHostToDevice, hs) != hipSuccess))
    st = fail(SKL_ERR_HIP, "stencil backward: probe upload failed");
  if (st == SKL_OK) {
    skl_arg a{SKL_ARG_BUFFER, dw, 0, 0};
    st = skl_mapoverlap_run(s, dout, dimg, rows, cols, &a, 1, stream);
  }
  if (st == SKL_OK &&
      (hipMemcpyAsync(out.data(), dout, out.size() * 4, hipMemcpyDeviceToHost, hs) != hipSuccess ||
       hipStreamSynchronize(hs) != hipSuccess))
    st = fail(SKL_ERR_HIP, "stencil backward: probe download failed");
  if (dimg) hipFree(dimg);
  if (dw) hipFree(dw);
  if (dout) hipFree(dout);
  if (st != SKL_OK) return st;
  const bool nearest = s->padding == SKL_PADDING_NEAREST;
  double wabs = 0.0;
  for (float v : w) wabs += std::fabs((double)v);
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      double ref = 0.0;
      for (int dy = -R; dy <= R; ++dy)
        for (int dx = -R; dx <= R; ++dx) {
          int rr = r + dy, cc = c + dx;
          double v = 0.0;
          if (nearest) {
            rr = rr < 0 ? 0 : (rr >= rows ? rows - 1 : rr);
            cc = cc < 0 ? 0 : (cc >= cols ? cols - 1 : cc);
            v = img[(size_t)rr * cols + cc];
          } else if (rr >= 0 && rr < rows && cc >= 0 && cc < cols) {
            v = img[(size_t)rr * cols + cc];
          }
          ref += v * (double)w[(size_t)(dy + R) * W + (dx + R)];
        }
      const double got = out[(size_t)r * cols + c];
      if (!(std::fabs(got - ref) <= 1e-4 * 2.0 * wabs))
        return reject("the probe's forward differs from the weighted sum");
    }
  s->stencil_checked = 1;
  return SKL_OK;
}
}  // namespace

extern "C" int skl_mapoverlap_stencil_backward(skl_skeleton s, void* d_in, const void* d_out,
                                               const float* weights, size_t rows, size_t cols,
                                               skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!s || s->kind != SKL_MAPOVERLAP) return fail(SKL_ERR_INVALID, "not a MapOverlap skeleton");
  if (s->types[0].name != "float" || s->types[1].name != "float")
    return fail(SKL_ERR_UNSUPPORTED, "stencil backward is defined for float MapOverlap");
  if (s->padding == SKL_PADDING_NEUTRAL && s->neutral != "0" && s->neutral != "0.0f" &&
      s->neutral != "0.0" && s->neutral != "0.f")
    return fail(SKL_ERR_UNSUPPORTED, "stencil backward needs a zero neutral value");
  if (rows == 0 || cols == 0) return SKL_OK;
  if (!d_in || !d_out || !weights) return fail(SKL_ERR_INVALID, "NULL buffer");
  if (rows > 0x7FFFFFFF || cols > 0x7FFFFFFF) return fail(SKL_ERR_INVALID, "matrix too large");
  SKL_TRY(check_weighted_stencil(s, stream));
  return skl_internal_stencil_transpose((float*)d_in, (const float*)d_out, weights, (int)rows,
                                        (int)cols, s->overlap,
                                        s->padding == SKL_PADDING_NEAREST ? 1 : 0,
                                        resolve_stream(stream));
}

extern "C" int skl_reduce_segmented_run(skl_skeleton s, void* out, const void* in, size_t n,
                                        const int64_t* offsets, int nseg, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!s || s->kind != SKL_REDUCE) return fail(SKL_ERR_INVALID, "not a Reduce skeleton");
  const int ty = segment_add_type(s, SKL_REDUCE);
  if (!ty) return fail(SKL_ERR_UNSUPPORTED, "segmented Reduce is available for float or int Reduce(+)");
  if (nseg < 0) return fail(SKL_ERR_INVALID, "negative segment count");
  if (nseg == 0) return SKL_OK;
  if (!out || !offsets || (n && !in)) return fail(SKL_ERR_INVALID, "NULL buffer");
  SKL_TRY(seg_scratch(s, skl_internal_segmented_workspace(n, nseg)));
  // int: exact modulo 2^32; float precision FP64: FP64 chunk sums;
  // otherwise compensated FP32 in a chunk
  return skl_internal_segmented_sum((float*)out, (const float*)in, n, offsets, nseg,
                                    ty == 2 ? 2 : s->precision == SKL_PREC_FP64 ? 1 : 0,
                                    s->seg_scratch, resolve_stream(stream));
}

extern "C" int skl_scan_segmented_run(skl_skeleton s, void* out, const void* in, size_t n,
                                      const int64_t* offsets, int nseg, int inclusive,
                                      skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (!s || s->kind != SKL_SCAN) return fail(SKL_ERR_INVALID, "not a Scan skeleton");
  const int ty = segment_add_type(s, SKL_SCAN);
  if (!ty) return fail(SKL_ERR_UNSUPPORTED, "segmented Scan is available for float or int Scan(+)");
  if (nseg < 0) return fail(SKL_ERR_INVALID, "negative segment count");
  if (nseg == 0 || n == 0) return SKL_OK;
  if (!out || !in || !offsets) return fail(SKL_ERR_INVALID, "NULL buffer");
  SKL_TRY(seg_scratch(s, skl_internal_segmented_workspace(n, nseg)));
  return skl_internal_segmented_scan((float*)out, (const float*)in, n, offsets, nseg,
                                     inclusive, ty == 2 ? 2 : s->precision == SKL_PREC_FP64 ? 1 : 0,
                                     s->seg_scratch, resolve_stream(stream));
}

extern "C" int skl_matmul_batched_f32(const skl_gemm_desc* descs_device, int count,
                                      int precision, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (count <= 0) return SKL_OK;
  if (!descs_device) return fail(SKL_ERR_INVALID, "descs is NULL");
  if (precision != SKL_PREC_DEFAULT && precision != SKL_PREC_FP32 &&
      precision != SKL_PREC_BF16X3)
    return fail(SKL_ERR_UNSUPPORTED, "batched matmul supports FP32 and BF16X3");
  void* ws = nullptr;
  SKL_TRY(workspace(skl_internal_batched_workspace(count), &ws));
  return skl_internal_batched_gemm_ws(descs_device, count, precision, ws, resolve_stream(stream));
}

extern "C" int skl_batch_run(skl_skeleton reduce, void* sums, skl_skeleton scan, void* scan_out,
                             int inclusive, const void* in, size_t n, const int64_t* offsets,
                             int nseg, const skl_gemm_desc* descs, int count, int precision,
                             int workers_per_cu, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  if (nseg < 0 || count < 0) return fail(SKL_ERR_INVALID, "negative segment or product count");
  if (!sums) reduce = nullptr;
  if (!scan_out) scan = nullptr;
  const int tr = reduce ? segment_add_type(reduce, SKL_REDUCE) : 0;
  const int ts = scan ? segment_add_type(scan, SKL_SCAN) : 0;
  if (reduce && !tr)
    return fail(SKL_ERR_UNSUPPORTED, "batch: segment sums need a float or int Reduce(+)");
  if (scan && !ts)
    return fail(SKL_ERR_UNSUPPORTED, "batch: the segmented scan needs a float or int Scan(+)");
  if (reduce && scan && tr != ts)
    return fail(SKL_ERR_INVALID, "batch: the Reduce and the Scan must share an element type");
  const bool segs = nseg > 0 && (reduce || scan);
  const bool ints = (tr ? tr : ts) == 2;
  const int fp64r = reduce && reduce->precision == SKL_PREC_FP64;
  const int fp64s = scan && scan->precision == SKL_PREC_FP64;
  if (!ints && reduce && scan && fp64r != fp64s)
    return fail(SKL_ERR_INVALID, "batch: the Reduce and the Scan must share a precision form");
  if (segs && (!offsets || (n && !in)))
    return fail(SKL_ERR_INVALID, "batch: NULL input or offsets");
  if (count > skl_internal_batch_max_products())
    return fail(SKL_ERR_INVALID, "batch: at most %d products per launch",
                skl_internal_batch_max_products());
  if (count > 0 && !descs) return fail(SKL_ERR_INVALID, "batch: descs is NULL");
  if (precision == SKL_PREC_DEFAULT) precision = SKL_PREC_FP32;
  if (count > 0 && precision != SKL_PREC_FP32 && precision != SKL_PREC_BF16X3)
    return fail(SKL_ERR_UNSUPPORTED, "batch: products run in FP32 or BF16X3");
  if (workers_per_cu < 0 || workers_per_cu > 2)
    return fail(SKL_ERR_INVALID, "batch: product work-groups per CU must be 0, 1 or 2");
  if (!segs && count == 0) return SKL_OK;
  void* ws = nullptr;
  if (segs) {
    skl_skeleton_s* owner = scan ? scan : reduce;
    SKL_TRY(batch_scratch(owner, skl_internal_batch_workspace(n, nseg)));
    ws = owner->batch_scratch;
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  // default two product work-groups per CU: the batched workload's first 64
  // segments and products 0.40 against 0.48 ms with one, all of it 3.26
  // against 3.36 ms (the first 8: 0.119 against 0.118)
  const int workers = count > 0 ? cus * (workers_per_cu ? workers_per_cu : 2) : 0;
  // no elements: the tile reads nothing, but its loads need a valid address
  const float* src = static_cast<const float*>(n ? in : ws);
  const int st = skl_internal_batch_run(src, segs ? static_cast<float*>(scan_out) : nullptr,
                                        segs ? static_cast<float*>(sums) : nullptr, n,
                                        offsets, segs ? nseg : 0, inclusive,
                                        ints ? 2 : (fp64r || fp64s ? 1 : 0), descs, count,
                                        precision, workers, ws, resolve_stream(stream));
  if (st != SKL_OK) return fail(st, "batch launch failed");
  return SKL_OK;
}

extern "C" int skl_batch_check(skl_skeleton s, int* invalid_offsets)
{
  SKL_TRY(ensure_initialized());
  if (!s || !invalid_offsets) return fail(SKL_ERR_INVALID, "NULL argument");
  *invalid_offsets = 0;
  if (!s->batch_scratch) return SKL_OK;
  SKL_HIP_TRY(hipDeviceSynchronize());
  return skl_internal_batch_status(s->batch_scratch, invalid_offsets);
}

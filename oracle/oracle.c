/*
 * oracle.c -- CPU restatement of SkelCL's skeleton algorithms (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the parity reference and by
 * bench.py as the CPU baseline.  The product never links it.
 *
 * Built with -ffp-contract=off so every float expression rounds exactly as
 * written (no fused multiply-add), which is the evaluation the reference's
 * OpenCL kernels would give with contraction disabled.
 *
 * Pinning -- what checks each function, besides its own formula
 * (tests/test_oracle.py):
 *   orc_splitmix64 / orc_gen_*     published SplitMix64 outputs (seed
 *                                  1234567), exact-grid and range invariants,
 *                                  committed first-64 fixtures per seed.
 *   orc_reduce_add_*_reforder      reference KATs ReduceTests.cpp:77-170
 *                                  (4950, 1258491, 1e8 ones, sizes 1..255 and
 *                                  8192..8447); brute-force sequential fold
 *                                  on exactly-representable inputs; error vs
 *                                  orc_sum_f64 bounded by the fp32 pairwise
 *                                  bound; int result equals numpy sum mod
 *                                  2^32.
 *   orc_scan_excl_add_*_reforder   ScanTests.cpp:71-107 KATs (ones -> i);
 *                                  equality with orc_scan_add_u32 (closed
 *                                  form mod 2^32) over many sizes across
 *                                  1..4 passes; float version exact on
 *                                  integer-valued inputs, within bound vs
 *                                  FP64 otherwise.
 *   orc_scan_add_f64 / _adjoint    numpy cumsum; adjoint against FP64 central
 *                                  finite differences of sum(g*scan(x)).
 *   orc_mapoverlap_stencil_f32_fma the same chain with every step rounded
 *                                  exactly (Python Fractions, round to
 *                                  nearest even f32) on random inputs.
 *   orc_saxpy_f32 / _fma, orc_scale_f32, orc_zip_mul_f32
 *                                  numpy float32 arithmetic (two roundings)
 *                                  and exactly rounded a*x+y (Fractions).
 *   orc_gen_loguniform             the committed batched_shapes.json segment
 *                                  lengths (written from Python integers).
 *   orc_mapoverlap_*               MapOverlapTests.cpp:187-447 shift KATs
 *                                  (through orc_getdata_i32); a delta-image
 *                                  reproduces the flipped weight table; the
 *                                  stencil equals scipy.signal.correlate2d
 *                                  (zero fill) / numpy edge-padded
 *                                  correlation; adjoint vs FD and vs
 *                                  <W, A x> = <A^T W, x>.
 *   orc_allpairs_*                 AllPairsTests.cpp:77-133 integer-valued
 *                                  KATs (64^3, 32x128x32, 100x60x11) vs
 *                                  numpy matmul; sqdist vs
 *                                  |a|^2+|b|^2-2ab in FP64.
 *   orc_segmented_*                numpy per-segment sum/cumsum; the u32
 *                                  forms also against orc_scan_add_u32 per
 *                                  segment and a Python loop mod 2^32.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* generic pthread fan-out                                                  */
/* ------------------------------------------------------------------------ */
typedef void (*orc_task_fn)(void* ctx, size_t begin, size_t end);
typedef struct {
  orc_task_fn fn;
  void* ctx;
  size_t begin, end;
} orc_task;

static void* orc_task_tramp(void* p)
{
  orc_task* t = (orc_task*)p;
  t->fn(t->ctx, t->begin, t->end);
  return NULL;
}

static void orc_parallel_for(size_t n, int nthreads, orc_task_fn fn, void* ctx)
{
  if (nthreads <= 1 || n < 2) {
    fn(ctx, 0, n);
    return;
  }
  if ((size_t)nthreads > n) nthreads = (int)n;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  orc_task* tk = (orc_task*)malloc(sizeof(orc_task) * (size_t)nthreads);
  size_t chunk = (n + (size_t)nthreads - 1) / (size_t)nthreads;
  int started = 0;
  for (int t = 0; t < nthreads; ++t) {
    size_t b = (size_t)t * chunk;
    size_t e = b + chunk < n ? b + chunk : n;
    if (b >= e) break;
    tk[t].fn = fn;
    tk[t].ctx = ctx;
    tk[t].begin = b;
    tk[t].end = e;
    if (pthread_create(&th[t], NULL, orc_task_tramp, &tk[t]) != 0) {
      fn(ctx, b, e); /* fall back to running inline */
      tk[t].fn = NULL;
    }
    started = t + 1;
  }
  for (int t = 0; t < started; ++t)
    if (tk[t].fn) pthread_join(th[t], NULL);
  free(th);
  free(tk);
}

/* ------------------------------------------------------------------------ */
/* generator                                                                */
/* ------------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t seed, uint64_t i)
{
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_gen_u64(uint64_t seed, size_t n, uint64_t* out)
{
  for (size_t i = 0; i < n; ++i) out[i] = orc_splitmix64(seed, i);
}

void orc_gen_f32_sym(uint64_t seed, size_t n, float* out)
{
  for (size_t i = 0; i < n; ++i) {
    float u = (float)(orc_splitmix64(seed, i) >> 40) * 0x1p-24f;
    out[i] = 2.0f * u - 1.0f;
  }
}

void orc_gen_f32_unit(uint64_t seed, size_t n, float scale, float* out)
{
  for (size_t i = 0; i < n; ++i) {
    float u = (float)(orc_splitmix64(seed, i) >> 40) * 0x1p-24f;
    out[i] = scale * u;
  }
}

void orc_gen_i32_below(uint64_t seed, size_t n, uint32_t m, int32_t* out)
{
  for (size_t i = 0; i < n; ++i) {
    uint64_t hi = orc_splitmix64(seed, i) >> 32;
    out[i] = (int32_t)((hi * (uint64_t)m) >> 32);
  }
}

void orc_gen_loguniform(uint64_t seed, size_t n, double e0, double span,
                        int64_t* out)
{
  for (size_t i = 0; i < n; ++i) {
    double d = (double)(orc_splitmix64(seed, i) >> 11) * 0x1p-53;
    out[i] = (int64_t)floor(pow(10.0, e0 + span * d));
  }
}

/* ------------------------------------------------------------------------ */
/* Map / Zip                                                                */
/* examples/saxpy/main.cpp:77-88 (Zip "a*x+y"), MapDef.h:182-196 (SCL_MAP), */
/* ZipDef.h:191-207 (SCL_ZIP): one independent evaluation per element.      */
/* ------------------------------------------------------------------------ */
typedef struct {
  float a;
  const float *x, *y;
  float* out;
} saxpy_ctx;

static void saxpy_task(void* p, size_t b, size_t e)
{
  saxpy_ctx* c = (saxpy_ctx*)p;
  for (size_t i = b; i < e; ++i) c->out[i] = c->a * c->x[i] + c->y[i];
}

void orc_saxpy_f32(float a, const float* x, const float* y, float* out,
                   size_t n, int nthreads)
{
  saxpy_ctx c = {a, x, y, out};
  orc_parallel_for(n, nthreads, saxpy_task, &c);
}

/* the contracted form fma(a, x, y): one rounding, as a compiler that fuses
 * a*x+y within the expression evaluates the saxpy user function */
void orc_saxpy_f32_fma(float a, const float* x, const float* y, float* out,
                       size_t n)
{
  for (size_t i = 0; i < n; ++i) out[i] = fmaf(a, x[i], y[i]);
}

void orc_scale_f32(float a, const float* x, float* out, size_t n)
{
  for (size_t i = 0; i < n; ++i) out[i] = a * x[i];
}

void orc_zip_mul_f32(const float* x, const float* y, float* out, size_t n)
{
  for (size_t i = 0; i < n; ++i) out[i] = x[i] * y[i];
}

/* ------------------------------------------------------------------------ */
/* Reduce: ReduceDef.h:96-127 (8192 partials, then one work-group),         */
/* ReduceKernel.cl:51-72 (SCL_REDUCE_1: strided fold, stride GLOBAL_SIZE),  */
/* ReduceKernel.cl:78-128 (SCL_REDUCE_2: strided fold to LOCAL_SIZE, then   */
/* the halving tree buf[i] = f(buf[i], buf[i + ceil(n/2)]), i < floor(n/2)), */
/* ReduceDef.h:205-245 (LOCAL_SIZE = min(data, kernel work-group size)).    */
/* ------------------------------------------------------------------------ */
#define ORC_REDUCE_GLOBAL 8192u

typedef struct {
  const float* in;
  size_t n;
  float* part;
} red1_ctx;

static void red1_task(void* p, size_t b, size_t e)
{
  red1_ctx* c = (red1_ctx*)p;
  for (size_t g = b; g < e; ++g) {
    float res = c->in[g];
    for (size_t i = g + ORC_REDUCE_GLOBAL; i < c->n; i += ORC_REDUCE_GLOBAL)
      res = res + c->in[i];
    c->part[g] = res;
  }
}

float orc_reduce_add_f32_reforder(const float* in, size_t n, size_t wg,
                                  int nthreads)
{
  if (n == 0) return 0.0f;
  size_t D = n < ORC_REDUCE_GLOBAL ? n : ORC_REDUCE_GLOBAL;
  float* p = (float*)malloc(sizeof(float) * D);
  red1_ctx c = {in, n, p};
  orc_parallel_for(D, nthreads, red1_task, &c);
  size_t L = D < wg ? D : wg;
  float* b = (float*)malloc(sizeof(float) * L);
  for (size_t t = 0; t < L; ++t) {
    float res = p[t];
    for (size_t i = t + L; i < D; i += L) res = res + p[i];
    b[t] = res;
  }
  size_t m = L;
  while (m > 1) {
    size_t modul = (m + 1) / 2, active = m / 2;
    for (size_t i = 0; i < active; ++i) b[i] = b[i] + b[i + modul];
    m = (m + 1) / 2;
  }
  float r = b[0];
  free(p);
  free(b);
  return r;
}

uint32_t orc_reduce_add_u32_reforder(const uint32_t* in, size_t n, size_t wg)
{
  if (n == 0) return 0u;
  size_t D = n < ORC_REDUCE_GLOBAL ? n : ORC_REDUCE_GLOBAL;
  uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * D);
  for (size_t g = 0; g < D; ++g) {
    uint32_t res = in[g];
    for (size_t i = g + ORC_REDUCE_GLOBAL; i < n; i += ORC_REDUCE_GLOBAL)
      res += in[i];
    p[g] = res;
  }
  size_t L = D < wg ? D : wg;
  uint32_t* b = (uint32_t*)malloc(sizeof(uint32_t) * L);
  for (size_t t = 0; t < L; ++t) {
    uint32_t res = p[t];
    for (size_t i = t + L; i < D; i += L) res += p[i];
    b[t] = res;
  }
  size_t m = L;
  while (m > 1) {
    size_t modul = (m + 1) / 2, active = m / 2;
    for (size_t i = 0; i < active; ++i) b[i] += b[i + modul];
    m = (m + 1) / 2;
  }
  uint32_t r = b[0];
  free(p);
  free(b);
  return r;
}

/* Neumaier-compensated sum in double: an FP64 reference value whose own
 * error is far below any fp32 result's. */
double orc_sum_f64(const float* in, size_t n)
{
  double s = 0.0, c = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double x = (double)in[i];
    double t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  return s + c;
}

double orc_dot_f64(const float* x, const float* y, size_t n)
{
  double s = 0.0, c = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double v = (double)x[i] * (double)y[i]; /* exact: 24+24 bits < 53 */
    double t = s + v;
    if (fabs(s) >= fabs(v))
      c += (s - t) + v;
    else
      c += (v - t) + s;
    s = t;
  }
  return s + c;
}

/* ------------------------------------------------------------------------ */
/* Scan: ScanDef.h:108-259 (passes = ceil(log_wg n); level buffers n,       */
/* ceil(n/wg), ...; uniform combination from the top level down) and        */
/* ScanKernel.cl:74-196 (exclusive Blelloch scan of wg elements per work-   */
/* group: up-sweep buf[bi] = f(buf[bi], buf[ai]); block total kept; last    */
/* element cleared to the identity; down-sweep tmp = buf[ai]; buf[ai] =     */
/* buf[bi]; buf[bi] = f(buf[bi], tmp)), ScanKernel.cl:206-250 (uniform add  */
/* out = f(out, blockSums[block])).                                         */
/* ------------------------------------------------------------------------ */
#define ORC_DEFINE_BLELLOCH(NAME, T, ADD)                                      \
  static void NAME(T* buf, size_t L, T id, T* total)                           \
  {                                                                            \
    size_t offset = 1;                                                         \
    for (size_t d = L >> 1; d > 0; d >>= 1) {                                  \
      for (size_t tid = 0; tid < d; ++tid) {                                   \
        size_t ai = offset * (2 * tid + 1) - 1;                                \
        size_t bi = offset * (2 * tid + 2) - 1;                                \
        buf[bi] = ADD(buf[bi], buf[ai]);                                       \
      }                                                                        \
      offset <<= 1;                                                            \
    }                                                                          \
    *total = buf[L - 1];                                                       \
    buf[L - 1] = id;                                                           \
    for (size_t d = 1; d < L; d <<= 1) {                                       \
      offset >>= 1;                                                            \
      for (size_t tid = 0; tid < d; ++tid) {                                   \
        size_t ai = offset * (2 * tid + 1) - 1;                                \
        size_t bi = offset * (2 * tid + 2) - 1;                                \
        T tmp = buf[ai];                                                       \
        buf[ai] = buf[bi];                                                     \
        buf[bi] = ADD(buf[bi], tmp);                                           \
      }                                                                        \
    }                                                                          \
  }

#define ORC_ADD_F(a, b) ((a) + (b))
#define ORC_ADD_U(a, b) ((uint32_t)((a) + (b)))
ORC_DEFINE_BLELLOCH(blelloch_f32, float, ORC_ADD_F)
ORC_DEFINE_BLELLOCH(blelloch_u32, uint32_t, ORC_ADD_U)

#define ORC_DEFINE_SCAN_LEVEL(NAME, BLL, T)                                    \
  typedef struct {                                                             \
    const T* in;                                                               \
    T* out;                                                                    \
    T* sums;                                                                   \
    size_t n, L;                                                               \
  } NAME##_ctx;                                                                \
  static void NAME##_task(void* p, size_t b, size_t e)                         \
  {                                                                            \
    NAME##_ctx* c = (NAME##_ctx*)p;                                            \
    T* buf = (T*)malloc(sizeof(T) * c->L);                                     \
    for (size_t blk = b; blk < e; ++blk) {                                     \
      size_t base = blk * c->L;                                                \
      for (size_t j = 0; j < c->L; ++j)                                        \
        buf[j] = base + j < c->n ? c->in[base + j] : (T)0;                     \
      T tot;                                                                   \
      BLL(buf, c->L, (T)0, &tot);                                              \
      c->sums[blk] = tot;                                                      \
      for (size_t j = 0; j < c->L && base + j < c->n; ++j)                     \
        c->out[base + j] = buf[j];                                             \
    }                                                                          \
    free(buf);                                                                 \
  }                                                                            \
  typedef struct {                                                             \
    T* out;                                                                    \
    const T* sums;                                                             \
    size_t n, L;                                                               \
  } NAME##_uc_ctx;                                                             \
  static void NAME##_uc_task(void* p, size_t b, size_t e)                      \
  {                                                                            \
    NAME##_uc_ctx* c = (NAME##_uc_ctx*)p;                                      \
    for (size_t blk = b; blk < e; ++blk) {                                     \
      size_t base = blk * c->L;                                                \
      for (size_t j = 0; j < c->L && base + j < c->n; ++j)                     \
        c->out[base + j] = c->out[base + j] + c->sums[blk];                    \
    }                                                                          \
  }                                                                            \
  static void NAME(const T* in, T* out, size_t n, size_t L, int nthreads)      \
  {                                                                            \
    size_t nb = (n + L - 1) / L;                                               \
    T* sums = (T*)malloc(sizeof(T) * (nb ? nb : 1));                           \
    NAME##_ctx c = {in, out, sums, n, L};                                      \
    orc_parallel_for(nb, nthreads, NAME##_task, &c);                           \
    if (nb > 1) {                                                              \
      NAME(sums, sums, nb, L, nthreads); /* scan block sums in place */        \
      NAME##_uc_ctx u = {out, sums, n, L};                                     \
      orc_parallel_for(nb, nthreads, NAME##_uc_task, &u);                      \
    }                                                                          \
    free(sums);                                                                \
  }

ORC_DEFINE_SCAN_LEVEL(scan_level_f32, blelloch_f32, float)
ORC_DEFINE_SCAN_LEVEL(scan_level_u32, blelloch_u32, uint32_t)

void orc_scan_excl_add_f32_reforder(const float* in, float* out, size_t n,
                                    size_t wg, int nthreads)
{
  if (n == 0) return;
  scan_level_f32(in, out, n, wg, nthreads);
}

void orc_scan_excl_add_u32_reforder(const uint32_t* in, uint32_t* out,
                                    size_t n, size_t wg, int nthreads)
{
  if (n == 0) return;
  scan_level_u32(in, out, n, wg, nthreads);
}

void orc_scan_add_u32(const uint32_t* in, uint32_t* out, size_t n,
                      int inclusive)
{
  uint32_t acc = 0;
  for (size_t i = 0; i < n; ++i) {
    uint32_t x = in[i];
    if (inclusive) {
      acc += x;
      out[i] = acc;
    } else {
      out[i] = acc;
      acc += x;
    }
  }
}

void orc_scan_add_f64(const float* in, double* out, size_t n, int inclusive)
{
  double s = 0.0, c = 0.0;
  for (size_t i = 0; i < n; ++i) {
    if (!inclusive) out[i] = s + c;
    double x = (double)in[i];
    double t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
    if (inclusive) out[i] = s + c;
  }
}

/* exclusive: dx_i = sum_{j>i} g_j ; inclusive: dx_i = sum_{j>=i} g_j */
void orc_scan_add_adjoint_f64(const float* g, double* dx, size_t n,
                              int inclusive)
{
  double s = 0.0;
  for (size_t k = n; k-- > 0;) {
    if (inclusive) {
      s += (double)g[k];
      dx[k] = s;
    } else {
      dx[k] = s;
      s += (double)g[k];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* MapOverlap: getData(m, x, y) reads the element x columns right and y     */
/* rows down of the centre (MapOverlapDef.h:216-222).  Rows outside the     */
/* matrix come from the padding rows built by OLDistributionDef.h:263-343   */
/* (NEUTRAL: filled with the neutral value; NEAREST: copies of the edge     */
/* row); columns outside come from MapOverlapKernel.cl:96-141 (NEUTRAL, or  */
/* the edge column of the same row).  The stencil user function sums        */
/* getData(img, dx, dy) * w[(dy+R)*(2R+1) + (dx+R)] with dy outer, dx inner.*/
/* ------------------------------------------------------------------------ */
static inline int clampi(int v, int lo, int hi)
{
  return v < lo ? lo : (v > hi ? hi : v);
}

int32_t orc_getdata_i32(const int32_t* in, int rows, int cols, int r, int c,
                        int dx, int dy, int padding, int32_t neutral)
{
  int rr = r + dy, cc = c + dx;
  if (rr < 0 || rr >= rows || cc < 0 || cc >= cols) {
    if (padding == 0) return neutral;
    rr = clampi(rr, 0, rows - 1);
    cc = clampi(cc, 0, cols - 1);
  }
  return in[(size_t)rr * (size_t)cols + (size_t)cc];
}

static inline float getdata_f32(const float* in, int rows, int cols, int r,
                                int c, int dx, int dy, int padding,
                                float neutral)
{
  int rr = r + dy, cc = c + dx;
  if (rr < 0 || rr >= rows || cc < 0 || cc >= cols) {
    if (padding == 0) return neutral;
    rr = clampi(rr, 0, rows - 1);
    cc = clampi(cc, 0, cols - 1);
  }
  return in[(size_t)rr * (size_t)cols + (size_t)cc];
}

typedef struct {
  const float* in;
  float* out;
  int rows, cols, R, padding;
  const float* w;
  float neutral;
} mo_ctx;

static void mo_task(void* p, size_t b, size_t e)
{
  mo_ctx* c = (mo_ctx*)p;
  int W = 2 * c->R + 1;
  for (size_t r = b; r < e; ++r) {
    for (int col = 0; col < c->cols; ++col) {
      float sum = 0.0f;
      for (int dy = -c->R; dy <= c->R; ++dy)
        for (int dx = -c->R; dx <= c->R; ++dx)
          sum = sum + getdata_f32(c->in, c->rows, c->cols, (int)r, col, dx, dy,
                                  c->padding, c->neutral) *
                          c->w[(dy + c->R) * W + (dx + c->R)];
      c->out[r * (size_t)c->cols + (size_t)col] = sum;
    }
  }
}

void orc_mapoverlap_stencil_f32(const float* in, float* out, int rows,
                                int cols, const float* w, int R, int padding,
                                float neutral, int nthreads)
{
  mo_ctx c = {in, out, rows, cols, R, padding, w, neutral};
  orc_parallel_for((size_t)rows, nthreads, mo_task, &c);
}

void orc_mapoverlap_stencil_f64(const float* in, double* out, int rows,
                                int cols, const float* w, int R, int padding,
                                float neutral, const int64_t* pix, int npix)
{
  int W = 2 * R + 1;
  for (int k = 0; k < npix; ++k) {
    int r = (int)(pix[k] / cols), c = (int)(pix[k] % cols);
    double s = 0.0;
    for (int dy = -R; dy <= R; ++dy)
      for (int dx = -R; dx <= R; ++dx)
        s += (double)getdata_f32(in, rows, cols, r, c, dx, dy, padding,
                                 neutral) *
             (double)w[(dy + R) * W + (dx + R)];
    out[k] = s;
  }
}

/* The user function `sum += getData(img, dx, dy) * w[...]` as a compiler
 * that contracts it evaluates it: one fmaf per term, dy outer, dx inner. */
void orc_mapoverlap_stencil_f32_fma(const float* in, float* out, int rows,
                                    int cols, const float* w, int R,
                                    int padding, float neutral,
                                    const int64_t* pix, int npix)
{
  int W = 2 * R + 1;
  for (int k = 0; k < npix; ++k) {
    int r = (int)(pix[k] / cols), c = (int)(pix[k] % cols);
    float sum = 0.0f;
    for (int dy = -R; dy <= R; ++dy)
      for (int dx = -R; dx <= R; ++dx)
        sum = fmaf(getdata_f32(in, rows, cols, r, c, dx, dy, padding, neutral),
                   w[(dy + R) * W + (dx + R)], sum);
    out[k] = sum;
  }
}

/* out[r][c] = sum_{dy,dx} w[dy][dx] * in(view(r+dy, c+dx)), so
 * d/d in[p][q] of sum g*out = sum over (r,c,dy,dx) whose view maps to (p,q)
 * of g[r][c]*w[dy][dx].  NEUTRAL: only in-range views; NEAREST: clamped. */
void orc_mapoverlap_stencil_adjoint_f64(const float* g, double* din,
                                        int rows, int cols, const float* w,
                                        int R, int padding,
                                        const int64_t* pix, int npix)
{
  int W = 2 * R + 1;
  for (int k = 0; k < npix; ++k) {
    int p = (int)(pix[k] / cols), q = (int)(pix[k] % cols);
    double s = 0.0;
    /* Only outputs within R of (p,q) can read it: clamping maps r+dy onto
     * row p only when |r-p| <= |dy| <= R (same for columns). */
    for (int r = p - R; r <= p + R; ++r) {
      if (r < 0 || r >= rows) continue;
      for (int c = q - R; c <= q + R; ++c) {
        if (c < 0 || c >= cols) continue;
        for (int dy = -R; dy <= R; ++dy)
          for (int dx = -R; dx <= R; ++dx) {
            int rr = r + dy, cc = c + dx;
            if (rr < 0 || rr >= rows || cc < 0 || cc >= cols) {
              if (padding == 0) continue;
              rr = clampi(rr, 0, rows - 1);
              cc = clampi(cc, 0, cols - 1);
            }
            if (rr == p && cc == q)
              s += (double)g[(size_t)r * (size_t)cols + (size_t)c] *
                   (double)w[(dy + R) * W + (dx + R)];
          }
      }
    }
    din[k] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* AllPairs(Reduce(+), Zip(*)): AllPairsKernel.cl:47-135 -- the output      */
/* starts at SCL_IDENTITY and folds k = 0..K-1 in order:                    */
/* result = USR_REDUCE(result, USR_ZIP(M[row][k], N[k][col])).              */
/* ------------------------------------------------------------------------ */
void orc_allpairs_mm_f32_seqk(const float* A, const float* B, float* C,
                              int M, int N, int K, float id,
                              const int32_t* rows, int nrows)
{
  int cnt = rows ? nrows : M;
  for (int ii = 0; ii < cnt; ++ii) {
    int i = rows ? rows[ii] : ii;
    for (int j = 0; j < N; ++j) {
      float res = id;
      for (int k = 0; k < K; ++k) {
        float z = A[(size_t)i * K + k] * B[(size_t)k * N + j];
        res = res + z;
      }
      C[(size_t)ii * N + j] = res;
    }
  }
}

void orc_allpairs_mm_f64(const float* A, const float* B, double* C, int M,
                         int N, int K, const int32_t* rows, int nrows)
{
  int cnt = rows ? nrows : M;
  double* acc = (double*)malloc(sizeof(double) * (size_t)N);
  for (int ii = 0; ii < cnt; ++ii) {
    int i = rows ? rows[ii] : ii;
    for (int j = 0; j < N; ++j) acc[j] = 0.0;
    for (int k = 0; k < K; ++k) {
      double a = (double)A[(size_t)i * K + k];
      const float* b = B + (size_t)k * N;
      for (int j = 0; j < N; ++j) acc[j] += a * (double)b[j];
    }
    for (int j = 0; j < N; ++j) C[(size_t)ii * N + j] = acc[j];
  }
  free(acc);
}

/* AllPairs general (AllPairsKernel2.cl:42-65) with the squared-Euclidean
 * user function: getElementFromRow(row,k) = L[i][k], getElementFromColumn(
 * col,k) = Rt[k][j] (AllPairsDef.h:292-317); s += t*t in k order. */
void orc_allpairs_sqdist_f32(const float* L, const float* Rt, float* C,
                             int nl, int nr, int d, const int32_t* rows,
                             int nrows)
{
  int cnt = rows ? nrows : nl;
  for (int ii = 0; ii < cnt; ++ii) {
    int i = rows ? rows[ii] : ii;
    for (int j = 0; j < nr; ++j) {
      float s = 0.0f;
      for (int k = 0; k < d; ++k) {
        float t = L[(size_t)i * d + k] - Rt[(size_t)k * nr + j];
        s = s + t * t;
      }
      C[(size_t)ii * nr + j] = s;
    }
  }
}

void orc_allpairs_sqdist_f64(const float* L, const float* Rt, double* C,
                             int nl, int nr, int d, const int32_t* rows,
                             int nrows)
{
  int cnt = rows ? nrows : nl;
  (void)nl;
  for (int ii = 0; ii < cnt; ++ii) {
    int i = rows ? rows[ii] : ii;
    for (int j = 0; j < nr; ++j) {
      double s = 0.0;
      for (int k = 0; k < d; ++k) {
        double t = (double)L[(size_t)i * d + k] - (double)Rt[(size_t)k * nr + j];
        s += t * t;
      }
      C[(size_t)ii * nr + j] = s;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* segmented sums / scans (batched extension), FP64                         */
/* ------------------------------------------------------------------------ */
void orc_segmented_sum_f64(const float* in, const int64_t* offsets, int nseg,
                           double* out)
{
  for (int s = 0; s < nseg; ++s)
    out[s] = orc_sum_f64(in + offsets[s], (size_t)(offsets[s + 1] - offsets[s]));
}

void orc_segmented_scan_f64(const float* in, const int64_t* offsets, int nseg,
                            int inclusive, double* out)
{
  for (int s = 0; s < nseg; ++s)
    orc_scan_add_f64(in + offsets[s], out + offsets[s],
                     (size_t)(offsets[s + 1] - offsets[s]), inclusive);
}

/* the same over 32-bit integers, modulo 2^32 (an int Reduce(+) / Scan(+)
 * folds that wrap, ReduceKernel.cl:51-128 / ScanKernel.cl:74-250 on ints):
 * each segment folded in element order; scan outputs outside
 * [offsets[0], offsets[nseg]) are left as they are */
void orc_segmented_sum_u32(const uint32_t* in, const int64_t* offsets, int nseg,
                           uint32_t* out)
{
  for (int s = 0; s < nseg; ++s) {
    uint32_t acc = 0;
    for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i) acc += in[i];
    out[s] = acc;
  }
}

void orc_segmented_scan_u32(const uint32_t* in, const int64_t* offsets, int nseg,
                            int inclusive, uint32_t* out)
{
  for (int s = 0; s < nseg; ++s)
    orc_scan_add_u32(in + offsets[s], out + offsets[s],
                     (size_t)(offsets[s + 1] - offsets[s]), inclusive);
}


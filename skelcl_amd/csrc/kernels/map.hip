// ---------------------------------------------------------------------------
// Map / Zip skeleton kernels (MapDef.h:182-196, ZipDef.h:191-207 are the
// reference's one-element-per-work-item forms).
//
// Fast path: HBM streaming.  Each thread moves SKL_UNROLL packs of SKL_V
// elements (16-byte global_load/store_dwordx4 for 4-byte types) per batch,
// grid-stride over the whole array; the next batch's loads are issued
// before the current batch is evaluated and stored.  Used when the user function does not ask for its
// work-item id, barriers or __local memory.
// Simple path: one element per work-item with the configured work-group
// size -- get_global_id(0) is the element index, as in the reference.
//
// Generated before this file: SKL_TIN, SKL_TL, SKL_TR, SKL_TOUT, SKL_WG,
// SKL_V, SKL_UNROLL, SKL_CALL1(x) / SKL_CALL2(l, r) (the user function
// with the extra arguments appended), SKL_EXTRA_PARAMS, SKL_EXTRA_SETUP.
// ---------------------------------------------------------------------------

#ifdef SKL_KIND_MAP
extern "C" __global__ __launch_bounds__(SKL_WG) void skl_map(
    const SKL_TIN* in_, SKL_TOUT* out_, unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  typedef skl_pack<SKL_TIN, SKL_V> pin_t;
  typedef skl_pack<SKL_TOUT, SKL_V> pout_t;
  const unsigned long long npk = n / SKL_V;
  constexpr unsigned long long U = SKL_UNROLL, W = SKL_WG;
  const unsigned long long step = (unsigned long long)gridDim.x * W * U;
  auto ld = [&](unsigned long long p) { return skl_load_pack(reinterpret_cast<const pin_t*>(in_) + p); };
  auto emit = [&](unsigned long long p, const pin_t& x) {
    pout_t r;
#pragma unroll
    for (int k = 0; k < SKL_V; ++k) r.v[k] = SKL_CALL1(x.v[k]);
    skl_store_pack(reinterpret_cast<pout_t*>(out_) + p, r);
  };
  unsigned long long p0 = (unsigned long long)blockIdx.x * W * U + threadIdx.x;
  // full batches: unconditional loads, the next batch's loads issued before
  // this batch's stores (one counter tracks loads and stores alike, so a
  // load issued after a store would also wait for the store)
  if (p0 + (U - 1) * W < npk) {
    pin_t a[SKL_UNROLL], b[SKL_UNROLL];
#pragma unroll
    for (int u = 0; u < SKL_UNROLL; ++u) a[u] = ld(p0 + u * W);
    for (; p0 + step + (U - 1) * W < npk; p0 += step) {
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) b[u] = ld(p0 + step + u * W);
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) emit(p0 + u * W, a[u]);
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) a[u] = b[u];
    }
#pragma unroll
    for (int u = 0; u < SKL_UNROLL; ++u) emit(p0 + u * W, a[u]);
    p0 += step;
  }
  if (p0 < npk)  // a partial batch
    for (int u = 0; u < SKL_UNROLL; ++u)
      if (p0 + u * W < npk) emit(p0 + u * W, ld(p0 + u * W));
  if (blockIdx.x == gridDim.x - 1)
    for (unsigned long long i = npk * SKL_V + threadIdx.x; i < n; i += SKL_WG)
      out_[i] = SKL_CALL1(in_[i]);
}

extern "C" __global__ void skl_map_simple(const SKL_TIN* in_, SKL_TOUT* out_,
                                          unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out_[i] = SKL_CALL1(in_[i]);
}
#endif

#ifdef SKL_KIND_MAP_VOID
extern "C" __global__ void skl_map_void(const SKL_TIN* in_,
                                        unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) SKL_CALL1(in_[i]);
}
#endif

#ifdef SKL_KIND_MAP_INDEX
// Map<Tout(Index)>: out[i] = f(offset + i)   (MapDef.h:438-451)
extern "C" __global__ __launch_bounds__(SKL_WG) void skl_map_index(
    SKL_TOUT* out_, unsigned long long n, unsigned long long offset SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  typedef skl_pack<SKL_TOUT, SKL_V> pout_t;
  const unsigned long long npk = n / SKL_V;
  const unsigned long long step = (unsigned long long)gridDim.x * SKL_WG;
  for (unsigned long long p = (unsigned long long)blockIdx.x * SKL_WG + threadIdx.x; p < npk;
       p += step) {
    pout_t r;
#pragma unroll
    for (int k = 0; k < SKL_V; ++k) r.v[k] = SKL_CALL1((Index)(offset + p * SKL_V + k));
    reinterpret_cast<pout_t*>(out_)[p] = r;
  }
  if (blockIdx.x == gridDim.x - 1)
    for (unsigned long long i = npk * SKL_V + threadIdx.x; i < n; i += SKL_WG)
      out_[i] = SKL_CALL1((Index)(offset + i));
}

extern "C" __global__ void skl_map_index_simple(SKL_TOUT* out_, unsigned long long n,
                                                unsigned long long offset SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out_[i] = SKL_CALL1((Index)(offset + i));
}
#endif

#ifdef SKL_KIND_MAP_INDEX_VOID
extern "C" __global__ void skl_map_index_void(unsigned long long n,
                                              unsigned long long offset SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) SKL_CALL1((Index)(offset + i));
}
#endif

#ifdef SKL_KIND_MAP_INDEXPOINT
// Map<Tout(IndexPoint)>: out[y][x] = f({x, y + row_offset})
// (MapDef.h:649-666: work-item dim 0 walks rows, dim 1 walks columns).
extern "C" __global__ void skl_map_indexpoint(SKL_TOUT* out_, unsigned long long rows,
                                              unsigned long long cols,
                                              unsigned long long row_offset SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long r = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long c = (unsigned long long)blockIdx.y * blockDim.y + threadIdx.y;
  if (r < rows && c < cols) {
    IndexPoint p;
    p.x = c;
    p.y = r + row_offset;
#ifdef SKL_VOID_OUT
    SKL_CALL1(p);
#else
    out_[r * cols + c] = SKL_CALL1(p);
#endif
  }
}

// Column-fast variant for user functions that do not read work-item ids:
// threadIdx.x walks columns so stores coalesce.
extern "C" __global__ __launch_bounds__(256) void skl_map_indexpoint_fast(
    SKL_TOUT* out_, unsigned long long rows, unsigned long long cols,
    unsigned long long row_offset SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long c = (unsigned long long)blockIdx.x * 64 + (threadIdx.x & 63);
  const unsigned long long r = (unsigned long long)blockIdx.y * 4 + (threadIdx.x >> 6);
  if (r < rows && c < cols) {
    IndexPoint p;
    p.x = c;
    p.y = r + row_offset;
#ifdef SKL_VOID_OUT
    SKL_CALL1(p);
#else
    out_[r * cols + c] = SKL_CALL1(p);
#endif
  }
}
#endif

#ifdef SKL_KIND_ZIP
extern "C" __global__ __launch_bounds__(SKL_WG) void skl_zip(
    const SKL_TL* l_, const SKL_TR* r_, SKL_TOUT* out_, unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  typedef skl_pack<SKL_TL, SKL_V> pl_t;
  typedef skl_pack<SKL_TR, SKL_V> pr_t;
  typedef skl_pack<SKL_TOUT, SKL_V> pout_t;
  const unsigned long long npk = n / SKL_V;
  constexpr unsigned long long U = SKL_UNROLL, W = SKL_WG;
  const unsigned long long step = (unsigned long long)gridDim.x * W * U;
  auto emit = [&](unsigned long long p, const pl_t& x, const pr_t& y) {
    pout_t r;
#pragma unroll
    for (int k = 0; k < SKL_V; ++k) r.v[k] = SKL_CALL2(x.v[k], y.v[k]);
    skl_store_pack(reinterpret_cast<pout_t*>(out_) + p, r);
  };
  auto ldl = [&](unsigned long long p) { return skl_load_pack(reinterpret_cast<const pl_t*>(l_) + p); };
  auto ldr = [&](unsigned long long p) { return skl_load_pack(reinterpret_cast<const pr_t*>(r_) + p); };
  unsigned long long p0 = (unsigned long long)blockIdx.x * W * U + threadIdx.x;
  // full batches: unconditional loads, the next batch's loads issued before
  // this batch's stores
  if (p0 + (U - 1) * W < npk) {
    pl_t a[SKL_UNROLL], na[SKL_UNROLL];
    pr_t b[SKL_UNROLL], nb[SKL_UNROLL];
#pragma unroll
    for (int u = 0; u < SKL_UNROLL; ++u) {
      a[u] = ldl(p0 + u * W);
      b[u] = ldr(p0 + u * W);
    }
    for (; p0 + step + (U - 1) * W < npk; p0 += step) {
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) {
        na[u] = ldl(p0 + step + u * W);
        nb[u] = ldr(p0 + step + u * W);
      }
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) emit(p0 + u * W, a[u], b[u]);
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) {
        a[u] = na[u];
        b[u] = nb[u];
      }
    }
#pragma unroll
    for (int u = 0; u < SKL_UNROLL; ++u) emit(p0 + u * W, a[u], b[u]);
    p0 += step;
  }
  if (p0 < npk)  // a partial batch
    for (int u = 0; u < SKL_UNROLL; ++u)
      if (p0 + u * W < npk)
        emit(p0 + u * W, reinterpret_cast<const pl_t*>(l_)[p0 + u * W],
             reinterpret_cast<const pr_t*>(r_)[p0 + u * W]);
  if (blockIdx.x == gridDim.x - 1)
    for (unsigned long long i = npk * SKL_V + threadIdx.x; i < n; i += SKL_WG)
      out_[i] = SKL_CALL2(l_[i], r_[i]);
}

extern "C" __global__ void skl_zip_simple(const SKL_TL* l_, const SKL_TR* r_, SKL_TOUT* out_,
                                          unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out_[i] = SKL_CALL2(l_[i], r_[i]);
}
#endif

#ifdef SKL_KIND_ZIP_VOID
extern "C" __global__ void skl_zip_void(const SKL_TL* l_, const SKL_TR* r_,
                                        unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) SKL_CALL2(l_[i], r_[i]);
}
#endif

#ifdef SKL_DERIV_MAP
// Map adjoint: d_in[i] = d_out[i] * f'(in[i], extras...)
extern "C" __global__ __launch_bounds__(SKL_WG) void skl_map_backward(
    const SKL_TIN* in_, const SKL_TOUT* dout_, SKL_TIN* din_, unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long stride = (unsigned long long)gridDim.x * SKL_WG;
  for (unsigned long long i = (unsigned long long)blockIdx.x * SKL_WG + threadIdx.x; i < n;
       i += stride)
    din_[i] = dout_[i] * SKL_DCALL1(in_[i]);
}
#endif

#ifdef SKL_DERIV_ZIP
// Zip adjoint: d_l = d_out * df/dl, d_r = d_out * df/dr (either may be null)
extern "C" __global__ __launch_bounds__(SKL_WG) void skl_zip_backward(
    const SKL_TL* l_, const SKL_TR* r_, const SKL_TOUT* dout_, SKL_TL* dl_, SKL_TR* dr_,
    unsigned long long n SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned long long stride = (unsigned long long)gridDim.x * SKL_WG;
  for (unsigned long long i = (unsigned long long)blockIdx.x * SKL_WG + threadIdx.x; i < n;
       i += stride) {
    SKL_TL a = l_[i];
    SKL_TR b = r_[i];
    SKL_TOUT g = dout_[i];
    if (dl_) dl_[i] = g * SKL_DCALL2L(a, b);
    if (dr_) dr_[i] = g * SKL_DCALL2R(a, b);
  }
}
#endif

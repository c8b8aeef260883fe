// One launch for a batch of uneven work (beyond the reference): the
// segmented Reduce(+) sums and the segmented Scan(+) of one float or int32
// array over the same segments, and a list of small float products, in one
// grid.
//
// Work-groups [0, W) are product workers: each numbers the tiles of all
// products itself (a block scan of the per-product tile counts into LDS) and
// walks the 128x128 tiles W apart, FP32 or BF16X3 (gemm_tile.h).  The other
// work-groups are segment tiles (24576 elements beside products, 32768 in a
// launch of segments only), ordered by a ticket and
// run as one pass of the segmented Scan with a decoupled look-back
// (seg_tile.h); the same pass writes each segment's sum, so the array is
// read once for both.  The tiles find their segments by searching the
// offsets and check a slice of them, so no prologue launch is needed.
//
// Product workers never wait for anything, and a segment tile waits only
// for tiles that drew earlier tickets (work-groups already running), so the
// launch finishes however the work-groups are scheduled.  Both kinds share
// one work-group size and the registers and LDS of the larger (the BF16X3
// tile's: two work-groups per CU), so at full size the specialised launches
// (segmented.hip, matmul.hip) stay faster; one launch wins where launches
// and their gaps dominate (small batches).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"
#include "ld_util.h"
#include "gemm_tile.h"
#include "seg_tile.h"


namespace {

// Segment tiles: 256 threads x 24 packs of 4 (24576 elements) beside the
// products, whose 256-thread work-groups fix the size (their registers
// leave room for 24 packs, the LDS of the products' tile for the per-pack
// prefixes of 24; at the batched workload 3.24 ms against 3.58 with 16
// packs); 512 threads x 16 packs (32768 elements, two work-groups per CU)
// when a launch holds segments only.
constexpr int NT = 256, FJ_P = 24, FJ = 16;
constexpr long long TILE = (long long)NT * FJ_P * 4;
constexpr int NT_SEG = 512;
constexpr long long TILE_SEG = (long long)NT_SEG * FJ * 4;
constexpr int MAXP = 2048;                             // products per launch
constexpr int PREC_NONE = -1;                          // a launch without products

// segment ends per round of the sums: what the LDS leaves
template <int PREC>
constexpr int slots_for() { return PREC == -1 ? 256 : 2048; }

struct BatchJob {
  const float* in;
  float* scan_out;
  float* sums;
  unsigned long long n;
  const int64_t* off;
  int nseg;
  int inclusive;
  unsigned long long* ticket;  // epoch << 32 | next segment tile
  unsigned* err;
  unsigned long long* pub;
  long long ntiles;            // segment tiles
  const skl_gemm_desc* descs;
  int count;                   // products
  int workers;                 // product work-groups
};

template <int PREC>
struct GemmLds;
template <>
struct GemmLds<SKL_PREC_BF16X3> {
  __bf16 As[3][BM][BLD];
  __bf16 Bs[3][BN][BLD];
};
template <>
struct GemmLds<SKL_PREC_FP32> {
  float As[FK][BM + 4];
  float Bs[FK][BN + 4];
};

template <int PREC>
struct WorkerLds {
  GemmLds<PREC> g;
  long long tpre[MAXP + 1];  // tiles of the products before p
};

template <typename A, int PREC>
union BatchLds {
  WorkerLds<PREC> w;
  segtile::Smem<A, NT, FJ_P, slots_for<PREC>()> s;
};
template <typename A>
union BatchLds<A, PREC_NONE> {
  segtile::Smem<A, NT_SEG, FJ, slots_for<PREC_NONE>()> s;
};

__device__ __forceinline__ long long product_tiles(const skl_gemm_desc& g)
{
  if (g.M <= 0 || g.N <= 0) return 0;
  return (long long)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
}

template <int PREC>
__device__ __forceinline__ void product_worker(const BatchJob& jb, WorkerLds<PREC>& L)
{
  // tpre[p] = tiles of products [0, p): thread t counts products
  // [t * per, (t + 1) * per), then one block scan
  __shared__ long long wsum[NT / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int per = (jb.count + NT - 1) / NT;
  const int p0 = (int)threadIdx.x * per;
  const int p1 = p0 + per < jb.count ? p0 + per : jb.count;
  long long c = 0;
  for (int p = p0; p < p1; ++p) c += product_tiles(jb.descs[p]);
  long long v = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[wv] = v;
  __syncthreads();
  long long base = v - c;
  for (int w = 0; w < wv; ++w) base += wsum[w];
  for (int p = p0; p < p1; ++p) {
    L.tpre[p] = base;
    base += product_tiles(jb.descs[p]);
  }
  if (p1 == jb.count && p0 <= p1) L.tpre[jb.count] = base;
  __syncthreads();
  const long long total = L.tpre[jb.count];
  for (long long t = blockIdx.x; t < total; t += jb.workers) {
    int lo = 0, hi = jb.count;  // the product p with tpre[p] <= t < tpre[p + 1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (L.tpre[mid] <= t) lo = mid;
      else hi = mid;
    }
    const skl_gemm_desc g = jb.descs[lo];
    const long long lt = t - L.tpre[lo];
    const int ntn = (g.N + BN - 1) / BN;
    const int m0 = (int)(lt / ntn) * BM, n0 = (int)(lt % ntn) * BN;
    if constexpr (PREC == SKL_PREC_BF16X3)
      bf16x3_tile<2>(g, m0, n0, L.g.As, L.g.Bs);
    else
      f32_tile(g.A, g.B, g.C, g.M, g.N, g.K, 0, 0, g.N, m0, n0, L.g.As, L.g.Bs);
  }
}

template <typename A, int PREC, int NTT, int FJT>
__device__ __forceinline__ void batch_body(const BatchJob& jb, BatchLds<A, PREC>& lds)
{
  __shared__ unsigned long long s_tile;
  __shared__ unsigned s_epoch;
  if constexpr (PREC != PREC_NONE) {
    if ((int)blockIdx.x < jb.workers) {
      product_worker<PREC>(jb, lds.w);
      return;
    }
  }
  // a segment tile: ticket word = epoch << 32 | tile; the work-group
  // drawing the last tile re-arms it for the next launch
  if (threadIdx.x == 0) {
    const unsigned long long w =
        __hip_atomic_fetch_add(jb.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t = w & 0xFFFFFFFFull;
    const unsigned ep = (unsigned)(w >> 32);
    if (t == (unsigned long long)jb.ntiles - 1)
      __hip_atomic_store(jb.ticket, (unsigned long long)((ep + 1u) & 0x3FFFFFFFu) << 32,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_tile = t;
    s_epoch = ep;
  }
  segtile::tile_clear(lds.s);
  __syncthreads();
  segtile::Job j{jb.in,  jb.scan_out, jb.sums,    jb.n,      jb.off,      jb.nseg,
                 nullptr, jb.pub,     jb.err,     jb.ntiles, jb.inclusive};
  segtile::tile_run<A, NTT, FJT, slots_for<PREC>(), true>(j, s_tile, s_epoch, lds.s);
}

template <typename A, int PREC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2))) void batch_kernel(
    BatchJob jb)
{
  __shared__ BatchLds<A, PREC> lds;
  batch_body<A, PREC, NT, FJ_P>(jb, lds);
}

template <typename A>
__global__ __launch_bounds__(NT_SEG) __attribute__((amdgpu_waves_per_eu(4, 4))) void
batch_segments_kernel(BatchJob jb)
{
  __shared__ BatchLds<A, PREC_NONE> lds;
  batch_body<A, PREC_NONE, NT_SEG, FJ>(jb, lds);
}

}  // namespace

// Scratch of a batch launch: ticket (256 B) | error word (256 B) | two
// status words per segment tile.
extern "C" size_t skl_internal_batch_workspace(unsigned long long n, int nseg)
{
  const long long tiles = nseg > 0 ? (long long)((n + TILE - 1) / TILE) : 0;  // the smaller tile
  return 512 + (size_t)(tiles > 0 ? tiles : 1) * 16;
}

extern "C" int skl_internal_batch_max_products() { return MAXP; }

// Epoch of the last batch launch on this scratch and whether its tiles
// found invalid offsets: reads the ticket and error words (synchronously).
extern "C" int skl_internal_batch_status(const void* ws, int* bad)
{
  unsigned long long w[2] = {0, 0};
  if (hipMemcpy(w, ws, 8, hipMemcpyDeviceToHost) != hipSuccess) return SKL_ERR_HIP;
  unsigned e = 0;
  if (hipMemcpy(&e, static_cast<const char*>(ws) + 256, 4, hipMemcpyDeviceToHost) != hipSuccess)
    return SKL_ERR_HIP;
  const unsigned last = ((unsigned)(w[0] >> 32) - 1u) & 0x3FFFFFFFu;
  *bad = (e >> 2) == last && (e & 3u) != 0;
  return SKL_OK;
}

// workers: product work-groups (0: none).  form: the segments' element
// type and accumulator: 0 float, compensated pairs in a tile and FP64
// across tiles; 1 float, FP64 throughout; 2 int32, exact modulo 2^32.
extern "C" int skl_internal_batch_run(const float* in, float* scan_out, float* sums,
                                      unsigned long long n, const int64_t* offsets, int nseg,
                                      int inclusive, int form, const skl_gemm_desc* descs,
                                      int count, int precision, int workers, void* ws,
                                      hipStream_t s)
{
  if (count > MAXP || count < 0 || nseg < 0) return SKL_ERR_INVALID;
  const bool segs = nseg > 0 && (scan_out || sums);
  if (count == 0) workers = 0;
  if (workers < 0 || (workers == 0 && count > 0)) return SKL_ERR_INVALID;
  const long long tile = count > 0 ? TILE : TILE_SEG;
  long long ntiles = segs ? (long long)((n + tile - 1) / tile) : 0;
  if (segs && ntiles == 0) ntiles = 1;  // empty segments still get their zero sums
  if (ntiles > 0x7FFFFFFE) return SKL_ERR_INVALID;
  const long long grid = (long long)workers + ntiles;
  if (grid == 0) return SKL_OK;
  if (grid > 0x7FFFFFFF) return SKL_ERR_INVALID;
  char* p = static_cast<char*>(ws);
  BatchJob jb{in,
              scan_out,
              sums,
              n,
              offsets,
              nseg,
              inclusive ? 1 : 0,
              reinterpret_cast<unsigned long long*>(p),
              reinterpret_cast<unsigned*>(p + 256),
              reinterpret_cast<unsigned long long*>(p + 512),
              ntiles,
              descs,
              count,
              workers};
  const dim3 g((unsigned)grid);
  if (form < 0 || form > 2) return SKL_ERR_INVALID;
  if (count == 0) {
    if (form == 2) batch_segments_kernel<acc_u32><<<g, NT_SEG, 0, s>>>(jb);
    else if (form == 1) batch_segments_kernel<double><<<g, NT_SEG, 0, s>>>(jb);
    else batch_segments_kernel<acc_df><<<g, NT_SEG, 0, s>>>(jb);
  } else if (precision == SKL_PREC_BF16X3) {
    if (form == 2) batch_kernel<acc_u32, SKL_PREC_BF16X3><<<g, NT, 0, s>>>(jb);
    else if (form == 1) batch_kernel<double, SKL_PREC_BF16X3><<<g, NT, 0, s>>>(jb);
    else batch_kernel<acc_df, SKL_PREC_BF16X3><<<g, NT, 0, s>>>(jb);
  } else if (precision == SKL_PREC_FP32) {
    if (form == 2) batch_kernel<acc_u32, SKL_PREC_FP32><<<g, NT, 0, s>>>(jb);
    else if (form == 1) batch_kernel<double, SKL_PREC_FP32><<<g, NT, 0, s>>>(jb);
    else batch_kernel<acc_df, SKL_PREC_FP32><<<g, NT, 0, s>>>(jb);
  } else {
    return SKL_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

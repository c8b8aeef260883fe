// Batched segmented Reduce(+) / Scan(+) on float (north-star extension):
// many uneven segments, given by device offsets, in one persistent launch
// after a one-work-group prologue.
//
// The prologue cuts segment s (length L) into chunks of CH(s) =
// max(4096, roundup(ceil(L / 1024), 1024)) elements (at most 1024 chunks,
// CH <= 16384), writes the chunk prefix sums and re-arms the work counter;
// empty segments get the identity.  The persistent work-groups then take
// chunks from a counter (global chunk order), so a chunk only ever waits on
// chunks handed out earlier.
//   Reduce: each chunk's compensated float sum is stored as an FP64
//           partial; the last chunk of a segment to finish (per-segment
//           counter) adds the segment's partials in chunk order (FP64) and
//           writes out[s].
//   Scan:   a chunk is staged in LDS (64 KB max), its total published with
//           the launch epoch; it then adds the published totals of the
//           segment's earlier chunks in order (FP64 carry) and writes
//           carry + local prefix.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"
#include "work_ticket.h"

namespace {

constexpr int MAXC = 1024;
constexpr long long MINCH = 4096, MAXCH = 16384;
constexpr int WG = 256;

struct SegCtl {
  unsigned long long work;   // chunk ticket
  unsigned long long total;  // number of chunks
  unsigned int epoch;
  unsigned int error;
};

__device__ __forceinline__ long long chunk_size(long long len)
{
  long long ch = (len + MAXC - 1) / MAXC;
  ch = (ch + 1023) / 1024 * 1024;
  return ch < MINCH ? MINCH : ch;
}

// one work-group of 1024 threads
__global__ __launch_bounds__(1024) void seg_prologue(const int64_t* __restrict__ off, int nseg,
                                                     long long* __restrict__ cstart,
                                                     unsigned* __restrict__ done,
                                                     float* __restrict__ out_empty,
                                                     SegCtl* ctl, int is_scan)
{
  __shared__ long long part[1024];
  __shared__ long long carry;
  __shared__ unsigned err;
  if (threadIdx.x == 0) {
    carry = 0;
    err = 0;
  }
  __syncthreads();
  for (int base = 0; base < nseg; base += 1024) {
    const int s = base + threadIdx.x;
    long long cnt = 0;
    if (s < nseg) {
      const long long len = off[s + 1] - off[s];
      if (len < 0) {
        atomicOr(&err, 1u);
      } else if (len > 0) {
        const long long ch = chunk_size(len);
        if (ch > MAXCH) atomicOr(&err, 2u);
        cnt = (len + ch - 1) / ch;
      } else if (!is_scan) {
        out_empty[s] = 0.0f;  // identity of (+)
      }
      done[s] = 0;
    }
    part[threadIdx.x] = cnt;
    __syncthreads();
    // inclusive scan of the 1024 counts (Hillis-Steele)
    for (int d = 1; d < 1024; d <<= 1) {
      long long v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    if (s < nseg) cstart[s] = carry + part[threadIdx.x] - cnt;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cstart[nseg] = carry;
    ctl->work = 0;
    ctl->total = err ? 0 : (unsigned long long)carry;
    ctl->error = err;
    ctl->epoch = (ctl->epoch + 1u) & 0x3FFFFFFFu;
  }
}

__device__ __forceinline__ int find_segment(const long long* cstart, int nseg, long long c)
{
  int lo = 0, hi = nseg;  // cstart[lo] <= c < cstart[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (cstart[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// compensated (two-sum) accumulation in float, returned as hi + lo in double
struct DF {
  float hi, lo;
};
__device__ __forceinline__ DF df_add(DF a, float x)
{
  float s = a.hi + x, bb = s - a.hi;
  float e = (a.hi - (s - bb)) + (x - bb);
  return DF{s, a.lo + e};
}

__device__ __forceinline__ double block_sum_double(double v, double* red)
{
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < WG / 64; ++i) t += red[i];
  __syncthreads();
  return t;  // valid in thread 0
}

__global__ __launch_bounds__(WG) void seg_sum_kernel(const float* __restrict__ in,
                                                     const int64_t* __restrict__ off, int nseg,
                                                     const long long* __restrict__ cstart,
                                                     unsigned* __restrict__ done,
                                                     double* __restrict__ partials,
                                                     float* __restrict__ out, SegCtl* ctl)
{
  __shared__ long long s_c;
  __shared__ double red[WG / 64];
  const long long total = (long long)ctl->total;
  if (threadIdx.x == 0) s_c = ticket_draw(&ctl->work);
  __syncthreads();
  for (;;) {
    const long long c = ticket_read(s_c);
    __syncthreads();
    if (c >= total) break;
    const int s = find_segment(cstart, nseg, c);
    const long long k = c - cstart[s];
    const long long b0 = off[s], len = off[s + 1] - b0;
    const long long ch = chunk_size(len);
    const long long lo = b0 + k * ch;
    const long long hi = lo + ch < b0 + len ? lo + ch : b0 + len;
    DF acc{0.0f, 0.0f};
    long long i = lo + threadIdx.x;
    for (; i + 3 * WG < hi; i += 4 * WG) {
      float x0 = in[i], x1 = in[i + WG], x2 = in[i + 2 * WG], x3 = in[i + 3 * WG];
      acc = df_add(acc, x0);
      acc = df_add(acc, x1);
      acc = df_add(acc, x2);
      acc = df_add(acc, x3);
    }
    for (; i < hi; i += WG) acc = df_add(acc, in[i]);
    const double tot = block_sum_double((double)acc.hi + (double)acc.lo, red);
    if (threadIdx.x == 0) {
      const long long cnt = cstart[s + 1] - cstart[s];
      partials[(size_t)s * MAXC + k] = tot;
      unsigned d = __hip_atomic_fetch_add(&done[s], 1u, __ATOMIC_ACQ_REL,
                                          __HIP_MEMORY_SCOPE_AGENT);
      if ((long long)d == cnt - 1) {  // last chunk of the segment: ordered FP64 combine
        double sum = 0.0;
        for (long long j = 0; j < cnt; ++j)
          sum += __hip_atomic_load(&partials[(size_t)s * MAXC + j], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        out[s] = (float)sum;
      }
      s_c = ticket_draw(&ctl->work);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(WG) void seg_scan_kernel(const float* __restrict__ in,
                                                      const int64_t* __restrict__ off, int nseg,
                                                      const long long* __restrict__ cstart,
                                                      unsigned long long* __restrict__ pub,
                                                      float* __restrict__ out, int inclusive,
                                                      SegCtl* ctl)
{
  __shared__ float tile[MAXCH];
  __shared__ long long s_c;
  __shared__ double red[WG / 64];
  __shared__ double s_thr[WG];
  __shared__ double s_carry;
  const long long total = (long long)ctl->total;
  const unsigned epoch = ctl->epoch;
  if (threadIdx.x == 0) s_c = ticket_draw(&ctl->work);
  __syncthreads();
  for (;;) {
    const long long c = ticket_read(s_c);
    __syncthreads();
    if (c >= total) break;
    const int s = find_segment(cstart, nseg, c);
    const long long k = c - cstart[s];
    const long long b0 = off[s], len = off[s + 1] - b0;
    const long long ch = chunk_size(len);
    const long long lo = b0 + k * ch;
    const int m = (int)((lo + ch < b0 + len ? lo + ch : b0 + len) - lo);
    for (int i = threadIdx.x; i < m; i += WG) tile[i] = in[lo + i];
    __syncthreads();
    // each thread: a contiguous run of per = ceil(m / WG) elements
    const int per = (m + WG - 1) / WG;
    const int t0 = threadIdx.x * per;
    const int t1 = t0 + per < m ? t0 + per : m;
    DF run{0.0f, 0.0f};
    for (int i = t0; i < t1; ++i) run = df_add(run, tile[i]);
    const double mine = (double)run.hi + (double)run.lo;
    s_thr[threadIdx.x] = mine;
    const double tot = block_sum_double(mine, red);
    // publish the chunk total (value bits + epoch tag in two words)
    if (threadIdx.x == 0) {
      const unsigned long long bits = __double_as_longlong(tot);
      unsigned long long* p = pub + ((size_t)s * MAXC + k) * 2;
      const unsigned long long tag = (unsigned long long)epoch << 32;
      __hip_atomic_store(p, tag | (bits & 0xFFFFFFFFull), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(p + 1, tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // carry: the segment's earlier chunks, in order
      double carry = 0.0;
      for (long long j = 0; j < k; ++j) {
        unsigned long long* q = pub + ((size_t)s * MAXC + j) * 2;
        unsigned long long w0, w1;
        do {
          w0 = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          w1 = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } while ((unsigned)(w0 >> 32) != epoch || (unsigned)(w1 >> 32) != epoch);
        carry += __longlong_as_double((long long)((w1 << 32) | (w0 & 0xFFFFFFFFull)));
      }
      s_carry = carry;
    }
    __syncthreads();
    // exclusive scan of the per-thread totals (serial in thread 0: 256 adds)
    if (threadIdx.x == 0) {
      double acc = s_carry;
      for (int t = 0; t < WG; ++t) {
        double v = s_thr[t];
        s_thr[t] = acc;
        acc += v;
      }
    }
    __syncthreads();
    const double base = s_thr[threadIdx.x];
    DF p{0.0f, 0.0f};
    for (int i = t0; i < t1; ++i) {
      const float x = tile[i];
      if (inclusive) {
        p = df_add(p, x);
        tile[i] = (float)(base + ((double)p.hi + (double)p.lo));
      } else {
        tile[i] = (float)(base + ((double)p.hi + (double)p.lo));
        p = df_add(p, x);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += WG) out[lo + i] = tile[i];
    if (threadIdx.x == 0) s_c = ticket_draw(&ctl->work);
    __syncthreads();
  }
}

int grid_size()
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
  }
  return cus * 4;
}

}  // namespace

// workspace: ctl | cstart (nseg+1) | done (nseg) | partials / publish words
extern "C" size_t skl_internal_segmented_scan_workspace(int nseg)
{
  return 256 + (size_t)(nseg + 1) * 8 + (size_t)nseg * 4 + 256 + (size_t)nseg * MAXC * 16;
}

static void carve(void* ws, int nseg, SegCtl** ctl, long long** cstart, unsigned** done,
                  void** data)
{
  char* p = static_cast<char*>(ws);
  *ctl = reinterpret_cast<SegCtl*>(p);
  *cstart = reinterpret_cast<long long*>(p + 256);
  *done = reinterpret_cast<unsigned*>(p + 256 + (size_t)(nseg + 1) * 8);
  size_t o = 256 + (size_t)(nseg + 1) * 8 + (size_t)nseg * 4;
  o = (o + 255) & ~(size_t)255;
  *data = p + o;
}

extern "C" int skl_internal_segmented_sum(float* out, const float* in, const int64_t* offsets,
                                          int nseg, void* ws, hipStream_t s)
{
  SegCtl* ctl;
  long long* cstart;
  unsigned* done;
  void* data;
  carve(ws, nseg, &ctl, &cstart, &done, &data);
  seg_prologue<<<1, 1024, 0, s>>>(offsets, nseg, cstart, done, out, ctl, 0);
  seg_sum_kernel<<<grid_size(), WG, 0, s>>>(in, offsets, nseg, cstart, done,
                                            static_cast<double*>(data), out, ctl);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

extern "C" int skl_internal_segmented_scan(float* out, const float* in, const int64_t* offsets,
                                           int nseg, int inclusive, void* ws, size_t ws_bytes,
                                           hipStream_t s)
{
  (void)ws_bytes;
  SegCtl* ctl;
  long long* cstart;
  unsigned* done;
  void* data;
  carve(ws, nseg, &ctl, &cstart, &done, &data);
  seg_prologue<<<1, 1024, 0, s>>>(offsets, nseg, cstart, done, nullptr, ctl, 1);
  seg_scan_kernel<<<grid_size(), WG, 0, s>>>(in, offsets, nseg, cstart,
                                             static_cast<unsigned long long*>(data), out,
                                             inclusive ? 1 : 0, ctl);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

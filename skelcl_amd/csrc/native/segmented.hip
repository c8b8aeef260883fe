// Batched segmented Reduce(+) / Scan(+) on float (north-star extension):
// many uneven segments, given by device offsets, in one launch after a
// one-work-group prologue.
//
// The prologue validates the offsets (non-decreasing, inside [0, n]), cuts
// segment s (length L) into ceil(L / CH) chunks and numbers the chunks
// class-major: segment s belongs to class s % 64, and position
// p = (s % 64) * M + s / 64 (M = ceil(nseg / 64)) holds its chunk count;
// cstart[] is the prefix over positions, so a class's chunks form one range
// of global chunk ids and a segment's chunks are consecutive.
//   Reduce (CH = 16384): a static grid-stride loop over chunks -- chunks do
//     not wait on each other -- stores each chunk's sum as an FP64 partial
//     (compensated float accumulation, FP64 work-group combine); a second
//     kernel, one wave per segment, adds the segment's partials in a fixed
//     order and writes out[s].  No ticket, no acquire/release.
//   Scan (CH = 4096 = one register tile of 256 lanes x 16 elements): each
//     class has its own relaxed ticket counter (64 counters 256 bytes apart:
//     one hot counter would serialise every draw at ~12 ns), handing out the
//     class's chunks in order; a work-group starts on class blockIdx % 64 and
//     moves on when that class runs dry.  A chunk only waits on earlier
//     chunks of its own segment, drawn earlier from the same counter.
//     Tile-local prefixes in FP64; chunk aggregates and inclusive prefixes
//     are published in epoch-tagged status words (AGG / INCL, 32 bits of the
//     FP64 value per word) and wave 0 looks back over windows of 64
//     predecessors within the segment.  Every output is
//     float(carry + local prefix), all in FP64.
// Invalid offsets set an error flag: no chunk runs, Reduce writes NaN.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "skelcl_rt.h"
#include "ld_util.h"
#include "work_ticket.h"

namespace {

constexpr int WG = 256;
constexpr int ROWS = 16;                              // scan rows of 64 per wave
constexpr long long SCAN_CH = (long long)WG * ROWS;  // 4096 elements per scan chunk
constexpr long long SUM_CH = 16384;                  // elements per reduce chunk
constexpr unsigned ST_AGG = 1u, ST_INCL = 2u;

constexpr int NCLASS = 64;
constexpr int CTR_STRIDE = 32;  // counters 256 bytes apart (in 64-bit words)

struct SegCtl {
  unsigned long long total;  // number of chunks (0 after an error)
  unsigned int epoch;        // scan: tags this launch's status words
  unsigned int error;        // 1: negative length, 2: offsets outside [0, n]
};

__device__ __forceinline__ int seg_of_pos(int p, int m) { return (p / m) + NCLASS * (p % m); }
__device__ __forceinline__ int pos_of_seg(int s, int m) { return (s % NCLASS) * m + s / NCLASS; }

// one work-group of 1024 threads; npos = NCLASS * M positions
__global__ __launch_bounds__(1024) void seg_prologue(const int64_t* __restrict__ off, int nseg,
                                                     int m, unsigned long long n, long long ch,
                                                     long long* __restrict__ cstart,
                                                     unsigned long long* __restrict__ ctr,
                                                     SegCtl* ctl)
{
  __shared__ long long part[1024];
  __shared__ long long carry;
  __shared__ unsigned err;
  const int npos = NCLASS * m;
  if (threadIdx.x == 0) {
    carry = 0;
    err = (off[0] < 0 || (unsigned long long)off[nseg] > n) ? 2u : 0u;
  }
  if (threadIdx.x < NCLASS) ctr[threadIdx.x * CTR_STRIDE] = 0;
  __syncthreads();
  for (int base = 0; base < npos; base += 1024) {
    const int p = base + threadIdx.x;
    long long cnt = 0;
    if (p < npos) {
      const int s = seg_of_pos(p, m);
      if (s < nseg) {
        const long long len = off[s + 1] - off[s];
        if (len < 0) atomicOr(&err, 1u);
        else cnt = (len + ch - 1) / ch;
      }
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
    if (p < npos) cstart[p] = carry + part[threadIdx.x] - cnt;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cstart[npos] = carry;
    ctl->total = err ? 0 : (unsigned long long)carry;
    ctl->error = err;
    ctl->epoch = (ctl->epoch + 1u) & 0x3FFFFFFFu;
  }
}

// position of global chunk c: cstart[lo] <= c < cstart[lo + 1] (positions
// with no chunks are skipped because cstart[lo + 1] > c)
__device__ __forceinline__ int find_position(const long long* cstart, int npos, long long c)
{
  int lo = 0, hi = npos;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (cstart[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// compensated (two-sum) accumulation in float
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
                                                     const int64_t* __restrict__ off, int m,
                                                     const long long* __restrict__ cstart,
                                                     double* __restrict__ partials,
                                                     const SegCtl* __restrict__ ctl)
{
  __shared__ double red[WG / 64];
  const long long total = (long long)ctl->total;
  for (long long c = blockIdx.x; c < total; c += gridDim.x) {
    const int pos = find_position(cstart, NCLASS * m, c);
    const int s = seg_of_pos(pos, m);
    const long long k = c - cstart[pos];
    const long long lo = off[s] + k * SUM_CH;
    const long long end = off[s + 1];
    const long long hi = lo + SUM_CH < end ? lo + SUM_CH : end;
    DF acc{0.0f, 0.0f};
    long long i = lo + threadIdx.x;
    for (; i + 3 * WG < hi; i += 4 * WG) {
      const float x0 = in[i], x1 = in[i + WG], x2 = in[i + 2 * WG], x3 = in[i + 3 * WG];
      acc = df_add(acc, x0);
      acc = df_add(acc, x1);
      acc = df_add(acc, x2);
      acc = df_add(acc, x3);
    }
    for (; i < hi; i += WG) acc = df_add(acc, in[i]);
    const double tot = block_sum_double((double)acc.hi + (double)acc.lo, red);
    if (threadIdx.x == 0) partials[c] = tot;
  }
}

// one wave per segment: partials in a fixed order (lane-strided FP64 sums,
// then a fixed shuffle tree)
__global__ __launch_bounds__(WG) void seg_combine_kernel(const long long* __restrict__ cstart,
                                                         int nseg, int m,
                                                         const double* __restrict__ partials,
                                                         float* __restrict__ out,
                                                         const SegCtl* __restrict__ ctl)
{
  const int s = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= nseg) return;
  if (ctl->error) {
    if (lane == 0) out[s] = __builtin_nanf("");
    return;
  }
  const int pos = pos_of_seg(s, m);
  const long long c0 = cstart[pos], c1 = cstart[pos + 1];
  double v = 0.0;
  for (long long c = c0 + lane; c < c1; c += 64) v += partials[c];
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) out[s] = (float)v;
}

__device__ __forceinline__ void publish(unsigned long long* pub, long long c, unsigned state,
                                        double v, unsigned epoch)
{
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned long long tag = (unsigned long long)((epoch << 2) | state) << 32;
  __hip_atomic_store(&pub[2 * c], tag | (bits & 0xFFFFFFFFull), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&pub[2 * c + 1], tag | (bits >> 32), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// state (AGG / INCL) once both words carry the same tag of this launch, else 0
__device__ __forceinline__ unsigned read_status(const unsigned long long* pub, long long c,
                                                unsigned epoch, double* v)
{
  const unsigned long long w0 =
      __hip_atomic_load(&pub[2 * c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long w1 =
      __hip_atomic_load(&pub[2 * c + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned t0 = (unsigned)(w0 >> 32), t1 = (unsigned)(w1 >> 32);
  if (t0 != t1 || (t0 >> 2) != epoch) return 0;
  *v = __longlong_as_double((long long)((w1 << 32) | (w0 & 0xFFFFFFFFull)));
  return t0 & 3u;
}

// thread 0: the next chunk of class q or, once q runs dry, of the following
// classes; -1 when every class is done.  A dry class is recognised by a
// plain read first, so its counter takes no further RMWs.
__device__ __forceinline__ long long draw_chunk(unsigned long long* ctr,
                                                const long long* cstart, int m, long long total,
                                                int& q)
{
  if (total == 0) return -1;
  for (int tries = 0; tries < NCLASS; ++tries) {
    const long long lo = cstart[q * m], hi = cstart[(q + 1) * m];
    unsigned long long* c = ctr + q * CTR_STRIDE;
    if (lo + (long long)__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < hi) {
      const long long j = ticket_draw(c);
      if (lo + j < hi) return lo + j;
    }
    q = (q + 1) & (NCLASS - 1);
  }
  return -1;
}

// wave 0: lane 0 draws the next chunk (draw_chunk), then the wave finds its
// position with a 64-way search inside the chunk's class -- one coalesced
// load of 64 prefix values per step, one step for classes of <= 64
// positions -- instead of a chain of dependent loads.  Writes *s_c, *s_pos.
__device__ __forceinline__ void next_chunk(unsigned long long* ctr, const long long* cstart,
                                           int m, long long total, int& q, long long* s_c,
                                           int* s_pos)
{
  const int lane = threadIdx.x & 63;
  long long c = -1;
  if (lane == 0) c = draw_chunk(ctr, cstart, m, total, q);
  c = __shfl(c, 0, 64);
  const int cq = __shfl(q, 0, 64);
  int pos = 0;
  if (c >= 0) {
    int lo = cq * m, hi = cq * m + m;  // cstart[lo] <= c < cstart[hi]
    while (hi - lo > 1) {
      const int step = (hi - lo + 63) / 64;
      const int idx = lo + lane * step;
      const bool in = idx < hi;
      const long long v = cstart[in ? idx : lo];
      const unsigned long long mk = __ballot(in && v <= c);  // bit 0 is always set
      lo += (63 - __builtin_clzll(mk)) * step;
      hi = lo + step < hi ? lo + step : hi;
    }
    pos = lo;
  }
  if (lane == 0) {
    *s_c = c;
    *s_pos = pos;
  }
}

__global__ __launch_bounds__(WG) void seg_scan_kernel(const float* __restrict__ in,
                                                      const int64_t* __restrict__ off, int m,
                                                      const long long* __restrict__ cstart,
                                                      unsigned long long* __restrict__ ctr,
                                                      unsigned long long* __restrict__ pub,
                                                      float* __restrict__ out, int inclusive,
                                                      const SegCtl* __restrict__ ctl)
{
  __shared__ long long s_c;
  __shared__ int s_pos;
  __shared__ double s_wtot[WG / 64];
  __shared__ double s_carry;
  const long long total = (long long)ctl->total;
  const unsigned epoch = ctl->epoch;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int q = blockIdx.x & (NCLASS - 1);  // lane 0 of wave 0: current class
  if (wv == 0) next_chunk(ctr, cstart, m, total, q, &s_c, &s_pos);
  __syncthreads();
  for (;;) {
    const long long c = ticket_read(s_c);
    const int pos = (int)__builtin_amdgcn_readfirstlane((unsigned)s_pos);
    __syncthreads();
    if (c < 0) break;
    const int s = seg_of_pos(pos, m);
    const long long cfirst = cstart[pos];
    const long long k = c - cfirst;
    const long long lo = off[s] + k * SCAN_CH;
    const long long end = off[s + 1];
    const long long hi = lo + SCAN_CH < end ? lo + SCAN_CH : end;
    const long long wbase = lo + (long long)wv * (64 * ROWS);

    // row j of this wave = 64 consecutive elements, one per lane
    float x[ROWS];
    double inc[ROWS];
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
      const long long e = wbase + j * 64 + lane;
      x[j] = ld_or(in, (size_t)e, (size_t)lo, e < hi, 0.0f);
    }
    double wsum = 0.0;  // sum of the rows before row j
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
      double v = (double)x[j];
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
      }
      const double rt = __shfl(v, 63, 64);
      inc[j] = wsum + v;  // wave-local inclusive prefix
      wsum += rt;
    }
    if (lane == 0) s_wtot[wv] = wsum;
    __syncthreads();
    double wpre = 0.0, agg = 0.0;
#pragma unroll
    for (int w = 0; w < WG / 64; ++w) {
      const double t = s_wtot[w];
      if (w < wv) wpre += t;
      agg += t;
    }

    // decoupled look-back over this segment's earlier chunks (wave 0)
    if (wv == 0) {
      if (k == 0) {
        if (lane == 0) {
          publish(pub, c, ST_INCL, agg, epoch);
          s_carry = 0.0;
        }
      } else {
        if (lane == 0) publish(pub, c, ST_AGG, agg, epoch);
        double excl = 0.0;
        long long wend = c;  // window [wend - 64, wend)
        for (;;) {
          const long long idx = wend - 64 + lane;
          unsigned st;
          double v = 0.0;
          if (idx < cfirst) {
            st = ST_INCL;  // before the segment: identity
          } else {
            do {
              st = read_status(pub, idx, epoch, &v);
            } while (st == 0);
          }
          const unsigned long long im = __ballot(st == ST_INCL);
          const int h = im ? 63 - __builtin_clzll(im) : -1;
          // lanes [max(h, 0), 63] in order
          double r = lane >= h ? v : 0.0;
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o, 64);
          excl += r;
          if (h >= 0) break;
          wend -= 64;
        }
        if (lane == 0) {
          publish(pub, c, ST_INCL, excl + agg, epoch);
          s_carry = excl;
        }
      }
    }
    __syncthreads();
    const double base = s_carry + wpre;
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
      const long long e = wbase + j * 64 + lane;
      const double p = inclusive ? inc[j] : inc[j] - (double)x[j];
      if (e < hi) out[e] = (float)(base + p);
    }
    if (wv == 0) next_chunk(ctr, cstart, m, total, q, &s_c, &s_pos);
    __syncthreads();
  }
}

int device_cus()
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  return cus;
}

size_t max_chunks(unsigned long long n, int nseg, long long ch)
{
  return (size_t)(n / (unsigned long long)ch) + (size_t)nseg + 1;
}

int classes_m(int nseg) { return (nseg + NCLASS - 1) / NCLASS; }

// workspace: SegCtl (256 B) | 64 ticket counters, 256 B apart | cstart over
// the 64 * M positions (+1) | pad to 256 | per-chunk data
size_t head_bytes(int nseg)
{
  const size_t npos = (size_t)NCLASS * classes_m(nseg);
  return (256 + (size_t)NCLASS * CTR_STRIDE * 8 + (npos + 1) * 8 + 255) & ~(size_t)255;
}

void carve(void* ws, SegCtl** ctl, unsigned long long** ctr, long long** cstart, void** data,
           int nseg)
{
  char* p = static_cast<char*>(ws);
  *ctl = reinterpret_cast<SegCtl*>(p);
  *ctr = reinterpret_cast<unsigned long long*>(p + 256);
  *cstart = reinterpret_cast<long long*>(p + 256 + (size_t)NCLASS * CTR_STRIDE * 8);
  *data = p + head_bytes(nseg);
}

}  // namespace

extern "C" size_t skl_internal_segmented_workspace(unsigned long long n, int nseg)
{
  const size_t sum = max_chunks(n, nseg, SUM_CH) * 8;
  const size_t scan = max_chunks(n, nseg, SCAN_CH) * 16;
  return head_bytes(nseg) + (sum > scan ? sum : scan);
}

extern "C" int skl_internal_segmented_sum(float* out, const float* in, unsigned long long n,
                                          const int64_t* offsets, int nseg, void* ws,
                                          hipStream_t s)
{
  SegCtl* ctl;
  unsigned long long* ctr;
  long long* cstart;
  void* data;
  carve(ws, &ctl, &ctr, &cstart, &data, nseg);
  const int m = classes_m(nseg);
  seg_prologue<<<1, 1024, 0, s>>>(offsets, nseg, m, n, SUM_CH, cstart, ctr, ctl);
  const size_t chunks = max_chunks(n, nseg, SUM_CH);
  const size_t cap = (size_t)device_cus() * 8;
  const unsigned grid = (unsigned)(chunks < cap ? chunks : cap);
  seg_sum_kernel<<<grid, WG, 0, s>>>(in, offsets, m, cstart, static_cast<double*>(data), ctl);
  seg_combine_kernel<<<(unsigned)((nseg + WG / 64 - 1) / (WG / 64)), WG, 0, s>>>(
      cstart, nseg, m, static_cast<const double*>(data), out, ctl);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

extern "C" int skl_internal_segmented_scan(float* out, const float* in, unsigned long long n,
                                           const int64_t* offsets, int nseg, int inclusive,
                                           void* ws, hipStream_t s)
{
  SegCtl* ctl;
  unsigned long long* ctr;
  long long* cstart;
  void* data;
  carve(ws, &ctl, &ctr, &cstart, &data, nseg);
  const int m = classes_m(nseg);
  seg_prologue<<<1, 1024, 0, s>>>(offsets, nseg, m, n, SCAN_CH, cstart, ctr, ctl);
  const size_t chunks = max_chunks(n, nseg, SCAN_CH);
  const size_t cap = (size_t)device_cus() * 8;
  const unsigned grid = (unsigned)(chunks < cap ? chunks : cap);
  seg_scan_kernel<<<grid, WG, 0, s>>>(in, offsets, m, cstart, ctr,
                                      static_cast<unsigned long long*>(data), out,
                                      inclusive ? 1 : 0, ctl);
  return hipGetLastError() == hipSuccess ? SKL_OK : SKL_ERR_HIP;
}

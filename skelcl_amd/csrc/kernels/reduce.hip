// ---------------------------------------------------------------------------
// Reduce skeleton kernel: one launch, HBM streaming, no identity needed.
//
// The reference folds with 8192 strided work-items then one work-group
// tree (ReduceKernel.cl:51-128); like it, this kernel assumes the user
// function is associative and commutative and never uses the identity
// (ReduceDef.h:247-275).  Here:
//   1. each thread folds packs p, p+G, p+2G, ... (G = grid threads), issuing
//      SKL_UNROLL 16-byte loads before folding them;
//   2. a wave64 shuffle tree, then an LDS step across the waves, gives the
//      work-group partial;
//   3. the last work-group to finish (a two-level agent-scope completion
//      count, re-armed by the work-groups that complete it) folds all
//      partials and writes out[0].
// Only threads that own at least one element take part: the valid threads
// are a prefix of the grid, so every wave knows how many lanes are valid.
//
// Accumulator / carry hooks (generated): SKL_ACC_T, skl_lift(x),
// SKL_ACC_COMB(a, b); SKL_CARRY_T, skl_to_carry(acc), SKL_CARRY_COMB(a, b),
// skl_carry_out(c).  For user functions they are the element type and the
// user function.  For float addition (precision DEFAULT) the accumulator is
// a compensated double-float pair and the cross-block carry is FP64.
// For the fused Zip->Reduce form SKL_FUSED_ZIP is defined and element i is
// SKL_ZCALL(l[i], r[i]).
// ---------------------------------------------------------------------------

template <typename A, typename F>
__device__ __forceinline__ A skl_wave_reduce(A v, int nvalid, F comb)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    A o = skl_shfl_down(v, off);
    if (lane + off < nvalid) v = comb(v, o);
  }
  return v;  // lane 0 holds the fold of lanes [0, nvalid)
}

// Fold `v` over the first `nvalid` threads of the work-group; the result is
// valid in thread 0.  `buf` holds SKL_WG/64 entries.
template <typename A, typename F>
__device__ __forceinline__ A skl_block_reduce(A v, int nvalid, A* buf, F comb)
{
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  int wvalid = nvalid - wv * 64;
  wvalid = wvalid < 0 ? 0 : (wvalid > 64 ? 64 : wvalid);
  v = skl_wave_reduce(v, wvalid, comb);
  if (SKL_WG <= 64) return v;
  if (lane == 0 && wvalid > 0) buf[wv] = v;
  __syncthreads();
  const int nw = (nvalid + 63) / 64;
  if (wv == 0) {
    A x = lane < nw ? buf[lane] : v;
    x = skl_wave_reduce(x, nw, comb);
    v = x;
  }
  return v;
}

#ifdef SKL_FUSED_ZIP
#define SKL_REDUCE_PARAMS const SKL_TL* l_, const SKL_TR* r_
#define SKL_ELEM(i) SKL_ZCALL(l_[i], r_[i])
#else
#define SKL_REDUCE_PARAMS const SKL_T* in_
#define SKL_ELEM(i) in_[i]
#endif

extern "C" __global__ __launch_bounds__(SKL_WG) void skl_reduce(
    SKL_REDUCE_PARAMS, SKL_T* out_, unsigned long long n, SKL_CARRY_T* partials,
    unsigned long long* ticket SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  __shared__ SKL_ACC_T s_acc[SKL_WG / 64 > 0 ? SKL_WG / 64 : 1];
  __shared__ SKL_CARRY_T s_car[SKL_WG / 64 > 0 ? SKL_WG / 64 : 1];
  __shared__ int s_last;
  auto comb = [&](const SKL_ACC_T& a, const SKL_ACC_T& b) { return SKL_ACC_COMB(a, b); };
  auto ccomb = [&](const SKL_CARRY_T& a, const SKL_CARRY_T& b) { return SKL_CARRY_COMB(a, b); };

  const unsigned long long nfull = n / SKL_V;             // full packs
  const unsigned long long npk = (n + SKL_V - 1) / SKL_V; // incl. a partial one
  const unsigned long long G = (unsigned long long)gridDim.x * SKL_WG;
  const unsigned long long g = (unsigned long long)blockIdx.x * SKL_WG + threadIdx.x;

  SKL_ACC_T acc;
  __builtin_memset(&acc, 0, sizeof(acc));  // never folded in: defined bits only
  if (g < npk) {
    // first pack: its first element seeds the accumulator
    if (g < nfull) {
#ifdef SKL_FUSED_ZIP
      typedef skl_pack<SKL_TL, SKL_V> pl_t;
      typedef skl_pack<SKL_TR, SKL_V> pr_t;
      pl_t a = reinterpret_cast<const pl_t*>(l_)[g];
      pr_t b = reinterpret_cast<const pr_t*>(r_)[g];
      acc = skl_lift(SKL_ZCALL(a.v[0], b.v[0]));
#pragma unroll
      for (int k = 1; k < SKL_V; ++k) acc = comb(acc, skl_lift(SKL_ZCALL(a.v[k], b.v[k])));
#else
      typedef skl_pack<SKL_T, SKL_V> pin_t;
      pin_t a = reinterpret_cast<const pin_t*>(in_)[g];
      acc = skl_lift(a.v[0]);
#pragma unroll
      for (int k = 1; k < SKL_V; ++k) acc = comb(acc, skl_lift(a.v[k]));
#endif
    } else {
      const unsigned long long e0 = g * SKL_V;
      acc = skl_lift(SKL_ELEM(e0));
      for (unsigned long long e = e0 + 1; e < n; ++e) acc = comb(acc, skl_lift(SKL_ELEM(e)));
    }
    // remaining packs.  Main loop over strides whose SKL_UNROLL packs are
    // all full: the loads are unconditional (a guarded load makes the
    // compiler wait for the previous one) and software-pipelined -- the next
    // stride's loads are issued before the current stride is folded.
#ifdef SKL_FUSED_ZIP
    typedef skl_pack<SKL_TL, SKL_V> pl_t;
    typedef skl_pack<SKL_TR, SKL_V> pr_t;
    struct pk_t {
      pl_t l;
      pr_t r;
    };
    auto ldp = [&](unsigned long long q) {
      pk_t x;
      x.l = skl_load_pack(reinterpret_cast<const pl_t*>(l_) + q);
      x.r = skl_load_pack(reinterpret_cast<const pr_t*>(r_) + q);
      return x;
    };
    auto fold = [&](const pk_t& x) {
#pragma unroll
      for (int k = 0; k < SKL_V; ++k) acc = comb(acc, skl_lift(SKL_ZCALL(x.l.v[k], x.r.v[k])));
    };
#else
    typedef skl_pack<SKL_T, SKL_V> pk_t;
    auto ldp = [&](unsigned long long q) { return skl_load_pack(reinterpret_cast<const pk_t*>(in_) + q); };
    auto fold = [&](const pk_t& x) {
#pragma unroll
      for (int k = 0; k < SKL_V; ++k) acc = comb(acc, skl_lift(x.v[k]));
    };
#endif
    constexpr unsigned long long U = SKL_UNROLL;
    unsigned long long p = g + G;
    if (p + (U - 1) * G < nfull) {
      pk_t a[SKL_UNROLL], b[SKL_UNROLL];
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) a[u] = ldp(p + u * G);
      for (; p + (2 * U - 1) * G < nfull; p += U * G) {
#pragma unroll
        for (int u = 0; u < SKL_UNROLL; ++u) b[u] = ldp(p + (U + u) * G);
#pragma unroll
        for (int u = 0; u < SKL_UNROLL; ++u) fold(a[u]);
#pragma unroll
        for (int u = 0; u < SKL_UNROLL; ++u) a[u] = b[u];
      }
#pragma unroll
      for (int u = 0; u < SKL_UNROLL; ++u) fold(a[u]);
      p += U * G;
    }
    for (; p < nfull; p += G) fold(ldp(p));
    // the partial last pack belongs to the thread whose stride lands on it
    if (p < npk)
      for (unsigned long long e = p * SKL_V; e < n; ++e) acc = comb(acc, skl_lift(SKL_ELEM(e)));
  }
  // valid threads of this work-group: a prefix of it
  const unsigned long long first = (unsigned long long)blockIdx.x * SKL_WG;
  const unsigned long long nvt = npk < G ? npk : G;
  const int nvalid = first >= nvt ? 0 : (int)((nvt - first) < SKL_WG ? (nvt - first) : SKL_WG);
  SKL_ACC_T bacc = skl_block_reduce(acc, nvalid, s_acc, comb);

  // Completion count in two levels, so no single address takes one RMW per
  // work-group (same-address RMWs serialise at ~12 ns each): work-group b
  // counts itself in sub-counter b % 64 (256 bytes apart); the one that
  // completes its class counts the class in the top counter; the one that
  // completes the top counter folds.  Acquire-release RMWs carry every
  // partial to the folding work-group; the completing ones re-arm the
  // counters for the next launch.
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = skl_to_carry(bacc);
    const unsigned nb = gridDim.x, q = blockIdx.x & 63u;
    const unsigned long long members = (nb - 1 - q) / 64 + 1;
    unsigned long long* sub = ticket + 32 * (1 + q);
    const unsigned long long t =
        __hip_atomic_fetch_add(sub, 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    int last = 0;
    if (t == members - 1) {
      __hip_atomic_store(sub, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long classes = nb < 64 ? nb : 64;
      const unsigned long long u =
          __hip_atomic_fetch_add(ticket, 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (u == classes - 1) {
        __hip_atomic_store(ticket, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = 1;
      }
    }
    s_last = last;
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  // last work-group: fold the gridDim.x partials in the carry type
  SKL_CARRY_T c;
  const int np = (int)gridDim.x;
  __builtin_memset(&c, 0, sizeof(c));
  if ((int)threadIdx.x < np) {
    c = partials[threadIdx.x];
    for (int i = threadIdx.x + SKL_WG; i < np; i += SKL_WG) c = ccomb(c, partials[i]);
  }
  const int cvalid = np < SKL_WG ? np : SKL_WG;
  SKL_CARRY_T total = skl_block_reduce(c, cvalid, s_car, ccomb);
  if (threadIdx.x == 0) out_[0] = skl_carry_out(total);
}

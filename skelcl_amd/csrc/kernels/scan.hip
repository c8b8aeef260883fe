// ---------------------------------------------------------------------------
// Scan skeleton kernel: single-pass prefix scan with decoupled look-back.
//
// The reference runs ceil(log_wg n) Blelloch passes plus uniform-add passes
// (ScanDef.h:108-259, ScanKernel.cl:74-250), moving ~4N elements through
// HBM.  This kernel reads N and writes N once:
//   * a work-group takes the next tile (ticket counter, so it only ever
//     waits on tiles that were handed out earlier and are running);
//   * tile = SKL_WG threads x SKL_J packs x SKL_V elements, loaded with
//     coalesced 16-byte accesses: pack j of lane l sits at wave_base +
//     (j*64 + l)*SKL_V; every lane scans its packs, then one wave64 scan per
//     pack row and an LDS step across waves give the tile-local prefixes;
//   * wave 0 publishes the tile aggregate, looks back over windows of
//     64 x SKL_LB_K predecessor tiles, SKL_LB_K consecutive ones per lane
//     (one status word per 32-bit chunk of the carry,
//     tagged with the launch epoch, read and written with agent-scope
//     relaxed atomics) and publishes the inclusive prefix;
//   * outputs = carry (+) local prefix, exclusive (the reference's form,
//     ScanKernel.cl:142-153) or inclusive.
// The user function must be associative; SKL_IDENTITY is its identity.
//
// Hooks (generated): SKL_ACC_T, skl_lift, SKL_ACC_COMB, SKL_ACC_ID,
// SKL_CARRY_T, skl_to_carry, SKL_CARRY_COMB, SKL_CARRY_ID, SKL_PREFIX_T,
// skl_make_prefix(carry, accprefix), SKL_EMIT(prefix, acc),
// skl_emit0(prefix).  For float addition (precision DEFAULT) the tile-local
// prefixes and the carry are FP64.
// ---------------------------------------------------------------------------

#define SKL_SCAN_AGG 1u
#define SKL_SCAN_INCL 2u
#define SKL_CARRY_WORDS ((int)((sizeof(SKL_CARRY_T) + 3) / 4))
// predecessor tiles read per lane and look-back step (window = 64 x SKL_LB_K)
// optional: waves per SIMD the register allocation must leave room for
#if defined(SKL_SCAN_WAVES) && SKL_SCAN_WAVES > 0
#define SKL_SCAN_ATTR __attribute__((amdgpu_waves_per_eu(SKL_SCAN_WAVES)))
#else
#define SKL_SCAN_ATTR
#endif
#ifndef SKL_LB_K
#define SKL_LB_K 1
#endif

__device__ __forceinline__ void skl_scan_publish(unsigned long long* lb,
                                                 unsigned long long tile,
                                                 unsigned state,
                                                 const SKL_CARRY_T& v,
                                                 unsigned epoch)
{
  unsigned w[SKL_CARRY_WORDS];
  w[SKL_CARRY_WORDS - 1] = 0;
  __builtin_memcpy(w, &v, sizeof(SKL_CARRY_T));
  const unsigned long long tag = (unsigned long long)((epoch << 2) | state) << 32;
#pragma unroll
  for (int i = 0; i < SKL_CARRY_WORDS; ++i)
    __hip_atomic_store(&lb[tile * SKL_CARRY_WORDS + i], tag | (unsigned long long)w[i],
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns the state (AGG/INCL) once every word of the tile carries the same
// tag of this launch; 0 while the tile has not published.
__device__ __forceinline__ unsigned skl_scan_read(unsigned long long* lb,
                                                  unsigned long long tile,
                                                  unsigned epoch, SKL_CARRY_T* v)
{
  unsigned w[SKL_CARRY_WORDS];
  unsigned tag0 = 0;
#pragma unroll
  for (int i = 0; i < SKL_CARRY_WORDS; ++i) {
    unsigned long long x = __hip_atomic_load(&lb[tile * SKL_CARRY_WORDS + i],
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned tag = (unsigned)(x >> 32);
    if (i == 0) tag0 = tag;
    else if (tag != tag0) return 0;
    w[i] = (unsigned)x;
  }
  if ((tag0 >> 2) != epoch) return 0;
  __builtin_memcpy(v, w, sizeof(SKL_CARRY_T));
  return tag0 & 3u;
}

extern "C" __global__ __launch_bounds__(SKL_WG) SKL_SCAN_ATTR void skl_scan(
    const SKL_T* in_, SKL_T* out_, unsigned long long n, int inclusive,
    unsigned long long* lb, unsigned long long* ctl SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  constexpr int W = SKL_WG / 64;
  constexpr unsigned long long WAVE_ELEMS = 64ull * SKL_J * SKL_V;
  constexpr unsigned long long TILE = (unsigned long long)SKL_WG * SKL_J * SKL_V;
  typedef skl_pack<SKL_T, SKL_V> pack_t;
  __shared__ unsigned long long s_tile;
  __shared__ unsigned s_epoch;
  __shared__ SKL_ACC_T s_wagg[W];
  __shared__ SKL_CARRY_T s_carry;
  auto comb = [&](const SKL_ACC_T& a, const SKL_ACC_T& b) { return SKL_ACC_COMB(a, b); };
  auto ccomb = [&](const SKL_CARRY_T& a, const SKL_CARRY_T& b) { return SKL_CARRY_COMB(a, b); };
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;

  // ctl[0] = launch epoch << 32 | tile ticket.  One relaxed RMW hands out the
  // tile and the epoch together (the epoch tags the look-back words, so words
  // left by earlier launches never match); the work-group drawing the last
  // ticket re-arms the word for the next launch, after every other
  // work-group's draw.  Nothing else needs ordering: every look-back value
  // travels inside its tagged status word.  (An acquire-release RMW here
  // costs an L2 write-back + invalidate per tile on a multi-XCD part.)
  if (threadIdx.x == 0) {
    const unsigned long long w =
        __hip_atomic_fetch_add(&ctl[0], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t = w & 0xFFFFFFFFull;
    const unsigned ep = (unsigned)(w >> 32);
    if (t == (unsigned long long)gridDim.x - 1)
      __hip_atomic_store(&ctl[0], (unsigned long long)((ep + 1u) & 0x3FFFFFFFu) << 32,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_tile = t;
    s_epoch = ep;
  }
  __syncthreads();
  const unsigned long long tile = s_tile;
  const unsigned epoch = s_epoch;
  const unsigned long long tbase = tile * TILE;
  const unsigned long long wbase = tbase + (unsigned long long)wv * WAVE_ELEMS;
#ifdef SKL_SCAN_REVERSE
  // scan read from the end (the adjoint of a scan): element e of the scan
  // is element n-1-e of the arrays.  The V elements of a pack, e0 .. e0+V-1,
  // are the array run at n-e0-V read backwards: an aligned pack when
  // n % V == 0, otherwise only element-aligned, which for 4-byte multiples
  // is still one dword-aligned vector access.
  constexpr bool SKL_REV_ANY = sizeof(SKL_T) % 4 == 0;
  const bool full = tbase + TILE <= n && (SKL_REV_ANY || n % SKL_V == 0);
#define SKL_RIDX(e) (n - 1 - (e))
#define SKL_PCOMP(k) (SKL_V - 1 - (k))
#define SKL_PLOAD(e0)                                                              \
  (SKL_REV_ANY ? skl_load_pack_u<pack_t>(in_ + (n - (e0) - SKL_V))                   \
               : skl_load_pack(reinterpret_cast<const pack_t*>(in_) + (n - (e0) - SKL_V) / SKL_V))
#define SKL_PSTORE(e0, w)                                                          \
  do {                                                                             \
    if constexpr (SKL_REV_ANY) skl_store_pack_u(out_ + (n - (e0) - SKL_V), (w));     \
    else skl_store_pack(reinterpret_cast<pack_t*>(out_) + (n - (e0) - SKL_V) / SKL_V, (w)); \
  } while (0)
#else
  const bool full = tbase + TILE <= n;
#define SKL_RIDX(e) (e)
#define SKL_PCOMP(k) (k)
#define SKL_PLOAD(e0) skl_load_pack(reinterpret_cast<const pack_t*>(in_) + (e0) / SKL_V)
#define SKL_PSTORE(e0, w) skl_store_pack(reinterpret_cast<pack_t*>(out_) + (e0) / SKL_V, (w))
#endif

  // ---- load + per-lane inclusive scan of each pack
  SKL_ACC_T q[SKL_J][SKL_V];
  if (full) {
    pack_t pk[SKL_J];
#pragma unroll
    for (int j = 0; j < SKL_J; ++j)
      pk[j] = SKL_PLOAD(wbase + ((unsigned long long)j * 64 + lane) * SKL_V);
#pragma unroll
    for (int j = 0; j < SKL_J; ++j)
#pragma unroll
      for (int k = 0; k < SKL_V; ++k) q[j][k] = skl_lift(pk[j].v[SKL_PCOMP(k)]);
  } else {
    SKL_T v[SKL_J][SKL_V];
#pragma unroll
    for (int j = 0; j < SKL_J; ++j) {
      const unsigned long long e0 = wbase + ((unsigned long long)j * 64 + lane) * SKL_V;
#pragma unroll
      for (int k = 0; k < SKL_V; ++k) {
        const unsigned long long e = e0 + k < n ? e0 + k : n - 1;
        v[j][k] = in_[SKL_RIDX(e)];  // unconditional: all loads in flight together
      }
    }
#pragma unroll
    for (int j = 0; j < SKL_J; ++j) {
      const unsigned long long e0 = wbase + ((unsigned long long)j * 64 + lane) * SKL_V;
#pragma unroll
      for (int k = 0; k < SKL_V; ++k) q[j][k] = e0 + k < n ? skl_lift(v[j][k]) : SKL_ACC_ID;
    }
  }
#pragma unroll
  for (int j = 0; j < SKL_J; ++j)
#pragma unroll
    for (int k = 1; k < SKL_V; ++k) q[j][k] = comb(q[j][k - 1], q[j][k]);

  // ---- one wave scan per pack row
  SKL_ACC_T lp[SKL_J];
  SKL_ACC_T rowpre = SKL_ACC_ID;
#pragma unroll
  for (int j = 0; j < SKL_J; ++j) {
    const SKL_ACC_T inc = skl_wave_scan(q[j][SKL_V - 1], lane, comb);
    SKL_ACC_T exc = skl_wave_shr1(inc);
    if (lane == 0) exc = SKL_ACC_ID;
    const SKL_ACC_T rt = skl_readlane(inc, 63);
    lp[j] = comb(rowpre, exc);
    rowpre = comb(rowpre, rt);
  }
  if (lane == 0) s_wagg[wv] = rowpre;
  __syncthreads();
  SKL_ACC_T wpre = SKL_ACC_ID;
  SKL_ACC_T bagg = SKL_ACC_ID;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    SKL_ACC_T a = s_wagg[w];
    if (w < wv) wpre = comb(wpre, a);
    bagg = comb(bagg, a);
  }

  // ---- decoupled look-back (wave 0)
  if (wv == 0) {
    const SKL_CARRY_T cagg = skl_to_carry(bagg);
    if (tile == 0) {
      if (lane == 0) {
        skl_scan_publish(lb, tile, SKL_SCAN_INCL, cagg, epoch);
        s_carry = SKL_CARRY_ID;
      }
    } else {
      if (lane == 0) skl_scan_publish(lb, tile, SKL_SCAN_AGG, cagg, epoch);
      SKL_CARRY_T excl = SKL_CARRY_ID;
      bool have = false;
      long long end = (long long)tile;
      while (true) {
        // window [end - 64K, end): lane l reads the K consecutive tiles
        // starting at end - 64K + lK, all loads issued before any wait
        const long long base = end - 64 * SKL_LB_K + (long long)lane * SKL_LB_K;
        unsigned st[SKL_LB_K];
        SKL_CARRY_T v[SKL_LB_K];
#pragma unroll
        for (int k = 0; k < SKL_LB_K; ++k) {
          st[k] = base + k < 0 ? SKL_SCAN_INCL : 0u;
          v[k] = SKL_CARRY_ID;
        }
        bool pending = true;
        while (pending) {
          pending = false;
#pragma unroll
          for (int k = 0; k < SKL_LB_K; ++k)
            if (st[k] == 0) {
              st[k] = skl_scan_read(lb, (unsigned long long)(base + k), epoch, &v[k]);
              pending |= st[k] == 0;
            }
        }
        // this lane's tiles folded in order from its last inclusive one
        SKL_CARRY_T r = v[0];
        bool incl = st[0] == SKL_SCAN_INCL;
#pragma unroll
        for (int k = 1; k < SKL_LB_K; ++k) {
          if (st[k] == SKL_SCAN_INCL) {
            r = v[k];
            incl = true;
          } else {
            r = ccomb(r, v[k]);
          }
        }
        const unsigned long long im = __ballot(incl);
        const int h = im ? 63 - __builtin_clzll(im) : -1;
        // ordered fold of lanes [h, 63]: lane i ends with r_i (+) ... (+) r_63
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          SKL_CARRY_T o = skl_shfl_down(r, off);
          if (lane + off < 64) r = ccomb(r, o);
        }
        const SKL_CARRY_T ws = skl_shfl(r, h < 0 ? 0 : h);
        excl = have ? ccomb(ws, excl) : ws;
        have = true;
        if (h >= 0) break;
        end -= 64 * SKL_LB_K;
      }
      if (lane == 0) {
        skl_scan_publish(lb, tile, SKL_SCAN_INCL, ccomb(excl, cagg), epoch);
        s_carry = excl;
      }
    }
  }
  __syncthreads();
  const SKL_CARRY_T carry = s_carry;

  // ---- outputs
#pragma unroll
  for (int j = 0; j < SKL_J; ++j) {
    const SKL_PREFIX_T P = skl_make_prefix(carry, comb(wpre, lp[j]));
    pack_t o;
    if (inclusive) {
#pragma unroll
      for (int k = 0; k < SKL_V; ++k) o.v[k] = SKL_EMIT(P, q[j][k]);
    } else {
      o.v[0] = skl_emit0(P);
#pragma unroll
      for (int k = 1; k < SKL_V; ++k) o.v[k] = SKL_EMIT(P, q[j][k - 1]);
    }
    const unsigned long long e0 = wbase + ((unsigned long long)j * 64 + lane) * SKL_V;
    if (full) {
      pack_t w;
#pragma unroll
      for (int k = 0; k < SKL_V; ++k) w.v[SKL_PCOMP(k)] = o.v[k];
      SKL_PSTORE(e0, w);
    } else {
#pragma unroll
      for (int k = 0; k < SKL_V; ++k)
        if (e0 + k < n) out_[SKL_RIDX(e0 + k)] = o.v[k];
    }
  }
}

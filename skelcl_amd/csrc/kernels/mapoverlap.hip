// ---------------------------------------------------------------------------
// MapOverlap skeleton kernel: LDS-staged halo tiles, register-blocked rows.
//
// The reference stages a (L+2R)^2 tile per work-group and evaluates the user
// function once per work-item (MapOverlapKernel.cl:59-148), reading the
// vertical halo from R padding rows built on the host
// (OLDistributionDef.h:263-343).  Here no padded copy exists: a work-group
// of 64x4 threads stages a (SKL_MO_TH+2R) x (64+2R) tile straight from the
// matrix, applying NEUTRAL fill or NEAREST clamping while loading, and each
// thread evaluates the user function for SKL_MO_RB consecutive rows of one
// column.  All evaluations run before any store, so after inlining the
// compiler can reuse the LDS reads that neighbouring rows share.
//
// getData(m, x, y) is the element x columns right and y rows down of the
// centre (MapOverlapDef.h:216-222).
// Generated before this file: SKL_TIN, SKL_TOUT, SCL_OVERLAP_RANGE,
// SKL_PAD_NEUTRAL (+ SKL_NEUTRAL), SKL_MO_RB, SKL_CALL1, SKL_EXTRA_*.
// ---------------------------------------------------------------------------
#define SKL_MO_TW 64
#define SKL_MO_TY 4
#define SKL_MO_TH (SKL_MO_TY * SKL_MO_RB)
#define SCL_TILE_WIDTH (SKL_MO_TW + 2 * SCL_OVERLAP_RANGE)
// optional: waves per SIMD the register allocation must leave room for (the
// unrolled rows otherwise keep the whole input window in registers)
#if defined(SKL_MO_WAVES) && SKL_MO_WAVES > 0
#define SKL_MO_ATTR __attribute__((amdgpu_waves_per_eu(SKL_MO_WAVES)))
#else
#define SKL_MO_ATTR
#endif

typedef struct {
  const SKL_TIN* data;  // LDS tile
  int local_row;        // centre, tile coordinates
  int local_column;
} input_matrix_t;

__device__ __forceinline__ SKL_TIN getData(input_matrix_t m, int x, int y)
{
  return m.data[(m.local_row + y) * SCL_TILE_WIDTH + m.local_column + x];
}

extern "C" __global__ __launch_bounds__(256) SKL_MO_ATTR void skl_mapoverlap(
    const SKL_TIN* in_, SKL_TOUT* out_, int rows, int cols SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  constexpr int R = SCL_OVERLAP_RANGE;
  constexpr int TWP = SKL_MO_TW + 2 * R;
  constexpr int THP = SKL_MO_TH + 2 * R;
  __shared__ SKL_TIN tile[THP * TWP];
  const int c0 = blockIdx.x * SKL_MO_TW;
  const int r0 = blockIdx.y * SKL_MO_TH;
  // every load issues unconditionally from clamped coordinates (the NEAREST
  // rule), so a thread's loads are all in flight before the LDS writes
  constexpr int NLD = (THP * TWP + 255) / 256;
  SKL_TIN v[NLD];
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int idx = (int)threadIdx.x + 256 * u < THP * TWP ? (int)threadIdx.x + 256 * u : 0;
    const int tr = idx / TWP;
    const int tc = idx - tr * TWP;
    const int gr = r0 - R + tr;
    const int gc = c0 - R + tc;
    const int cr = gr < 0 ? 0 : (gr >= rows ? rows - 1 : gr);
    const int cc = gc < 0 ? 0 : (gc >= cols ? cols - 1 : gc);
    v[u] = in_[(size_t)cr * cols + cc];
#ifdef SKL_PAD_NEUTRAL
    if (gr != cr || gc != cc) v[u] = SKL_NEUTRAL;
#endif
  }
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int idx = (int)threadIdx.x + 256 * u;
    if (idx < THP * TWP) tile[idx] = v[u];
  }
  __syncthreads();
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const int c = c0 + tx;
  input_matrix_t m;
  m.data = tile;
  m.local_column = R + tx;
  SKL_TOUT res[SKL_MO_RB];
#pragma unroll
  for (int i = 0; i < SKL_MO_RB; ++i) {
    m.local_row = R + ty * SKL_MO_RB + i;
    res[i] = SKL_CALL1(m);
  }
  if (c < cols) {
    const int rb0 = r0 + ty * SKL_MO_RB;
    if (rb0 + SKL_MO_RB <= rows) {
      // every row in range (all but the last row tile): no per-row guard
#pragma unroll
      for (int i = 0; i < SKL_MO_RB; ++i) out_[(size_t)(rb0 + i) * cols + c] = res[i];
    } else {
#pragma unroll
      for (int i = 0; i < SKL_MO_RB; ++i)
        if (rb0 + i < rows) out_[(size_t)(rb0 + i) * cols + c] = res[i];
    }
  }
}

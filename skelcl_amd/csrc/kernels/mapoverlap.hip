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

typedef struct {
  const SKL_TIN* data;  // LDS tile
  int local_row;        // centre, tile coordinates
  int local_column;
} input_matrix_t;

__device__ __forceinline__ SKL_TIN getData(input_matrix_t m, int x, int y)
{
  return m.data[(m.local_row + y) * SCL_TILE_WIDTH + m.local_column + x];
}

extern "C" __global__ __launch_bounds__(256) void skl_mapoverlap(
    const SKL_TIN* in_, SKL_TOUT* out_, int rows, int cols SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  constexpr int R = SCL_OVERLAP_RANGE;
  constexpr int TWP = SKL_MO_TW + 2 * R;
  constexpr int THP = SKL_MO_TH + 2 * R;
  __shared__ SKL_TIN tile[THP * TWP];
  const int c0 = blockIdx.x * SKL_MO_TW;
  const int r0 = blockIdx.y * SKL_MO_TH;
  for (int idx = threadIdx.x; idx < THP * TWP; idx += 256) {
    const int tr = idx / TWP;
    const int tc = idx - tr * TWP;
    int gr = r0 - R + tr;
    int gc = c0 - R + tc;
    SKL_TIN v;
    if (gr >= 0 && gr < rows && gc >= 0 && gc < cols) {
      v = in_[(size_t)gr * cols + gc];
    } else {
#ifdef SKL_PAD_NEUTRAL
      v = SKL_NEUTRAL;
#else
      gr = gr < 0 ? 0 : (gr >= rows ? rows - 1 : gr);
      gc = gc < 0 ? 0 : (gc >= cols ? cols - 1 : gc);
      v = in_[(size_t)gr * cols + gc];
#endif
    }
    tile[idx] = v;
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
#pragma unroll
    for (int i = 0; i < SKL_MO_RB; ++i) {
      const int r = r0 + ty * SKL_MO_RB + i;
      if (r < rows) out_[(size_t)r * cols + c] = res[i];
    }
  }
}

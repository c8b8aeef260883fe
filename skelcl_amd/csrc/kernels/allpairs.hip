// ---------------------------------------------------------------------------
// AllPairs skeleton kernels (VALU paths).
//
// SKL_KIND_ALLPAIRS -- AllPairs(Reduce, Zip) (AllPairsKernel.cl:47-135):
//   P[i][j] = fold_k reduce(acc, zip(M[i][k], N[k][j])) from SCL_IDENTITY,
//   strictly in k order.  A 256-thread work-group computes a 64x64 block of
//   P; each thread a 4x4 register block (rows ty*4+i, columns tx*4+j).
//   A and B k-slices of 16 are staged in LDS, A transposed, so every inner
//   step is two 16-byte LDS reads shared by the wave.  SKL_AP_FMA (the
//   (+,*) contraction on float with FMA) replaces the user pair by fmaf.
//
// SKL_KIND_ALLPAIRS_GENERAL -- AllPairs(source, func) (AllPairsKernel2.cl:
//   42-65, helpers AllPairsDef.h:292-317): P[i][j] = f(&row_i, &col_j, dim).
//   When the 32 rows and 32 columns of a work-group fit in LDS (SKL_AP_LDS)
//   they are staged there and lmatrix_t / rmatrix_t point into LDS;
//   otherwise they point at the matrices in HBM as in the reference.
// ---------------------------------------------------------------------------

#ifdef SKL_KIND_ALLPAIRS
#define SKL_AP_BM 64
#define SKL_AP_BN 64
#define SKL_AP_BK 16

extern "C" __global__ __launch_bounds__(256) void skl_allpairs(
    const SKL_TL* A, const SKL_TR* B, SKL_TOUT* C, int M, int N, int K SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  __shared__ SKL_TL As[SKL_AP_BK][SKL_AP_BM];
  __shared__ SKL_TR Bs[SKL_AP_BK][SKL_AP_BN];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int m0 = blockIdx.y * SKL_AP_BM, n0 = blockIdx.x * SKL_AP_BN;
  SKL_TOUT acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = SCL_IDENTITY;

  for (int k0 = 0; k0 < K; k0 += SKL_AP_BK) {
    // A: 64 rows x 16 k, read along k; stored transposed
#pragma unroll
    for (int e = tid; e < SKL_AP_BM * SKL_AP_BK; e += 256) {
      const int r = e / SKL_AP_BK, kk = e - r * SKL_AP_BK;
      const int gr = m0 + r, gk = k0 + kk;
      SKL_TL z;
      __builtin_memset(&z, 0, sizeof(z));
      As[kk][r] = skl_ld_or(A, (size_t)gr * K + gk, 0, gr < M && gk < K, z);
    }
    // B: 16 k x 64 columns, read along columns
#pragma unroll
    for (int e = tid; e < SKL_AP_BK * SKL_AP_BN; e += 256) {
      const int kk = e / SKL_AP_BN, c = e - kk * SKL_AP_BN;
      const int gk = k0 + kk, gc = n0 + c;
      SKL_TR z;
      __builtin_memset(&z, 0, sizeof(z));
      Bs[kk][c] = skl_ld_or(B, (size_t)gk * N + gc, 0, gk < K && gc < N, z);
    }
    __syncthreads();
    const int kmax = K - k0 < SKL_AP_BK ? K - k0 : SKL_AP_BK;
    if (kmax == SKL_AP_BK) {
#pragma unroll
      for (int kk = 0; kk < SKL_AP_BK; ++kk) {
        SKL_TL a[4];
        SKL_TR b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = SKL_AP_STEP(acc[i][j], a[i], b[j]);
      }
    } else {
      for (int kk = 0; kk < kmax; ++kk) {
        SKL_TL a[4];
        SKL_TR b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = SKL_AP_STEP(acc[i][j], a[i], b[j]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = m0 + ty * 4 + i;
    if (r >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = n0 + tx * 4 + j;
      if (c < N) C[(size_t)r * N + c] = acc[i][j];
    }
  }
}
#endif

#ifdef SKL_KIND_ALLPAIRS_GENERAL
typedef struct {
  const SKL_TL* data;
  unsigned int dimension;
  unsigned int row;
} lmatrix_t;

typedef struct {
  const SKL_TR* data;
  unsigned int width;
  unsigned int column;
} rmatrix_t;

__device__ __forceinline__ SKL_TL getElementFromRow(lmatrix_t* m, const unsigned int k)
{
  return m->data[(size_t)m->row * m->dimension + k];
}

__device__ __forceinline__ SKL_TR getElementFromColumn(rmatrix_t* m, const unsigned int k)
{
  return m->data[(size_t)k * m->width + m->column];
}

#define SKL_APG_B 32
// LDS-staged: 32 rows of left and 32 columns of right per work-group; taken
// when 64 columns of right exceed 64 KB but this staging fits, which needs
// right elements wider than left ones.
extern "C" __global__ __launch_bounds__(256) void skl_allpairs_general_lds(
    const SKL_TL* A, const SKL_TR* B, SKL_TOUT* C, int M, int N, int K SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  SKL_TL* As = reinterpret_cast<SKL_TL*>(skl_dyn_lds + SKL_APG_LDS_BASE);
  SKL_TR* Bs = reinterpret_cast<SKL_TR*>(skl_dyn_lds + SKL_APG_LDS_BASE +
                                         ((SKL_APG_B * (size_t)K * sizeof(SKL_TL) + 15) & ~(size_t)15));
  const int tid = threadIdx.x;
  const int m0 = blockIdx.y * SKL_APG_B, n0 = blockIdx.x * SKL_APG_B;
  // element loop: the column-staged kernel takes the 4-byte cases
  for (int e = tid; e < SKL_APG_B * K; e += 256) {
    const int r = e / K, k = e - r * K;
    SKL_TL z;
    __builtin_memset(&z, 0, sizeof(z));
    As[e] = skl_ld_or(A, (size_t)(m0 + r) * K + k, 0, m0 + r < M, z);
  }
  for (int e = tid; e < SKL_APG_B * K; e += 256) {
    const int k = e / SKL_APG_B, c = e - k * SKL_APG_B;
    SKL_TR z;
    __builtin_memset(&z, 0, sizeof(z));
    Bs[e] = skl_ld_or(B, (size_t)k * N + n0 + c, 0, n0 + c < N, z);
  }
  __syncthreads();
  const int tx = tid & 31, ty = tid >> 5;
  rmatrix_t Nm;
  Nm.data = Bs;
  Nm.width = SKL_APG_B;
  Nm.column = tx;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int lr = ty + 8 * i;
    lmatrix_t Mm;
    Mm.data = As;
    Mm.dimension = (unsigned)K;
    Mm.row = lr;
    SKL_TOUT v = SKL_CALL3(&Mm, &Nm, (unsigned)SKL_APG_DIM);
    const int r = m0 + lr, c = n0 + tx;
    if (r < M && c < N) C[(size_t)r * N + c] = v;
  }
}

// Column-staged: 64 columns of right in LDS (K x 64), 64 rows per
// work-group of 16 waves.  A wave computes one row at a time with one column
// per lane, so the row's elements are wave-uniform: the compiler reads them
// with scalar loads straight from left (the operands do not alias the
// output: an output overlapping an input would race in any order), and the
// only LDS traffic is the column element, half of what the form above reads.
#define SKL_APG_COLS 64
#define SKL_APG_ROWS 64
extern "C" __global__ __launch_bounds__(1024) void skl_allpairs_general_cols(
    const SKL_TL* __restrict__ A, const SKL_TR* __restrict__ B, SKL_TOUT* __restrict__ C, int M,
    int N, int K SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  SKL_TR* Bs = reinterpret_cast<SKL_TR*>(skl_dyn_lds + SKL_APG_LDS_BASE);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = (int)__builtin_amdgcn_readfirstlane((unsigned)(tid >> 6));
  const int m0 = blockIdx.y * SKL_APG_ROWS, n0 = blockIdx.x * SKL_APG_COLS;
  const bool vec = sizeof(SKL_TR) == 4 && (N & 3) == 0 && ((size_t)B & 15) == 0;
  if (vec) {
    // pack i: k = i / 16, columns 4 (i % 16) .. +3; LDS pack i
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const int nv = 16 * K;
    const u4* gb = reinterpret_cast<const u4*>(B);
    for (int base = 0; base < nv; base += 1024 * 4) {
      u4 rb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + u * 1024 + tid;
        const int c = (i & 15) * 4;
        const bool ok = i < nv && n0 + c < N;
        rb[u] = gb[ok ? ((size_t)(i >> 4) * N + n0 + c) / 4 : 0];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + u * 1024 + tid;
        if (i < nv) {
          const u4 z = {0u, 0u, 0u, 0u};
          reinterpret_cast<u4*>(Bs)[i] = n0 + (i & 15) * 4 < N ? rb[u] : z;
        }
      }
    }
  } else {
    for (int e = tid; e < SKL_APG_COLS * K; e += 1024) {
      const int k = e >> 6, c = e & 63;
      SKL_TR z;
      __builtin_memset(&z, 0, sizeof(z));
      Bs[e] = skl_ld_or(B, (size_t)k * N + n0 + c, 0, n0 + c < N, z);
    }
  }
  __syncthreads();
  rmatrix_t Nm;
  Nm.data = Bs;
  Nm.width = SKL_APG_COLS;
  Nm.column = lane;
  for (int i = 0; i < SKL_APG_ROWS / 16; ++i) {
    const int r = m0 + wv + 16 * i;  // wave-uniform
    if (r >= M) break;
    lmatrix_t Mm;
    Mm.data = A;
    Mm.dimension = (unsigned)K;
    Mm.row = r;
    SKL_TOUT v = SKL_CALL3(&Mm, &Nm, (unsigned)SKL_APG_DIM);
    if (n0 + lane < N) C[(size_t)r * N + n0 + lane] = v;
  }
}

// Reference form: rows and columns read from HBM (any dimension).
extern "C" __global__ void skl_allpairs_general(const SKL_TL* A, const SKL_TR* B, SKL_TOUT* C,
                                                int M, int N, int K SKL_EXTRA_PARAMS)
{
  SKL_EXTRA_SETUP
  const unsigned col = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned row = blockIdx.y * blockDim.y + threadIdx.y;
  lmatrix_t Mm;
  Mm.data = A;
  Mm.dimension = (unsigned)K;
  Mm.row = row;
  rmatrix_t Nm;
  Nm.data = B;
  Nm.width = (unsigned)N;
  Nm.column = col;
  if (row < (unsigned)M && col < (unsigned)N) C[(size_t)row * N + col] = SKL_CALL3(&Mm, &Nm, (unsigned)K);
}
#endif

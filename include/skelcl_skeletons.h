/*
 * skelcl_skeletons.h -- C-ABI of the skeleton execution layer (the hot path).
 *
 * A skeleton object fuses a user function, given as source text, into a
 * hand-written gfx950 skeleton kernel (compiled at run time with hiprtc) and
 * launches it on device buffers.  It replaces what the reference's
 * skeleton templates do below their C++ signatures:
 *   Map<Tout(Tin)>         include/SkelCL/detail/MapDef.h:75-214
 *   Map<void(Tin)>         MapDef.h:218-332
 *   Map<Tout(Index)>       MapDef.h:334-469, src/Map.cpp:53-101
 *   Map<Tout(IndexPoint)>  MapDef.h:532-684, src/Map.cpp:104-159
 *   Zip<Tout(Tl,Tr)>       include/SkelCL/detail/ZipDef.h:70-442
 *   Reduce<T(T)>           include/SkelCL/detail/ReduceDef.h:79-293
 *   Scan<T(T)>             include/SkelCL/detail/ScanDef.h:70-327
 *   MapOverlap<Tout(Tin)>  include/SkelCL/detail/MapOverlapDef.h:71-279
 *   AllPairs(Reduce, Zip)  include/SkelCL/detail/AllPairsDef.h:77-91,212-280
 *   AllPairs(src, func)    AllPairsDef.h:93-107,282-344
 * User-function fusion replaces stooling's transferParameters /
 * transferArguments / renameFunction / redefineTypedef
 * (libraries/stooling/src/SourceCode.cpp:98-283).
 *
 * Extra user arguments follow the reference's marshalling
 * (KernelUtil.h:134-245): a Vector becomes a device pointer, a Matrix a
 * device pointer plus its column count (the user writes T_matrix_t and
 * get(m,row,col)), Local(bytes) a __local pointer to dynamic LDS, anything
 * else is passed by value.
 *
 * All sizes are element counts unless named *_bytes.  Inputs and outputs are
 * device pointers.  Every call is asynchronous on `stream` (NULL = library
 * stream).  A skeleton object keeps scratch memory (partials, scan tile
 * states); calls that use the same object on different streams at the same
 * time race.
 */
#ifndef SKELCL_SKELETONS_H
#define SKELCL_SKELETONS_H

#include "skelcl_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skl_skeleton_s* skl_skeleton;

enum skl_skeleton_kind {
  SKL_MAP = 0,             /* types {out, in}; out "void" = Map<void(Tin)> */
  SKL_MAP_INDEX = 1,       /* types {out};     Map<Tout(Index)>            */
  SKL_MAP_INDEXPOINT = 2,  /* types {out};     Map<Tout(IndexPoint)>       */
  SKL_ZIP = 3,             /* types {out, left, right}                     */
  SKL_REDUCE = 4,          /* types {T}                                    */
  SKL_SCAN = 5,            /* types {T}                                    */
  SKL_MAPOVERLAP = 6,      /* types {out, in}                              */
  SKL_ALLPAIRS = 7,        /* types {out, left, right}; Reduce + Zip       */
  SKL_ALLPAIRS_GENERAL = 8 /* types {out, left, right}; f(row*, col*, dim) */
};

enum skl_padding { SKL_PADDING_NEUTRAL = 0, SKL_PADDING_NEAREST = 1 };

/* Precision form of the float paths (beyond the reference):
 *  DEFAULT  Reduce(+) and Scan(+) on float: compensated FP32 (double-float
 *           pairs) inside each work-group / tile, FP64 only in the
 *           cross-block combine (Reduce partials, Scan look-back carries);
 *           each Scan output is one rounding of carry + tile prefix.  The
 *           batched segmented forms likewise.  AllPairs (+,*) on float:
 *           plain FP32.
 *  FP32     plain FP32 everywhere (user function, FP32 accumulation).
 *  BF16X3   AllPairs (+,*) on float: bf16x3 split on MFMA, FP32 accumulate.
 *  FP8      AllPairs (+,*) on float: FP8 e4m3 MFMA, per-128x128-tile scale.
 *  FP64     Reduce(+) and Scan(+) on float, plain and segmented: FP64
 *           accumulation throughout (every tile prefix and partial in FP64);
 *           other skeletons run as DEFAULT.
 * The AllPairs forms apply to float inputs and output with a zero Reduce
 * identity; a non-zero identity starts each fold (AllPairsKernel.cl:63), so
 * that contraction runs the in-order FP32 FMA fold at every setting. */
enum skl_precision {
  SKL_PREC_DEFAULT = 0,
  SKL_PREC_FP32 = 1,
  SKL_PREC_BF16X3 = 2,
  SKL_PREC_FP8 = 3,
  SKL_PREC_FP64 = 4
};

typedef struct {
  const char* name; /* type name as written in kernel source: "float",
                       "int", "unsigned int", "double", "float4", "void",
                       or a struct the user source defines */
  size_t size;      /* host sizeof (0 for "void") */
} skl_type;

typedef struct {
  int kind;                /* enum skl_skeleton_kind */
  const char* source;      /* user source (Reduce's source for SKL_ALLPAIRS) */
  const char* func;        /* function name; NULL = "func" */
  const char* source2;     /* SKL_ALLPAIRS: the Zip's source */
  const char* func2;       /* SKL_ALLPAIRS: the Zip's function name */
  const char* identity;    /* Reduce/Scan/AllPairs identity; NULL = "0" */
  skl_type types[3];
  int ntypes;
  int overlap_range;       /* MapOverlap: R (MapOverlap.h:119-122 default 1) */
  int padding;             /* MapOverlap: enum skl_padding (default NEAREST) */
  const char* neutral;     /* MapOverlap NEUTRAL value; NULL = "0" */
  int work_group_size;     /* 0 = 256 (src/Skeleton.cpp:48) */
  int precision;           /* enum skl_precision */
  const char* deriv_source;/* adjoint (extension): derivative functions */
  const char* deriv_func;  /* Map: d f/d x (x, extras...) ;
                              Zip: d f/d left (l, r, extras...) */
  const char* deriv_func2; /* Zip: d f/d right (l, r, extras...) */
  const char* build_options; /* extra hiprtc options, may be NULL */
} skl_skeleton_desc;

enum skl_arg_kind {
  SKL_ARG_VALUE = 0,   /* by value: `value` points at `size` bytes        */
  SKL_ARG_BUFFER = 1,  /* Vector: `value` is the device pointer           */
  SKL_ARG_MATRIX = 2,  /* Matrix: `value` device pointer, `cols` columns  */
  SKL_ARG_LOCAL = 3    /* Local(size): `size` bytes of work-group LDS     */
};

typedef struct {
  int kind;
  const void* value;
  size_t size;
  unsigned int cols;
} skl_arg;

/* CommonDefinitions (src/Source.cpp:102-142, Macros.h:55-61): source text
 * that every skeleton created afterwards places ahead of its element
 * typedefs and user function -- the types, functions and macros of
 * SKELCL_COMMON_DEFINITION / SKELCL_ADD_DEFINE.  level: 0 pragmas, 1 user
 * definitions, 2 generated definitions; the text is emitted level by level in
 * the order appended.  Needs no GPU and no skl_init (static initialisers
 * register definitions before main). */
int skl_common_definition_append(const char* text, int level);
/* all appended text, level by level; free with skl_free_string */
int skl_common_definitions(char** text);

/* Skeleton construction: parses the user function, generates the fused
 * kernel source and builds it (hiprtc, cached).  Mirrors the reference
 * constructors' createAndBuildProgram (e.g. MapDef.h:168-214). */
int skl_skeleton_create(const skl_skeleton_desc* desc, skl_skeleton* out);
int skl_skeleton_release(skl_skeleton s);
/* Skeleton::setWorkGroupSize (src/Skeleton.cpp:75-78) */
int skl_skeleton_set_work_group_size(skl_skeleton s, int wg);
/* the work-group size in use (Skeleton::workGroupSize,
 * include/SkelCL/detail/Skeleton.h:64) */
int skl_skeleton_work_group_size(skl_skeleton s, int* wg);
int skl_skeleton_set_precision(skl_skeleton s, int precision);
/* the generated kernel source (for inspection); free with skl_free_string */
int skl_skeleton_source(skl_skeleton s, char** source);
/* the extra parameters the user function declares after the skeleton's own
 * (what kernelUtil::setKernelArgs marshals, KernelUtil.h:134-245):
 * kind = enum skl_arg_kind expected, size = bytes of a by-value type (0 if
 * unknown), type = the declared type text. */
int skl_skeleton_extra_count(skl_skeleton s, int* count);
/* Reads the definition of `func` in `source` (no GPU needed) and returns one
 * line per item: "return <type>" then "param <kind> <size> <type>|<name>"
 * with kind as enum skl_arg_kind.  What stooling's parameterTypeNames /
 * transferParameters inspect (libraries/stooling/src/SourceCode.cpp).
 * Then "op <add|mul|max|min|other>", "work_item <0|1>" and "stencil <N>":
 * N >= 0 when the function is the weighted stencil of radius N
 *   float f(input_matrix_t m, __global float* w) { float s = 0.0f;
 *     for (int dy = -N; dy <= N; ++dy) for (int dx = -N; dx <= N; ++dx)
 *       s += getData(m, dx, dy) * w[(dy + N) * (2N+1) + (dx + N)];
 *     return s; }
 * (any names, token for token otherwise), which a float MapOverlap runs on
 * its native stencil kernel for N <= 5, with the same results; else -1. */
int skl_describe_user_function(const char* source, const char* func,
                               char** description);
int skl_skeleton_extra_info(skl_skeleton s, int index, int* kind, size_t* size,
                            char* type, size_t type_len);
/* like skl_skeleton_create but only generates + compiles (no GPU needed) */
int skl_skeleton_check(const skl_skeleton_desc* desc, char** source,
                       char** log);

/* Map<Tout(Tin)> / Map<void(Tin)> (out may be NULL for void):
 * out[i] = f(in[i], extras...), i < n  -- MapDef.h:116-166 */
int skl_map_run(skl_skeleton s, void* out, const void* in, size_t n,
                const skl_arg* args, int nargs, skl_stream stream);
/* Map<Tout(Index)>: out[i] = f(offset + i) -- MapDef.h:374-419 */
int skl_map_index_run(skl_skeleton s, void* out, size_t n, size_t offset,
                      const skl_arg* args, int nargs, skl_stream stream);
/* Map<Tout(IndexPoint)>: out[y][x] = f({x, y + row_offset}) --
 * MapDef.h:572-628 */
int skl_map_indexpoint_run(skl_skeleton s, void* out, size_t rows,
                           size_t cols, size_t row_offset,
                           const skl_arg* args, int nargs, skl_stream stream);
/* Zip: out[i] = f(left[i], right[i], extras...) -- ZipDef.h:116-173.
 * out may alias left or right (ZipTests.cpp:134-156). */
int skl_zip_run(skl_skeleton s, void* out, const void* left,
                const void* right, size_t n, const skl_arg* args, int nargs,
                skl_stream stream);
/* Reduce: out[0] = f-fold of in[0..n), n >= 1 -- ReduceDef.h:96-127.
 * The identity is not used (ReduceDef.h:247-275); f must be associative
 * and commutative, as the reference's strided fold assumes. */
int skl_reduce_run(skl_skeleton s, void* out, const void* in, size_t n,
                   const skl_arg* args, int nargs, skl_stream stream);
/* skl_reduce_run without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same kernel, grid and
 * result as skl_reduce_run; its scratch (completion counters + one partial
 * per work-group) is a 64 KiB static device array per GPU, held until the
 * process ends, instead of the skeleton's own.  Each call clears the
 * counters with a fill kernel on `stream`; calls that overlap on one GPU
 * race.  Fails with SKL_ERR_INVALID when the launch shape needs more than
 * 64 KiB (16640 bytes of counters + one partial per work-group: the default
 * shape needs about 18.7 KB) or when the skeleton's kernel is not built (a
 * created skeleton always holds it). */
int skl_reduce_run_noalloc(skl_skeleton s, void* out, const void* in,
                           size_t n, const skl_arg* args, int nargs,
                           skl_stream stream);
/* Scan: exclusive (inclusive = 0; the reference's form, ScanKernel.cl:
 * 142-153) or inclusive prefix under f with the identity; f must be
 * associative.  In-place (out == in) is allowed. */
int skl_scan_run(skl_skeleton s, void* out, const void* in, size_t n,
                 int inclusive, const skl_arg* args, int nargs,
                 skl_stream stream);
/* MapOverlap: out[r][c] = f(view centred on (r,c), extras...) over a
 * rows x cols row-major matrix; out-of-range reads give the neutral value
 * or the nearest edge element (MapOverlapKernel.cl:96-141,
 * OLDistributionDef.h:263-343).  No halo copy is made. */
int skl_mapoverlap_run(skl_skeleton s, void* out, const void* in,
                       size_t rows, size_t cols, const skl_arg* args,
                       int nargs, skl_stream stream);
/* AllPairs: out[i][j] over left (rows x dim) and right (dim x cols), both
 * row-major (AllPairsKernel.cl:47-135, AllPairsKernel2.cl:42-65). */
int skl_allpairs_run(skl_skeleton s, void* out, const void* left,
                     const void* right, size_t rows, size_t cols, size_t dim,
                     const skl_arg* args, int nargs, skl_stream stream);

/* ---- extensions (no reference counterpart) ------------------------------ */

/* Zip -> Reduce fused in one kernel: out[0] = reduce_i zip(l[i], r[i]).
 * The zip's output never goes to HBM. */
int skl_zip_reduce_run(skl_skeleton zip, skl_skeleton reduce, void* out,
                       const void* left, const void* right, size_t n,
                       skl_stream stream);

/* Batched segmented Reduce / Scan ((+) on float, or on int32 summed exactly
 * modulo 2^32): nseg segments of the n elements of `in`, described by device
 * offsets[0..nseg] (int64, segment s is offsets[s]..offsets[s+1]), processed
 * by small prologue launches and one main launch (no host
 * synchronisation).  Reduce writes out[s] (nseg elements; an empty segment
 * gives 0); Scan writes out[offsets[s] + k] (n elements; elements outside
 * [offsets[0], offsets[nseg]) are left as they are).  Offsets are validated
 * on the device (non-decreasing, inside [0, n]); invalid offsets make Reduce
 * write NaN (float) or INT_MIN (int32) to every out[s] and Scan write
 * nothing.  No counterpart in the reference. */
int skl_reduce_segmented_run(skl_skeleton s, void* out, const void* in, size_t n,
                             const int64_t* offsets, int nseg,
                             skl_stream stream);
int skl_scan_segmented_run(skl_skeleton s, void* out, const void* in, size_t n,
                           const int64_t* offsets, int nseg, int inclusive,
                           skl_stream stream);

/* Adjoints (vector-Jacobian products) of the float skeletons. */
/* Map: d_in[i] = d_out[i] * f'(in[i], extras...)  (deriv_func) */
int skl_map_backward(skl_skeleton s, void* d_in, const void* in,
                     const void* d_out, size_t n, const skl_arg* args,
                     int nargs, skl_stream stream);
/* Zip: d_left = d_out * df/dl, d_right = d_out * df/dr (either may be NULL) */
int skl_zip_backward(skl_skeleton s, void* d_left, void* d_right,
                     const void* left, const void* right, const void* d_out,
                     size_t n, const skl_arg* args, int nargs,
                     skl_stream stream);
/* Reduce(+): d_in[i] = d_out[0] (broadcast) */
int skl_reduce_backward(skl_skeleton s, void* d_in, const void* d_out,
                        size_t n, skl_stream stream);
/* Scan(+): reversed scan of d_out (exclusive: sum_{j>i}, inclusive:
 * sum_{j>=i}) */
int skl_scan_backward(skl_skeleton s, void* d_in, const void* d_out,
                      size_t n, int inclusive, skl_stream stream);
/* MapOverlap whose user function is a weighted stencil
 * sum_{dy,dx} w[(dy+R)(2R+1)+dx+R] * getData(img,dx,dy): the adjoint is the
 * transposed stencil of d_out under the same padding.  `weights` is the
 * device weight table the forward received.  The first call checks the
 * user function by running it on a small seeded probe against that sum on
 * the host (it allocates and synchronises once); any other user function
 * gives SKL_ERR_UNSUPPORTED. */
int skl_mapoverlap_stencil_backward(skl_skeleton s, void* d_in,
                                    const void* d_out, const float* weights,
                                    size_t rows, size_t cols,
                                    skl_stream stream);
/* AllPairs(+,*): dA = dC . B^T, dB = A^T . dC (either output may be NULL) */
int skl_allpairs_backward(skl_skeleton s, void* d_left, void* d_right,
                          const void* left, const void* right,
                          const void* d_out, size_t rows, size_t cols,
                          size_t dim, skl_stream stream);

/* Dense C = A.B on float (the AllPairs(+,*) contraction) with an explicit
 * precision form; row-major, lda = K, ldb = N, ldc = N.  trans_a / trans_b
 * read A as KxM / B as NxK row-major instead. */
int skl_matmul_f32(float* C, const float* A, const float* B, size_t M,
                   size_t N, size_t K, int trans_a, int trans_b,
                   int precision, skl_stream stream);

/* Many small AllPairs(+,*) products in one launch per kernel class (after a
 * one-work-group prologue that numbers the tiles).  Device arrays
 * of `count` descriptors (pointers + shapes). */
typedef struct {
  const float* A;
  const float* B;
  float* C;
  int M, N, K;
  int pad;
} skl_gemm_desc;
int skl_matmul_batched_f32(const skl_gemm_desc* descs_device, int count,
                           int precision, skl_stream stream);

/* One launch for a batch of uneven work: the segmented Reduce(+) sums
 * (`reduce`, into sums[nseg]) and the segmented Scan(+) (`scan`, into
 * scan_out[n], exclusive or inclusive) of the same array over the same
 * nseg segments (device offsets as above; float, or int32 summed exactly
 * modulo 2^32 -- both skeletons of one element type), and `count` small float
 * products (device descriptors, as for skl_matmul_batched_f32; at most
 * 2048; precision FP32 or BF16X3).  Any part may be left out (NULL output,
 * nseg or count 0).  One kernel launch, no prologue and no host
 * synchronisation: the segment part is one pass over tiles of 24576
 * elements beside products, 32768 alone (a decoupled look-back; the sums are the inclusive values at
 * the segments' ends, so the array is read once for both), and
 * workers_per_cu (1, or 2 = the default 0) work-groups per CU walk the product tiles at
 * the same time.  Float Reduce and Scan must share a precision form
 * (DEFAULT: compensated FP32 in a tile, FP64 across tiles; FP64).  The
 * offsets are checked by the tiles themselves: after invalid offsets the
 * outputs of the call are unspecified (in-bounds writes only) and
 * skl_batch_check reports it.  No counterpart in the reference: for
 * workloads of many uneven segments and small products. */
int skl_batch_run(skl_skeleton reduce, void* sums, skl_skeleton scan,
                  void* scan_out, int inclusive, const void* in, size_t n,
                  const int64_t* offsets, int nseg, const skl_gemm_desc* descs,
                  int count, int precision, int workers_per_cu,
                  skl_stream stream);
/* Synchronises the device, then *invalid_offsets = 1 when the last
 * skl_batch_run on this skeleton (the scan if one was given, else the
 * reduce) found invalid offsets, else 0. */
int skl_batch_check(skl_skeleton s, int* invalid_offsets);

/* Autotuner: time candidate launch shapes for this skeleton at size n
 * (elements; for MapOverlap an n-pixel square image, for AllPairs a product
 * with n output elements and square operands, sqrt(n) on each side) on
 * scratch buffers, keep the fastest and cache it per (kind, user source,
 * types, log2 size) in memory and in $SKELCL_CACHE_DIR/autotune_v2.txt.
 * Tuned: work-group size (256 / 512 for Map/Zip/Reduce, 256 / 512 / 1024 for
 * Scan) with packs in flight per thread and grid cap (Map/Zip/Reduce) or
 * packs per lane (Scan tile); rows per thread (MapOverlap); the MFMA
 * work-group tile of the AllPairs (+,*) products at BF16X3 (128 x 128,
 * 256 x 128) and FP8 (128 x 128, 256 x 128, 256 x 256).  Other AllPairs forms
 * and the Index maps have fixed shapes.  A tuned work-group size replaces
 * one set by skl_skeleton_set_work_group_size.  enable = 0 restores the
 * defaults.  skl_autotune needs a user function without extra parameters;
 * skl_autotune_args times with the caller's extra arguments. */
int skl_autotune(skl_skeleton s, size_t n, int enable);
int skl_autotune_args(skl_skeleton s, size_t n, const skl_arg* args, int nargs);
/* the tuned shape in use: {unroll, grid blocks per CU, scan packs per lane,
 * MapOverlap rows per thread} */
int skl_skeleton_launch_shape(skl_skeleton s, int shape[4]);
/* Pin a launch shape (same layout; ranges unroll 1-16, blocks per CU 1-1024,
 * scan packs 1-32, rows 1-16).  No counterpart in the reference (its launch
 * shapes are fixed): for callers that tune offline or pin a measured shape. */
int skl_skeleton_set_launch_shape(skl_skeleton s, const int shape[4]);
/* The MFMA work-group tile of an AllPairs (+,*) product at BF16X3 / FP8
 * (what skl_autotune tunes for it): 0 = chosen by the product's size (the
 * default), 1 = 128 x 128, 2 = 256 x 128, 3 = 256 x 256 (FP8; BF16X3 takes
 * 256 x 128).  Results do not depend on it.  No counterpart in the
 * reference: for callers that tune offline or pin a measured tile. */
int skl_skeleton_set_mma_tile(skl_skeleton s, int tile);
int skl_skeleton_mma_tile(skl_skeleton s, int* tile);

#ifdef __cplusplus
}
#endif
#endif /* SKELCL_SKELETONS_H */

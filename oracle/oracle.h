/*
 * oracle.h -- CPU restatement of SkelCL's skeleton algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (skelcl_amd/, include/)
 * links, loads or calls this code.  It is used by tests/ as the parity
 * reference and by bench.py as the timed CPU baseline ("port"), never as the
 * thing measured on the GPU.
 *
 * Each function names the reference file:line whose behaviour it restates.
 * Pinning (what checks each function against something other than itself)
 * is listed in the header comment of oracle.c.
 */
#ifndef SKELCL_ORACLE_H
#define SKELCL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based input generator (SplitMix64 finaliser) --------------- */
uint64_t orc_splitmix64(uint64_t seed, uint64_t i);
void orc_gen_u64(uint64_t seed, size_t n, uint64_t* out);
/* 2u-1 with u = (z>>40)*2^-24: uniform [-1,1) on the 2^-23 grid */
void orc_gen_f32_sym(uint64_t seed, size_t n, float* out);
/* scale*u, one rounding */
void orc_gen_f32_unit(uint64_t seed, size_t n, float scale, float* out);
/* ((z>>32)*m)>>32: uniform integer in [0,m) */
void orc_gen_i32_below(uint64_t seed, size_t n, uint32_t m, int32_t* out);
/* floor(10^(e0 + span*d)), d = (z>>11)*2^-53 */
void orc_gen_loguniform(uint64_t seed, size_t n, double e0, double span,
                        int64_t* out);

/* ---- Map / Zip (elementwise) -------------------------------------------- */
void orc_saxpy_f32(float a, const float* x, const float* y, float* out,
                   size_t n, int nthreads);
void orc_saxpy_f32_fma(float a, const float* x, const float* y, float* out,
                       size_t n);
void orc_scale_f32(float a, const float* x, float* out, size_t n);
void orc_zip_mul_f32(const float* x, const float* y, float* out, size_t n);

/* ---- Reduce ------------------------------------------------------------- */
float orc_reduce_add_f32_reforder(const float* in, size_t n, size_t wg,
                                  int nthreads);
uint32_t orc_reduce_add_u32_reforder(const uint32_t* in, size_t n, size_t wg);
double orc_sum_f64(const float* in, size_t n);
double orc_dot_f64(const float* x, const float* y, size_t n);

/* ---- Scan --------------------------------------------------------------- */
void orc_scan_excl_add_f32_reforder(const float* in, float* out, size_t n,
                                    size_t wg, int nthreads);
void orc_scan_excl_add_u32_reforder(const uint32_t* in, uint32_t* out,
                                    size_t n, size_t wg, int nthreads);
void orc_scan_add_u32(const uint32_t* in, uint32_t* out, size_t n,
                      int inclusive);
void orc_scan_add_f64(const float* in, double* out, size_t n, int inclusive);
/* adjoint of an additive scan: reversed scan of the upstream gradient */
void orc_scan_add_adjoint_f64(const float* g, double* dx, size_t n,
                              int inclusive);

/* ---- MapOverlap (2D stencil over an image with padding) ------------------ */
/* padding: 0 = NEUTRAL (value `neutral` outside), 1 = NEAREST (clamp)     */
void orc_mapoverlap_stencil_f32(const float* in, float* out, int rows,
                                int cols, const float* w, int R, int padding,
                                float neutral, int nthreads);
void orc_mapoverlap_stencil_f64(const float* in, double* out, int rows,
                                int cols, const float* w, int R, int padding,
                                float neutral, const int64_t* pix,
                                int npix);
/* the stencil with one fmaf per term (a contracting compiler's form of
 * `sum += getData(..) * w[..]`), at the listed pixels */
void orc_mapoverlap_stencil_f32_fma(const float* in, float* out, int rows,
                                    int cols, const float* w, int R,
                                    int padding, float neutral,
                                    const int64_t* pix, int npix);
/* d(sum W*out)/d(in) for the linear stencil, FP64 */
void orc_mapoverlap_stencil_adjoint_f64(const float* g, double* din,
                                        int rows, int cols, const float* w,
                                        int R, int padding,
                                        const int64_t* pix, int npix);
/* generic getData view (for shift/index tests): value at (r+dy, c+dx) */
int32_t orc_getdata_i32(const int32_t* in, int rows, int cols, int r, int c,
                        int dx, int dy, int padding, int32_t neutral);

/* ---- AllPairs ----------------------------------------------------------- */
void orc_allpairs_mm_f32_seqk(const float* A, const float* B, float* C,
                              int M, int N, int K, float id,
                              const int32_t* rows, int nrows);
void orc_allpairs_mm_f64(const float* A, const float* B, double* C, int M,
                         int N, int K, const int32_t* rows, int nrows);
void orc_allpairs_sqdist_f32(const float* L, const float* Rt, float* C,
                             int nl, int nr, int d, const int32_t* rows,
                             int nrows);
void orc_allpairs_sqdist_f64(const float* L, const float* Rt, double* C,
                             int nl, int nr, int d, const int32_t* rows,
                             int nrows);

/* ---- segmented (batched) reduce / scan, FP64 ----------------------------- */
void orc_segmented_sum_f64(const float* in, const int64_t* offsets,
                           int nseg, double* out);
void orc_segmented_scan_f64(const float* in, const int64_t* offsets,
                            int nseg, int inclusive, double* out);
void orc_segmented_sum_u32(const uint32_t* in, const int64_t* offsets, int nseg,
                           uint32_t* out);
void orc_segmented_scan_u32(const uint32_t* in, const int64_t* offsets, int nseg,
                            int inclusive, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif

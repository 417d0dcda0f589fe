/*
 * refcpu -- CPU restatement of Muesli's single-device skeletons (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline.  Only the test-suite, the smoke check and the CPU
 * baseline leg of bench.py may load it; the product library (libmsl_hip.so)
 * never links, calls or falls back to it.
 *
 * Semantics follow the reference skeleton loops (cited per function in
 * refcpu.c) with the clean-up rules below, which the GPU path shares:
 *   - x is the column, y the row (z the depth); DM index row*ncol+col,
 *     DC index (z*nrow+y)*ncol+x (nplmatrix.h:141-143, plcube.h:121-123).
 *   - one storage per container, no stale CPU shadow copy.
 *   - deterministic folds (fixed blocking), FP64 reference sums for floats.
 *   - floating point is IEEE binary32 with no FMA contraction
 *     (built with -ffp-contract=off) and correctly rounded division.
 *
 * Parity status of each function is recorded next to its declaration:
 * "pinned by" names the independent checks in tests/ that fix it.
 */
#ifndef MSL_REFCPU_H
#define MSL_REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded counter-based inputs -------------------------------------- */
/* SplitMix64 finaliser of seed + (idx+1)*golden-gamma.
 * pinned by: published SplitMix64 test vectors (tests/test_oracle.py). */
uint64_t orc_hash64(uint64_t seed, uint64_t idx);
/* out[i] = (float)(hash>>40) * scale  (24 random bits, one IEEE multiply).
 * pinned by: numpy restatement on the same hashes. */
void orc_fill_uniform_f32(float *out, size_t n, uint64_t seed, float scale);
/* out[i] = hash < threshold ? 1 : 0.  pinned by: numpy restatement. */
void orc_fill_bernoulli_i32(int32_t *out, size_t n, uint64_t seed, uint64_t threshold);
/* out[i] = (uint32)(hash >> 32).  pinned by: numpy restatement. */
void orc_fill_u32(uint32_t *out, size_t n, uint64_t seed);

/* ---- elementwise skeletons (src/ds.cpp:612-727, src/dm.cpp:370-429) ---- */
/* out[i] = in[i]*a + b, two roundings.  pinned by: numpy float32 ops. */
void orc_map_affine_f32(float *out, const float *in, size_t n, float a, float b);
/* out[i] = a[i] + b[i].  pinned by: numpy float32 add. */
void orc_zip_add_f32(float *out, const float *a, const float *b, size_t n);
/* out[row*cols+col] = (float)(row*cols+col), int product then RNE convert.
 * pinned by: numpy arange(...).astype(float32) and the closed form. */
void orc_mapindex_lin_f32(float *out, int rows, int cols);
/* out[i] = in[i]*a + b in uint32 (wrapping).  pinned by: numpy uint32. */
void orc_map_affine_u32(uint32_t *out, const uint32_t *in, size_t n, uint32_t a, uint32_t b);

/* ---- folds (src/ds.cpp:922-954, deterministic) ------------------------ */
/* FP64 sum of float inputs, fixed 4096-element blocks combined in order.
 * pinned by: math.fsum on the same data (tests). */
double orc_fold_sum_f64(const float *in, size_t n);
/* Plain single-chain FP32 sum (the reference's sequential foldCPU order).
 * pinned by: numpy cumulative float32 loop on small inputs; stagnation
 * closed form on a constant array. */
float orc_fold_sum_f32_seq(const float *in, size_t n);
/* Wrapping uint32 sum.  pinned by: numpy uint32 sum. */
uint32_t orc_fold_sum_u32(const uint32_t *in, size_t n);
/* Integer sum of int32 cells into int64.  pinned by: numpy sum. */
int64_t orc_fold_sum_i32(const int32_t *in, size_t n);
/* max over in[i] (n>0).  pinned by: numpy max. */
float orc_fold_max_f32(const float *in, size_t n);
/* max |a[i]-b[i]|.  pinned by: numpy. */
float orc_max_abs_diff_f32(const float *a, const float *b, size_t n);

/* ---- stencils ---------------------------------------------------------- */
/* 2-D mean blur of radius r, neutral value nv outside the grid:
 *   out(x,y) = (sum_{dy=-r..r} (sum_{dx=-r..r} in(x+dx,y+dy))) / (2r+1)^2
 * each inner sum starting from 0.0f, dx and dy ascending.
 * pinned by: brute-force Python loop on tiny grids, constant-field closed
 * forms (interior and border counts). */
void orc_stencil2d_mean_f32(float *out, const float *in, int rows, int cols, int r, float nv);
/* Same in FP64 (drift reference). */
void orc_stencil2d_mean_f64(double *out, const double *in, int rows, int cols, int r, double nv);
/* Game of Life B3/S23 on a Moore neighbourhood, cells outside are 0
 * (examples/gameoflife.cpp:18-43).  pinned by: block, blinker, glider
 * patterns and a brute-force Python loop. */
void orc_gol_i32(int32_t *out, const int32_t *in, int rows, int cols);
/* 3-D 7-point heat step (examples/heatdiffusion.cpp:55-65):
 *   v = 0; v += in(x-1); v += in(x+1); v += in(y-1); v += in(y+1);
 *   v += in(z-1); v += in(z+1); v /= 6      (nv outside the cube)
 * pinned by: brute-force Python loop; constant-field closed forms. */
void orc_heat7_f32(float *out, const float *in, int nx, int ny, int nz, float nv);
void orc_heat7_f64(double *out, const double *in, int nx, int ny, int nz, double nv);
/* heat7 step plus max |out-in| over all cells (fused residual). */
float orc_heat7_residual_f32(float *out, const float *in, int nx, int ny, int nz, float nv);
/* Generic 3-D neighbourhood sum of radius r in the order of the addStencil
 * functor (examples/dc_map_stencil_test.cpp:78-88): dx outermost, then dy,
 * then dz, each ascending, starting from 0.  pinned by: the all-ones closed
 * form 27/18/12/8 at r=1 and a brute-force Python loop. */
void orc_stencil3d_boxsum_f32(float *out, const float *in, int nx, int ny, int nz, int r, float nv);

/* ---- runtime helpers ---------------------------------------------------- */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif

/*
 * refcpu -- the oracle: OpenMP C restatement of Muesli's skeleton loops.
 *
 * TEST INFRASTRUCTURE ONLY (see refcpu.h).  Loaded by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py; never by the product path.
 *
 * Loop structure mirrors the reference's OpenMP build: one
 * `#pragma omp parallel for` over the flattened element index calling the
 * element function, e.g. src/ds.cpp:626-633 (map), src/dm.cpp:1582-1589
 * (2-D stencil, here with x = column as the GPU build of the reference uses,
 * include/detail/map_kernels.cuh:165-166) and src/dc.cpp:579-589 (3-D
 * stencil).  Built with -ffp-contract=off so every float operation is a
 * single IEEE rounding, as on the GPU side.
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Seeded inputs.  SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference
 * constants).  The reference draws unseeded rand() values
 * (examples/gameoflife.cpp:87-89); a counter-based hash lets the CPU and the
 * GPU produce bit-identical inputs independently. */
uint64_t orc_hash64(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_fill_uniform_f32(float *out, size_t n, uint64_t seed, float scale) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        uint32_t u = (uint32_t)(orc_hash64(seed, i) >> 40);
        out[i] = (float)u * scale;
    }
}

void orc_fill_bernoulli_i32(int32_t *out, size_t n, uint64_t seed, uint64_t threshold) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) out[i] = orc_hash64(seed, i) < threshold ? 1 : 0;
}

void orc_fill_u32(uint32_t *out, size_t n, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) out[i] = (uint32_t)(orc_hash64(seed, i) >> 32);
}

/* ------------------------------------------------------------------------ */
/* Elementwise skeletons.  DS::map writes this[k] = f(b[k]) (src/ds.cpp:638-663);
 * DS::zip writes this[k] = f(b[k], c[k]) in signature order (src/ds.cpp:696-727,
 * GPU operand order, src/zip_kernels.cu:41). */
void orc_map_affine_f32(float *out, const float *in, size_t n, float a, float b) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; k++) {
        float t = in[k] * a;
        out[k] = t + b;
    }
}

void orc_zip_add_f32(float *out, const float *a, const float *b, size_t n) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; k++) out[k] = a[k] + b[k];
}

/* DM::mapIndexInPlace: i = k / ncol (row), j = k % ncol (col), f(i, j, v)
 * (src/dm.cpp:386-396).  Functor: row*cols + col as int, stored as float. */
void orc_mapindex_lin_f32(float *out, int rows, int cols) {
    long n = (long)rows * cols;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        int i = (int)(k / cols);
        int j = (int)(k % cols);
        int v = i * cols + j;
        out[k] = (float)v;
    }
}

void orc_map_affine_u32(uint32_t *out, const uint32_t *in, size_t n, uint32_t a, uint32_t b) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; k++) out[k] = in[k] * a + b;
}

/* ------------------------------------------------------------------------ */
/* Folds.  The reference's foldCPU accumulates sequentially from element 0
 * (src/ds.cpp:922-954); its OpenMP loop races on the accumulator, so the
 * oracle uses a fixed blocking instead: 4096-element blocks, each summed in
 * order, block sums combined in block order. */
#define ORC_FOLD_BLOCK 4096

double orc_fold_sum_f64(const float *in, size_t n) {
    size_t nb = (n + ORC_FOLD_BLOCK - 1) / ORC_FOLD_BLOCK;
    double total = 0.0;
    double comp = 0.0;
    /* per-block FP64 sums, combined with Neumaier compensation in order */
    double *part = (double *)malloc(nb ? nb * sizeof(double) : sizeof(double));
#pragma omp parallel for schedule(static)
    for (size_t b = 0; b < nb; b++) {
        size_t lo = b * ORC_FOLD_BLOCK;
        size_t hi = lo + ORC_FOLD_BLOCK < n ? lo + ORC_FOLD_BLOCK : n;
        double s = 0.0;
        for (size_t k = lo; k < hi; k++) s += (double)in[k];
        part[b] = s;
    }
    for (size_t b = 0; b < nb; b++) {
        double t = total + part[b];
        if (fabs(total) >= fabs(part[b]))
            comp += (total - t) + part[b];
        else
            comp += (part[b] - t) + total;
        total = t;
    }
    free(part);
    return total + comp;
}

float orc_fold_sum_f32_seq(const float *in, size_t n) {
    float acc = 0.0f;
    for (size_t k = 0; k < n; k++) acc = acc + in[k];
    return acc;
}

uint32_t orc_fold_sum_u32(const uint32_t *in, size_t n) {
    uint32_t acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (size_t k = 0; k < n; k++) acc += in[k];
    return acc;
}

int64_t orc_fold_sum_i32(const int32_t *in, size_t n) {
    int64_t acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (size_t k = 0; k < n; k++) acc += in[k];
    return acc;
}

float orc_fold_max_f32(const float *in, size_t n) {
    float m = in[0];
#pragma omp parallel for reduction(max : m) schedule(static)
    for (size_t k = 0; k < n; k++) m = in[k] > m ? in[k] : m;
    return m;
}

float orc_max_abs_diff_f32(const float *a, const float *b, size_t n) {
    float m = 0.0f;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (size_t k = 0; k < n; k++) {
        float d = fabsf(a[k] - b[k]);
        m = d > m ? d : m;
    }
    return m;
}

/* ------------------------------------------------------------------------ */
/* 2-D stencils.  NPLMatrix::operator()(x, y) returns the neutral value for
 * x,y outside [0,width)x[0,height), else data[y*width + x]
 * (include/nplmatrix.h:123-143).  Loop: one iteration per flattened index k,
 * x = k % cols, y = k / cols (src/dm.cpp:1582-1589 with the GPU's x = col). */
static inline float npl_get_f32(const float *in, int rows, int cols, float nv, int x, int y) {
    if (x < 0 || y < 0 || x >= cols || y >= rows) return nv;
    return in[(long)y * cols + x];
}

static inline double npl_get_f64(const double *in, int rows, int cols, double nv, int x, int y) {
    if (x < 0 || y < 0 || x >= cols || y >= rows) return nv;
    return in[(long)y * cols + x];
}

void orc_stencil2d_mean_f32(float *out, const float *in, int rows, int cols, int r, float nv) {
    long n = (long)rows * cols;
    float denom = (float)((2 * r + 1) * (2 * r + 1));
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        int x = (int)(k % cols);
        int y = (int)(k / cols);
        float s = 0.0f;
        for (int dy = -r; dy <= r; dy++) {
            float rs = 0.0f;
            for (int dx = -r; dx <= r; dx++) rs = rs + npl_get_f32(in, rows, cols, nv, x + dx, y + dy);
            s = s + rs;
        }
        out[k] = s / denom;
    }
}

void orc_stencil2d_mean_f64(double *out, const double *in, int rows, int cols, int r, double nv) {
    long n = (long)rows * cols;
    double denom = (double)((2 * r + 1) * (2 * r + 1));
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        int x = (int)(k % cols);
        int y = (int)(k / cols);
        double s = 0.0;
        for (int dy = -r; dy <= r; dy++) {
            double rs = 0.0;
            for (int dx = -r; dx <= r; dx++) rs += npl_get_f64(in, rows, cols, nv, x + dx, y + dy);
            s += rs;
        }
        out[k] = s / denom;
    }
}

/* Game of Life functor (examples/gameoflife.cpp:18-43): sum of the eight
 * Moore neighbours read through the padded view (zero outside), then
 * live==1 && (sum==2||sum==3) -> 1, live==0 && sum==3 -> 1, else 0. */
static inline int32_t gol_get(const int32_t *in, int rows, int cols, int x, int y) {
    if (x < 0 || y < 0 || x >= cols || y >= rows) return 0;
    return in[(long)y * cols + x];
}

void orc_gol_i32(int32_t *out, const int32_t *in, int rows, int cols) {
    long n = (long)rows * cols;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        int x = (int)(k % cols);
        int y = (int)(k / cols);
        int32_t sum = gol_get(in, rows, cols, x, y - 1) + gol_get(in, rows, cols, x - 1, y - 1) +
                      gol_get(in, rows, cols, x + 1, y - 1) + gol_get(in, rows, cols, x, y + 1) +
                      gol_get(in, rows, cols, x - 1, y + 1) + gol_get(in, rows, cols, x + 1, y + 1) +
                      gol_get(in, rows, cols, x - 1, y) + gol_get(in, rows, cols, x + 1, y);
        int32_t live = gol_get(in, rows, cols, x, y);
        int32_t next = 0;
        if (live == 1 && (sum == 2 || sum == 3)) next = 1;
        if (live == 0 && sum == 3) next = 1;
        out[k] = next;
    }
}

/* ------------------------------------------------------------------------ */
/* 3-D stencils.  PLCube::operator()(x, y, z) returns the neutral value
 * outside the cube, else data[(z*height + y)*width + x]
 * (include/plcube.h:105-123).  Loop over k: x = k % ncol, y = (k / ncol) %
 * nrow, z = k / (ncol*nrow) (src/dc.cpp:579-589). */
static inline float plc_get_f32(const float *in, int nx, int ny, int nz, float nv, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return nv;
    return in[((long)z * ny + y) * nx + x];
}

static inline double plc_get_f64(const double *in, int nx, int ny, int nz, double nv, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return nv;
    return in[((long)z * ny + y) * nx + x];
}

static inline float heat7_at(const float *in, int nx, int ny, int nz, float nv, int x, int y, int z) {
    float v = 0.0f;
    v = v + plc_get_f32(in, nx, ny, nz, nv, x - 1, y, z);
    v = v + plc_get_f32(in, nx, ny, nz, nv, x + 1, y, z);
    v = v + plc_get_f32(in, nx, ny, nz, nv, x, y - 1, z);
    v = v + plc_get_f32(in, nx, ny, nz, nv, x, y + 1, z);
    v = v + plc_get_f32(in, nx, ny, nz, nv, x, y, z - 1);
    v = v + plc_get_f32(in, nx, ny, nz, nv, x, y, z + 1);
    v = v / 6.0f;
    return v;
}

void orc_heat7_f32(float *out, const float *in, int nx, int ny, int nz, float nv) {
    long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        int x = (int)(k % nx);
        int y = (int)((k / nx) % ny);
        int z = (int)(k / ((long)nx * ny));
        out[k] = heat7_at(in, nx, ny, nz, nv, x, y, z);
    }
}

void orc_heat7_f64(double *out, const double *in, int nx, int ny, int nz, double nv) {
    long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        int x = (int)(k % nx);
        int y = (int)((k / nx) % ny);
        int z = (int)(k / ((long)nx * ny));
        double v = 0.0;
        v += plc_get_f64(in, nx, ny, nz, nv, x - 1, y, z);
        v += plc_get_f64(in, nx, ny, nz, nv, x + 1, y, z);
        v += plc_get_f64(in, nx, ny, nz, nv, x, y - 1, z);
        v += plc_get_f64(in, nx, ny, nz, nv, x, y + 1, z);
        v += plc_get_f64(in, nx, ny, nz, nv, x, y, z - 1);
        v += plc_get_f64(in, nx, ny, nz, nv, x, y, z + 1);
        out[k] = v / 6.0;
    }
}

float orc_heat7_residual_f32(float *out, const float *in, int nx, int ny, int nz, float nv) {
    long n = (long)nx * ny * nz;
    float m = 0.0f;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long k = 0; k < n; k++) {
        int x = (int)(k % nx);
        int y = (int)((k / nx) % ny);
        int z = (int)(k / ((long)nx * ny));
        float v = heat7_at(in, nx, ny, nz, nv, x, y, z);
        out[k] = v;
        float d = fabsf(v - in[k]);
        m = d > m ? d : m;
    }
    return m;
}

void orc_stencil3d_boxsum_f32(float *out, const float *in, int nx, int ny, int nz, int r, float nv) {
    long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        int x = (int)(k % nx);
        int y = (int)((k / nx) % ny);
        int z = (int)(k / ((long)nx * ny));
        float s = 0.0f;
        for (int dx = -r; dx <= r; dx++)
            for (int dy = -r; dy <= r; dy++)
                for (int dz = -r; dz <= r; dz++) s = s + plc_get_f32(in, nx, ny, nz, nv, x + dx, y + dy, z + dz);
        out[k] = s;
    }
}

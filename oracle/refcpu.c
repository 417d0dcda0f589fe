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
 * (include/nplmatrix.h:123-143).  The reference loop is one iteration per
 * flattened index k, x = k % cols, y = k / cols (src/dm.cpp:1582-1589 with
 * the GPU's x = col). */
/* The loops below run over rows (one OpenMP iteration per row) and keep,
 * for every cell, exactly the per-cell operation sequence of the definition:
 * the running sums of all cells of a row are held in per-thread row buffers
 * and advanced one (dy, dx) term at a time in ascending order, so each cell
 * sees the same additions in the same order as a per-element loop would.
 * Reads outside the grid take the neutral value; the bounds test is split
 * off the interior range instead of being done per access.  Values are
 * identical to the per-element form (checked by tests/test_oracle.py
 * against brute-force and numpy restatements). */
#define ORC_ROWSUM_2D(T, NAME)                                                                         \
    void NAME(T *out, const T *in, int rows, int cols, int r, T nv) {                                   \
        T denom = (T)((2 * r + 1) * (2 * r + 1));                                                       \
        _Pragma("omp parallel")                                                                         \
        {                                                                                               \
            T *rs = (T *)malloc((size_t)cols * sizeof(T));                                              \
            T *s = (T *)malloc((size_t)cols * sizeof(T));                                               \
            _Pragma("omp for schedule(static)")                                                         \
            for (int y = 0; y < rows; y++) {                                                            \
                for (int x = 0; x < cols; x++) s[x] = (T)0;                                             \
                for (int dy = -r; dy <= r; dy++) {                                                      \
                    int yy = y + dy;                                                                    \
                    for (int x = 0; x < cols; x++) rs[x] = (T)0;                                        \
                    if (yy < 0 || yy >= rows) {                                                         \
                        for (int dx = -r; dx <= r; dx++)                                                \
                            for (int x = 0; x < cols; x++) rs[x] = rs[x] + nv;                          \
                    } else {                                                                            \
                        const T *row = in + (long)yy * cols;                                            \
                        for (int dx = -r; dx <= r; dx++) {                                              \
                            int lo = dx < 0 ? -dx : 0;               /* first x with x+dx >= 0 */      \
                            int hi = dx > 0 ? cols - dx : cols;      /* end of x with x+dx < cols */   \
                            if (lo > cols) lo = cols;                                                   \
                            if (hi < lo) hi = lo;                                                       \
                            for (int x = 0; x < lo; x++) rs[x] = rs[x] + nv;                            \
                            for (int x = lo; x < hi; x++) rs[x] = rs[x] + row[x + dx];                  \
                            for (int x = hi; x < cols; x++) rs[x] = rs[x] + nv;                         \
                        }                                                                               \
                    }                                                                                   \
                    for (int x = 0; x < cols; x++) s[x] = s[x] + rs[x];                                 \
                }                                                                                       \
                T *o = out + (long)y * cols;                                                            \
                for (int x = 0; x < cols; x++) o[x] = s[x] / denom;                                     \
            }                                                                                           \
            free(rs);                                                                                   \
            free(s);                                                                                    \
        }                                                                                               \
    }

/* 2-D mean blur (the reference has no 2-D mean blur; the definition is
 * out(x,y) = (sum_dy (sum_dx in(x+dx, y+dy))) / (2r+1)^2 with both sums
 * from 0 and ascending, evaluated like the NPLMatrix functor loop of
 * src/dm.cpp:1582-1589 with x = col). */
ORC_ROWSUM_2D(float, orc_stencil2d_mean_f32)
ORC_ROWSUM_2D(double, orc_stencil2d_mean_f64)

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
/* heatdiffusion.cpp:55-65 per cell: v = 0; v += in(x-1); v += in(x+1);
 * v += in(y-1); v += in(y+1); v += in(z-1); v += in(z+1); v /= 6.
 * The loop runs over (z, y) rows; the y/z neighbour rows outside the cube
 * point at a per-thread row of neutral values, the x neighbours outside at
 * nv.  Each cell gets exactly the sequence above. */
#define ORC_HEAT7_CELL(T, x, XM, XP, RESID)                                                          \
    {                                                                                                   \
        T v = (T)0;                                                                                     \
        v = v + (XM);                                                                                   \
        v = v + (XP);                                                                                   \
        v = v + ym[x];                                                                                  \
        v = v + yp[x];                                                                                  \
        v = v + zm[x];                                                                                  \
        v = v + zp[x];                                                                                  \
        v = v / (T)6;                                                                                   \
        o[x] = v;                                                                                       \
        if (RESID) {                                                                                    \
            T d = v - c[x];                                                                             \
            d = d < 0 ? -d : d;                                                                         \
            mres = d > mres ? d : mres;                                                                 \
        }                                                                                               \
    }

#define ORC_HEAT7(T, NAME, RESID, RET, RETV)                                                            \
    RET NAME(T *out, const T *in, int nx, int ny, int nz, T nv) {                                       \
        T mres = (T)0;                                                                                  \
        long nrows = (long)ny * nz;                                                                     \
        _Pragma("omp parallel reduction(max : mres)")                                                   \
        {                                                                                               \
            T *nvrow = (T *)malloc((size_t)nx * sizeof(T));                                             \
            for (int x = 0; x < nx; x++) nvrow[x] = nv;                                                 \
            _Pragma("omp for schedule(static)")                                                         \
            for (long q = 0; q < nrows; q++) {                                                          \
                int y = (int)(q % ny), z = (int)(q / ny);                                               \
                const T *c = in + q * nx;                                                               \
                const T *ym = y > 0 ? c - nx : nvrow;                                                   \
                const T *yp = y + 1 < ny ? c + nx : nvrow;                                              \
                const T *zm = z > 0 ? c - (long)nx * ny : nvrow;                                        \
                const T *zp = z + 1 < nz ? c + (long)nx * ny : nvrow;                                   \
                T *o = out + q * nx;                                                                    \
                if (nx == 1) {                                                                          \
                    ORC_HEAT7_CELL(T, 0, nv, nv, RESID)                                                 \
                } else {                                                                                \
                    ORC_HEAT7_CELL(T, 0, nv, c[1], RESID)                                               \
                    for (int x = 1; x < nx - 1; x++) ORC_HEAT7_CELL(T, x, c[x - 1], c[x + 1], RESID)    \
                    ORC_HEAT7_CELL(T, nx - 1, c[nx - 2], nv, RESID)                                     \
                }                                                                                       \
            }                                                                                           \
            free(nvrow);                                                                                \
        }                                                                                               \
        (void)mres;                                                                                     \
        return RETV;                                                                                    \
    }

/* PLCube::operator()(x, y, z) returns the neutral value outside the cube,
 * else data[(z*height + y)*width + x] (include/plcube.h:105-123); the
 * reference loop is src/dc.cpp:579-589. */
ORC_HEAT7(float, orc_heat7_f32, 0, void, )
ORC_HEAT7(double, orc_heat7_f64, 0, void, )
/* heat7 step plus max |out-in| over all cells (fused residual). */
ORC_HEAT7(float, orc_heat7_residual_f32, 1, float, mres)

static inline float plc_get_f32(const float *in, int nx, int ny, int nz, float nv, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return nv;
    return in[((long)z * ny + y) * nx + x];
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

/*
 * skeleton_kernels.h -- HIP kernels instantiated per user functor by the
 * msl::DA / DM / DC templates (the generic counterpart of the fixed-operation
 * kernels in libmsl_hip.so).
 *
 * Reference kernels: src/map_kernels.cu:36-101 (map, mapIndex), 140-168
 * (stencils), src/zip_kernels.cu:37-148 (zip, zipIndex),
 * src/fold_kernels.cu:51-129 (fold).  Differences: grid-stride loops sized to
 * the device instead of one 1024-thread block per 1024 elements; user
 * functors are called through their own class (f.F::operator()), so
 * polymorphic functor objects never dispatch through a host vtable; stencil
 * views stage their tile in LDS; all launches go to the library stream.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "../muesli.h"
#include "../nplmatrix.h"
#include "../plcube.h"
#include "../plmatrix.h"

namespace msl {
namespace detail {

constexpr int kThreads = 256;

inline int grid_for(size_t n, int per_cu = 8) {
    size_t want = (n + kThreads - 1) / kThreads;
    size_t cap = (size_t)num_cus() * per_cu;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

#define MSL_GRID_STRIDE(i, n) \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

// ---- fills and element-wise skeletons ---------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void fill_kernel(T *out, size_t n, T v) {
    MSL_GRID_STRIDE(i, n) out[i] = v;
}

// out[i] = f(in[i])
template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void map_kernel(R *out, const T *in, size_t n, F f) {
    MSL_GRID_STRIDE(i, n) out[i] = f.F::operator()(in[i]);
}

// out[i] = f(a[i], b[i])
template <typename A, typename B, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_kernel(R *out, const A *a, const B *b, size_t n, F f) {
    MSL_GRID_STRIDE(i, n) out[i] = f.F::operator()(a[i], b[i]);
}

// out[i] = f(a[i], b[i], c[i])
template <typename A, typename B, typename C, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip3_kernel(R *out, const A *a, const B *b, const C *c, size_t n, F f) {
    MSL_GRID_STRIDE(i, n) out[i] = f.F::operator()(a[i], b[i], c[i]);
}

// DA: out[i] = f(first + i, in[i])
template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void map_index_1d_kernel(R *out, const T *in, size_t n, int first, F f) {
    MSL_GRID_STRIDE(i, n) out[i] = f.F::operator()((int)i + first, in[i]);
}

// DA: out[i] = f(first + i, a[i], b[i])
template <typename A, typename B, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_index_1d_kernel(R *out, const A *a, const B *b, size_t n, int first,
                                                                F f) {
    MSL_GRID_STRIDE(i, n) out[i] = f.F::operator()((int)i + first, a[i], b[i]);
}

// DM: out[k] = f(row, col, in[k]), row = k / cols, col = k % cols
template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void map_index_2d_kernel(R *out, const T *in, size_t n, int cols, F f) {
    MSL_GRID_STRIDE(k, n) {
        const int row = (int)(k / cols), col = (int)(k % cols);
        out[k] = f.F::operator()(row, col, in[k]);
    }
}

template <typename A, typename B, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_index_2d_kernel(R *out, const A *a, const B *b, size_t n, int cols,
                                                                F f) {
    MSL_GRID_STRIDE(k, n) {
        const int row = (int)(k / cols), col = (int)(k % cols);
        out[k] = f.F::operator()(row, col, a[k], b[k]);
    }
}

// DC: out[k] = f(col, row, depth, in[k])
template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void map_index_3d_kernel(R *out, const T *in, size_t n, int cols, int rows,
                                                                F f) {
    MSL_GRID_STRIDE(k, n) {
        const size_t plane = (size_t)cols * rows;
        const int depth = (int)(k / plane);
        const size_t rem = k - (size_t)depth * plane;
        const int row = (int)(rem / cols), col = (int)(rem % cols);
        out[k] = f.F::operator()(col, row, depth, in[k]);
    }
}

template <typename A, typename B, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_index_3d_kernel(R *out, const A *a, const B *b, size_t n, int cols,
                                                                int rows, F f) {
    MSL_GRID_STRIDE(k, n) {
        const size_t plane = (size_t)cols * rows;
        const int depth = (int)(k / plane);
        const size_t rem = k - (size_t)depth * plane;
        const int row = (int)(rem / cols), col = (int)(rem % cols);
        out[k] = f.F::operator()(col, row, depth, a[k], b[k]);
    }
}

// ---- fold ----------------------------------------------------------------------
// Every thread folds its grid-stride elements in index order; threads and
// blocks then combine pairwise in a fixed tree.  No identity element is
// needed: an empty thread carries a "has no value" flag.
template <typename T, typename F>
__device__ __forceinline__ void fold_block(T acc, bool has, F &f, T *out, int *out_has) {
    __shared__ T sv[kThreads];
    __shared__ int sh[kThreads];
    sv[threadIdx.x] = acc;
    sh[threadIdx.x] = has ? 1 : 0;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < (unsigned)s) {
            const int o = threadIdx.x + s;
            if (sh[o]) {
                sv[threadIdx.x] = sh[threadIdx.x] ? f.F::operator()(sv[threadIdx.x], sv[o]) : sv[o];
                sh[threadIdx.x] = 1;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out = sv[0];
        *out_has = sh[0];
    }
}

template <typename T, typename F>
__global__ __launch_bounds__(kThreads) void fold_stage1_kernel(const T *in, size_t n, T *partials, int *has, F f) {
    T acc{};
    bool h = false;
    MSL_GRID_STRIDE(i, n) {
        acc = h ? f.F::operator()(acc, in[i]) : in[i];
        h = true;
    }
    fold_block(acc, h, f, partials + blockIdx.x, has + blockIdx.x);
}

template <typename T, typename F>
__global__ __launch_bounds__(kThreads) void fold_stage2_kernel(const T *partials, const int *has, int nparts, T *result,
                                                               int *result_has, F f) {
    T acc{};
    bool h = false;
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
        if (has[i]) {
            acc = h ? f.F::operator()(acc, partials[i]) : partials[i];
            h = true;
        }
    }
    fold_block(acc, h, f, result, result_has);
}

// ---- stencils ---------------------------------------------------------------------
// 2-D: a 256-thread block computes a 64 x 16 output tile (4 rows per thread)
// after staging the tile plus a halo of the stencil radius in LDS.
constexpr int kTileX = 64;
constexpr int kTileY = 16;

inline size_t stencil2d_lds_bytes(int r, size_t elem) { return (size_t)(kTileX + 2 * r) * (kTileY + 2 * r) * elem; }
constexpr size_t kMaxLdsBytes = 48 * 1024;

template <typename T>
__device__ __forceinline__ void stage_tile_2d(T *tile, const T *in, int rows, int cols, int gx0, int gy0, int tw,
                                              int th, T nv) {
    for (int idx = threadIdx.x; idx < tw * th; idx += kThreads) {
        const int lx = idx % tw, ly = idx / tw;
        const int gx = gx0 + lx, gy = gy0 + ly;
        tile[idx] = ((unsigned)gx < (unsigned)cols && (unsigned)gy < (unsigned)rows) ? in[(size_t)gy * cols + gx] : nv;
    }
}

template <typename T, T (*f)(const NPLMatrix<T> &, int, int)>
__global__ __launch_bounds__(kThreads) void dm_stencil_kernel(T *out, const T *in, int rows, int cols, int r, T nv,
                                                              int tiles_x, int use_lds) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T *tile = reinterpret_cast<T *>(smem_raw);
    const int bx = blockIdx.x % tiles_x, by = blockIdx.x / tiles_x;
    const int x0 = bx * kTileX, y0 = by * kTileY;
    NPLMatrix<T> view;
    view.width = cols;
    view.height = rows;
    view.stencilSize = r;
    view.neutralValue = nv;
    view.data = in;
    if (use_lds) {
        const int tw = kTileX + 2 * r, th = kTileY + 2 * r;
        stage_tile_2d(tile, in, rows, cols, x0 - r, y0 - r, tw, th, nv);
        __syncthreads();
        view.tile = tile;
        view.tileX = x0 - r;
        view.tileY = y0 - r;
        view.tileW = tw;
        view.tileH = th;
    }
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x = x0 + tx;
#pragma unroll
    for (int k = 0; k < kTileY / 4; k++) {
        const int y = y0 + ty + 4 * k;
        if (x < cols && y < rows) out[(size_t)y * cols + x] = f(view, x, y);
    }
}

// Legacy mapStencilMM: result(row, col) = f(row, col, &view, ncol, nrow).
template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void dm_stencil_mm_kernel(R *out, const T *in, int rows, int cols, int r, T nv,
                                                                 const T *border_data, int border, int tiles_x,
                                                                 int use_lds, F f) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T *tile = reinterpret_cast<T *>(smem_raw);
    const int bx = blockIdx.x % tiles_x, by = blockIdx.x / tiles_x;
    const int x0 = bx * kTileX, y0 = by * kTileY;
    PLMatrix<T> view;
    view.rows = rows;
    view.cols = cols;
    view.stencil_size = r;
    view.neutral_value = nv;
    view.data = in;
    view.border_data = border_data;
    view.border = border;
    if (use_lds) {
        const int tw = kTileX + 2 * r, th = kTileY + 2 * r;
        stage_tile_2d(tile, in, rows, cols, x0 - r, y0 - r, tw, th, nv);
        __syncthreads();
        view.tile = tile;
        view.tileRow = y0 - r;
        view.tileCol = x0 - r;
        view.tileRows = th;
        view.tileCols = tw;
    }
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int col = x0 + tx;
#pragma unroll
    for (int k = 0; k < kTileY / 4; k++) {
        const int row = y0 + ty + 4 * k;
        if (col < cols && row < rows) out[(size_t)row * cols + col] = f.F::operator()(row, col, &view, cols, rows);
    }
}

// Border band of width b around a rows x cols grid, value nvf(row, col), in a
// (rows+2b) x (cols+2b) buffer (interior cells are not read).
template <typename T, typename NVF>
__global__ __launch_bounds__(kThreads) void border_kernel(T *buf, int rows, int cols, int b, NVF nvf) {
    const size_t pw = (size_t)cols + 2 * b;
    const size_t n = pw * ((size_t)rows + 2 * b);
    MSL_GRID_STRIDE(i, n) {
        const int row = (int)(i / pw) - b, col = (int)(i % pw) - b;
        if (row < 0 || row >= rows || col < 0 || col >= cols) buf[i] = (T)nvf.NVF::operator()(row, col);
    }
}

// Cyclic rotation of a rows x cols matrix: out(i, j) = in(src_row(i), src_col(j))
// with src_row(i) = (i - row_shift) mod rows and src_col(j) = (j + col_shift)
// mod cols; both shifts are already reduced into [0, rows) and [0, cols).
template <typename T>
__global__ __launch_bounds__(kThreads) void rotate_kernel(T *out, const T *in, int rows, int cols, int row_shift,
                                                          int col_shift) {
    const size_t n = (size_t)rows * cols;
    MSL_GRID_STRIDE(k, n) {
        const int i = (int)(k / (size_t)cols), j = (int)(k % (size_t)cols);
        int si = i - row_shift;
        if (si < 0) si += rows;
        int sj = j + col_shift;
        if (sj >= cols) sj -= cols;
        out[k] = in[(size_t)si * cols + sj];
    }
}

// 3-D: a 256-thread block (32 x 8) computes a 32 x 8 x 8 output tile (8 z
// cells per thread) after staging the tile plus halo in LDS.
constexpr int kTile3X = 32;
constexpr int kTile3Y = 8;
constexpr int kTile3Z = 8;

inline size_t stencil3d_lds_bytes(int r, size_t elem) {
    return (size_t)(kTile3X + 2 * r) * (kTile3Y + 2 * r) * (kTile3Z + 2 * r) * elem;
}

template <typename T, T (*f)(const PLCube<T> &, int, int, int)>
__global__ __launch_bounds__(kThreads) void dc_stencil_kernel(T *out, const T *in, int cols, int rows, int depth,
                                                              int r, T nv, int tiles_x, int tiles_y, int use_lds) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T *tile = reinterpret_cast<T *>(smem_raw);
    const int bx = blockIdx.x % tiles_x;
    const int rest = blockIdx.x / tiles_x;
    const int by = rest % tiles_y, bz = rest / tiles_y;
    const int x0 = bx * kTile3X, y0 = by * kTile3Y, z0 = bz * kTile3Z;
    PLCube<T> view;
    view.width = cols;
    view.height = rows;
    view.depth = depth;
    view.stencilSize = r;
    view.neutralValue = nv;
    view.data = in;
    if (use_lds) {
        const int tw = kTile3X + 2 * r, th = kTile3Y + 2 * r, td = kTile3Z + 2 * r;
        for (int idx = threadIdx.x; idx < tw * th * td; idx += kThreads) {
            const int lx = idx % tw, ly = (idx / tw) % th, lz = idx / (tw * th);
            const int gx = x0 - r + lx, gy = y0 - r + ly, gz = z0 - r + lz;
            tile[idx] = ((unsigned)gx < (unsigned)cols && (unsigned)gy < (unsigned)rows && (unsigned)gz < (unsigned)depth)
                            ? in[((size_t)gz * rows + gy) * cols + gx]
                            : nv;
        }
        __syncthreads();
        view.tile = tile;
        view.tileX = x0 - r;
        view.tileY = y0 - r;
        view.tileZ = z0 - r;
        view.tileW = tw;
        view.tileH = th;
        view.tileD = td;
    }
    const int x = x0 + (threadIdx.x & 31), y = y0 + (threadIdx.x >> 5);
#pragma unroll
    for (int k = 0; k < kTile3Z; k++) {
        const int z = z0 + k;
        if (x < cols && y < rows && z < depth) out[((size_t)z * rows + y) * cols + x] = f(view, x, y, z);
    }
}

}  // namespace detail
}  // namespace msl

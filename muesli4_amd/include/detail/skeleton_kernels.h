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
#include <type_traits>

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

// Four consecutive elements of a 4-byte type, moved with one 16-byte access.
template <typename T>
struct alignas(16) Quad4 {
    T v[4];
};
template <typename T>
constexpr bool quad_type() {
    return sizeof(T) == 4 && std::is_trivially_copyable<T>::value;
}
inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

#define MSL_GRID_STRIDE_FROM(i, start, n)                                                          \
    for (size_t i = (start) + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (size_t)gridDim.x * blockDim.x)

// Grid of the vectorised streams (map, zip, fold): every lane keeps
// kStreamUnroll 16-byte accesses in flight, so `per_cu` 256-thread blocks
// per CU are enough to saturate HBM (one for map/zip, two for the read-only
// fold stream, as the library's fixed-operation kernels).
constexpr int kStreamUnroll = 4;
// Two input streams: two quads of each in flight, the same four loads as a
// map (the library's zip: 5.68 TB/s with four of each, 6.01 with two).
constexpr int kZipUnroll = 2;
// Folds: three quads in flight (the library's compensated fold: 5.67 TB/s
// with four, 5.78 with three); a lane's elements keep their order.
constexpr int kFoldUnroll = 3;
inline int stream_grid(size_t n, int per_cu) {
    const size_t per_block = (size_t)kThreads * 4 * kStreamUnroll;
    size_t want = (n + per_block - 1) / per_block;
    size_t cap = (size_t)num_cus() * per_cu;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

// Non-temporal 16-byte store of a quad of a 4-byte type: streamed outputs do
// not evict the input being read from the Infinity Cache.
template <typename R>
__device__ __forceinline__ void store_quad_nt(Quad4<R> *p, const Quad4<R> &o) {
    typedef unsigned v4u32 __attribute__((ext_vector_type(4)));
    v4u32 v;
    __builtin_memcpy(&v, &o, sizeof v);
    __builtin_nontemporal_store(v, reinterpret_cast<v4u32 *>(p));
}

// out[i] = f(in[i]); with quad (4-byte types, 16-byte aligned buffers) four
// elements per 16-byte access and kStreamUnroll quads in flight per lane,
// the remainder element-wise.
template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void map_kernel(R *out, const T *in, size_t n, F f, int quad) {
    size_t start = 0;
    if constexpr (quad_type<T>() && quad_type<R>()) {
        if (quad) {
            const size_t n4 = n / 4;
            const Quad4<T> *in4 = reinterpret_cast<const Quad4<T> *>(in);
            Quad4<R> *out4 = reinterpret_cast<Quad4<R> *>(out);
            const size_t stride = (size_t)gridDim.x * blockDim.x;
            size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
            for (; i + (kStreamUnroll - 1) * stride < n4; i += kStreamUnroll * stride) {
                Quad4<T> a[kStreamUnroll];
#pragma unroll
                for (int u = 0; u < kStreamUnroll; u++) a[u] = in4[i + u * stride];
#pragma unroll
                for (int u = 0; u < kStreamUnroll; u++) {
                    Quad4<R> o;
#pragma unroll
                    for (int k = 0; k < 4; k++) o.v[k] = f.F::operator()(a[u].v[k]);
                    store_quad_nt(out4 + i + u * stride, o);
                }
            }
            for (; i < n4; i += stride) {
                const Quad4<T> a = in4[i];
                Quad4<R> o;
#pragma unroll
                for (int k = 0; k < 4; k++) o.v[k] = f.F::operator()(a.v[k]);
                store_quad_nt(out4 + i, o);
            }
            start = n4 * 4;
        }
    }
    MSL_GRID_STRIDE_FROM(i, start, n) out[i] = f.F::operator()(in[i]);
}

// out[i] = f(a[i], b[i])
template <typename A, typename B, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_kernel(R *out, const A *a, const B *b, size_t n, F f, int quad) {
    size_t start = 0;
    if constexpr (quad_type<A>() && quad_type<B>() && quad_type<R>()) {
        if (quad) {
            const size_t n4 = n / 4;
            const Quad4<A> *a4 = reinterpret_cast<const Quad4<A> *>(a);
            const Quad4<B> *b4 = reinterpret_cast<const Quad4<B> *>(b);
            Quad4<R> *out4 = reinterpret_cast<Quad4<R> *>(out);
            const size_t stride = (size_t)gridDim.x * blockDim.x;
            size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
            for (; i + (kZipUnroll - 1) * stride < n4; i += kZipUnroll * stride) {
                Quad4<A> x[kZipUnroll];
                Quad4<B> y[kZipUnroll];
#pragma unroll
                for (int u = 0; u < kZipUnroll; u++) {
                    x[u] = a4[i + u * stride];
                    y[u] = b4[i + u * stride];
                }
#pragma unroll
                for (int u = 0; u < kZipUnroll; u++) {
                    Quad4<R> o;
#pragma unroll
                    for (int k = 0; k < 4; k++) o.v[k] = f.F::operator()(x[u].v[k], y[u].v[k]);
                    store_quad_nt(out4 + i + u * stride, o);
                }
            }
            for (; i < n4; i += stride) {
                const Quad4<A> x = a4[i];
                const Quad4<B> y = b4[i];
                Quad4<R> o;
#pragma unroll
                for (int k = 0; k < 4; k++) o.v[k] = f.F::operator()(x.v[k], y.v[k]);
                store_quad_nt(out4 + i, o);
            }
            start = n4 * 4;
        }
    }
    MSL_GRID_STRIDE_FROM(i, start, n) out[i] = f.F::operator()(a[i], b[i]);
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

// ---- DM index skeletons with extra operands (reference src/dm.cpp:1757-1911,
// src/zip_kernels.cu:46-99): per-element operands index by k, per-row
// operands (a DA of one value per row) by the row.
// out[k] = f(row, col, a[k], b[k], c[k])
template <typename A, typename B, typename C, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip3_index_2d_kernel(R *out, const A *a, const B *b, const C *c, size_t n,
                                                                 int cols, F f) {
    MSL_GRID_STRIDE(k, n) {
        const int row = (int)(k / cols), col = (int)(k % cols);
        out[k] = f.F::operator()(row, col, a[k], b[k], c[k]);
    }
}

// out[k] = f(row, col, a[k], b[k], v[row])
template <typename A, typename B, typename V, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_index_2d_row_kernel(R *out, const A *a, const B *b, const V *v,
                                                                    size_t n, int cols, F f) {
    MSL_GRID_STRIDE(k, n) {
        const int row = (int)(k / cols), col = (int)(k % cols);
        out[k] = f.F::operator()(row, col, a[k], b[k], v[row]);
    }
}

// out[k] = f(row, col, a[k], v[row])
template <typename A, typename V, typename R, typename F>
__global__ __launch_bounds__(kThreads) void map_index_2d_row_kernel(R *out, const A *a, const V *v, size_t n, int cols,
                                                                    F f) {
    MSL_GRID_STRIDE(k, n) {
        const int row = (int)(k / cols), col = (int)(k % cols);
        out[k] = f.F::operator()(row, col, a[k], v[row]);
    }
}

// out[k] = f(row, col, a[k], u[row], v[row], w[row])
template <typename A, typename U, typename V, typename W, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_index_2d_rows3_kernel(R *out, const A *a, const U *u, const V *v,
                                                                      const W *w, size_t n, int cols, F f) {
    MSL_GRID_STRIDE(k, n) {
        const int row = (int)(k / cols), col = (int)(k % cols);
        out[k] = f.F::operator()(row, col, a[k], u[row], v[row], w[row]);
    }
}

// out[k] = f(row, col, a[k], whole) -- the functor sees the whole second
// matrix through its device pointer
template <typename A, typename B, typename R, typename F>
__global__ __launch_bounds__(kThreads) void cross_zip_index_2d_kernel(R *out, const A *a, B *whole, size_t n, int cols,
                                                                      F f) {
    MSL_GRID_STRIDE(k, n) {
        const int row = (int)(k / cols), col = (int)(k % cols);
        out[k] = f.F::operator()(row, col, a[k], whole);
    }
}

// out[row] = f(...f(f(in[row][0], in[row][1]), in[row][2])..., in[row][cols-1])
// in a fixed tree: one wave per row, lane l folds columns l, l+64, ... in
// order, then lane 0 folds the 64 lane results in lane order.  Any T (the
// lane results go through LDS, not shuffles).
template <typename T, typename F>
__global__ __launch_bounds__(kThreads) void reduce_rows_kernel(T *out, const T *in, int rows, int cols, F f) {
    __shared__ __attribute__((aligned(16))) unsigned char part_raw[kThreads * sizeof(T)];
    __shared__ int has[kThreads];
    T *part = reinterpret_cast<T *>(part_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = blockIdx.x * (kThreads / 64) + wave;
    if (row < rows) {  // uniform per wave
        const T *p = in + (size_t)row * cols;
        T acc{};
        int any = 0;
        for (int c = lane; c < cols; c += 64) {
            acc = any ? f.F::operator()(acc, p[c]) : p[c];
            any = 1;
        }
        part[threadIdx.x] = acc;
        has[threadIdx.x] = any;
    }
    __syncthreads();
    if (row < rows && lane == 0) {
        T acc = part[threadIdx.x];
        for (int l = 1; l < 64; l++)
            if (has[threadIdx.x + l]) acc = f.F::operator()(acc, part[threadIdx.x + l]);
        out[row] = acc;
    }
}

// out[row] = f(in[row][cols-1], in[row][0], ... f(in[row][1], in[row][2],
// f(in[row][0], in[row][1], R(0)))) -- the reference's sequential order
// (src/dm.cpp:1731-1754), one thread per row.
template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void reduce_two_columns_kernel(R *out, const T *in, int rows, int cols, F f) {
    MSL_GRID_STRIDE(row, (size_t)rows) {
        const T *p = in + row * (size_t)cols;
        R acc = 0;
        for (int j = 0; j < cols; j++) acc = f.F::operator()(p[j], p[(j + 1) % cols], acc);
        out[row] = acc;
    }
}

// m[row][col] = v[row]
template <typename T>
__global__ __launch_bounds__(kThreads) void set_column_kernel(T *m, const T *v, int rows, int cols, int col) {
    MSL_GRID_STRIDE(row, (size_t)rows) m[row * (size_t)cols + col] = v[row];
}

// DA with a matrix of rows (reference src/da.cpp:256-324,
// src/map_kernels.cu:80-86, src/zip_kernels.cu:73-79):
// out[i] = f(i, &m[i * mcols], in[i])
template <typename T, typename M, typename R, typename F>
__global__ __launch_bounds__(kThreads) void map_index_dmrows_kernel(R *out, const T *in, M *m, size_t n, int mcols,
                                                                    F f) {
    MSL_GRID_STRIDE(i, n) out[i] = f.F::operator()((int)i, m + i * (size_t)mcols, in[i]);
}

// out[i] = f(i, &m[i * mcols], b[i], in[i])
template <typename T, typename B, typename M, typename R, typename F>
__global__ __launch_bounds__(kThreads) void zip_index_dmrows_kernel(R *out, const T *in, const B *b, M *m, size_t n,
                                                                    int mcols, F f) {
    MSL_GRID_STRIDE(i, n) out[i] = f.F::operator()((int)i, m + i * (size_t)mcols, b[i], in[i]);
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
// Every thread folds its grid-stride elements in index order; the lanes of a
// wave then combine in a fixed pairwise tree by shuffles (lane l takes lane
// l+m at m = 1, 2, 4, ..., so lane 0 ends with the in-order tree), and the
// four wave results combine in wave order.  No identity element is needed:
// an empty thread carries a "has no value" flag.  Types whose size is not a
// multiple of 4 bytes combine through LDS in the same tree instead.
template <typename T>
constexpr bool shuffle_type() {
    return sizeof(T) % 4 == 0 && sizeof(T) <= 64 && std::is_trivially_copyable<T>::value;
}

template <typename T>
__device__ __forceinline__ T shfl_down_words(const T &v, int m) {
    constexpr int W = sizeof(T) / 4;
    int w[W];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int j = 0; j < W; j++) w[j] = __shfl_down(w[j], m);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

template <typename T, typename F>
__device__ __forceinline__ void fold_block(T acc, bool has, F &f, T *out, int *out_has) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ __attribute__((aligned(16))) unsigned char wv_raw[(kThreads / 64) * sizeof(T)];
    __shared__ int wh[kThreads / 64];
    T *wv = reinterpret_cast<T *>(wv_raw);
    int h = has ? 1 : 0;
    if constexpr (shuffle_type<T>()) {
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const T o = shfl_down_words(acc, m);
            const int oh = __shfl_down(h, m);
            if ((lane & (2 * m - 1)) == 0 && lane + m < 64 && oh) {
                acc = h ? f.F::operator()(acc, o) : o;
                h = 1;
            }
        }
    } else {
        __shared__ __attribute__((aligned(16))) unsigned char sv_raw[kThreads * sizeof(T)];
        __shared__ int sh[kThreads];
        T *sv = reinterpret_cast<T *>(sv_raw);
        sv[threadIdx.x] = acc;
        sh[threadIdx.x] = h;
        __syncthreads();
        for (int m = 1; m < 64; m <<= 1) {
            if ((lane & (2 * m - 1)) == 0 && sh[threadIdx.x + m]) {
                sv[threadIdx.x] = sh[threadIdx.x] ? f.F::operator()(sv[threadIdx.x], sv[threadIdx.x + m])
                                                  : sv[threadIdx.x + m];
                sh[threadIdx.x] = 1;
            }
            __syncthreads();
        }
        acc = sv[threadIdx.x];
        h = sh[threadIdx.x];
    }
    if (lane == 0) {
        wv[wave] = acc;
        wh[wave] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        T r = wv[0];
        int rh = wh[0];
        for (int w = 1; w < kThreads / 64; w++) {
            if (wh[w]) {
                r = rh ? f.F::operator()(r, wv[w]) : wv[w];
                rh = 1;
            }
        }
        *out = r;
        *out_has = rh;
    }
}

// Stage 1 of one block: kStreamUnroll 16-byte loads in flight per lane
// (4-byte types on aligned buffers), then the element tail; the block's
// partial goes to *part / *part_has.  A thread with at least one quad starts
// its accumulator from the first quad's first element, so the unrolled loop
// folds without a "has a value" flag.
template <typename T, typename F>
__device__ __forceinline__ void fold_stage1_block(const T *in, size_t n, T *part, int *part_has, F &f, int quad) {
    T acc{};
    bool h = false;
    size_t start = 0;
    if constexpr (quad_type<T>()) {
        if (quad) {  // this thread's quads in grid-stride order, then its tail elements
            const size_t n4 = n / 4;
            const Quad4<T> *in4 = reinterpret_cast<const Quad4<T> *>(in);
            const size_t stride = (size_t)gridDim.x * blockDim.x;
            size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
            if (i < n4) {
                const Quad4<T> a = in4[i];
                acc = a.v[0];
#pragma unroll
                for (int k = 1; k < 4; k++) acc = f.F::operator()(acc, a.v[k]);
                h = true;
                i += stride;
                for (; i + (kFoldUnroll - 1) * stride < n4; i += kFoldUnroll * stride) {
                    Quad4<T> q[kFoldUnroll];
#pragma unroll
                    for (int u = 0; u < kFoldUnroll; u++) q[u] = in4[i + u * stride];
#pragma unroll
                    for (int u = 0; u < kFoldUnroll; u++)
#pragma unroll
                        for (int k = 0; k < 4; k++) acc = f.F::operator()(acc, q[u].v[k]);
                }
                for (; i < n4; i += stride) {
                    const Quad4<T> q = in4[i];
#pragma unroll
                    for (int k = 0; k < 4; k++) acc = f.F::operator()(acc, q.v[k]);
                }
            }
            start = n4 * 4;
        }
    }
    MSL_GRID_STRIDE_FROM(i, start, n) {
        acc = h ? f.F::operator()(acc, in[i]) : in[i];
        h = true;
    }
    fold_block(acc, h, f, part, part_has);
}

// Stage 2 of one block: the nparts partials in index order per thread, then
// the block tree.
template <typename T, typename F>
__device__ __forceinline__ void fold_stage2_block(const T *partials, const int *has, int nparts, T *result,
                                                  int *result_has, F &f) {
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

template <typename T, typename F>
__global__ __launch_bounds__(kThreads) void fold_stage1_kernel(const T *in, size_t n, T *partials, int *has, F f,
                                                               int quad) {
    fold_stage1_block(in, n, partials + blockIdx.x, has + blockIdx.x, f, quad);
}

// *result may be host memory mapped for the device (the result slot).
template <typename T, typename F>
__global__ __launch_bounds__(kThreads) void fold_stage2_kernel(const T *partials, const int *has, int nparts, T *result,
                                                               int *result_has, F f) {
    fold_stage2_block(partials, has, nparts, result, result_has, f);
}

// ---- stencils ---------------------------------------------------------------------
// 2-D march: a 256-thread block owns a strip of kMarchW columns (one output
// column per thread) and a segment of rows.  At each step it stages kMarchK
// input rows of the window (the strip plus r halo columns each side, with
// the outside value for cells off the grid) into an LDS ring of
// H >= 2r + 2*kMarchK rows, the next step's rows being requested in
// registers meanwhile; after one barrier every thread computes the kMarchK
// output rows whose 2r+1 window rows are now staged.  The ring is deep
// enough that the rows staged at the next step never overwrite a row that
// the current step reads, so one barrier per step suffices.
constexpr int kMarchW = 256;
constexpr int kMarchK = 4;

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// waits for every outstanding global load of the wave, which would drain the
// rows (planes) that the marches request a step ahead at every step; the
// "local"-only fences make the barrier wait for LDS operations alone.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
constexpr size_t kMaxLdsBytes = 48 * 1024;

constexpr int pow2_at_least(int v) {
    int h = 1;
    while (h < v) h <<= 1;
    return h;
}

constexpr int march2d_ring_rows(int r) { return pow2_at_least(2 * r + 2 * kMarchK); }

// The generic stencil kernels take the radius as a template argument RC for
// the common small radii (RC = 0: at run time).  With a constant radius the
// view's window geometry is constant too, and the compiler folds the
// per-access window checks of constant-offset reads.

inline size_t march2d_lds_bytes(int r, size_t elem) {
    return (size_t)march2d_ring_rows(r) * (size_t)(kMarchW + 2 * r) * elem;
}

struct March2D {
    int strips_x, seg, blocks;
};

// Blocks per CU the 2-D march aims at (r = 1 blur functor on 8192^2: 2
// blocks 3.3 TB/s, 4 4.4, 6 4.6, 8 4.3).
constexpr int kMarch2DBlocksPerCu = 6;

inline March2D march2d_shape(int rows, int cols) {
    March2D m;
    m.strips_x = (cols + kMarchW - 1) / kMarchW;
    const long target = (long)num_cus() * kMarch2DBlocksPerCu;
    long segs = (target + m.strips_x - 1) / m.strips_x;
    if (segs < 1) segs = 1;
    int seg = (int)((rows + segs - 1) / segs);
    if (seg < kMarchK) seg = kMarchK;
    m.seg = seg;
    m.blocks = m.strips_x * ((rows + seg - 1) / seg);
    return m;
}

// Runs the march of output rows [y0, y1) of the strip at column xs and calls
// out_row(yo) once the window rows yo-r .. yo+r are staged.  outside(row, col)
// gives the staged value of cells off the grid.
template <typename T, typename Outside, typename OutRow>
__device__ __forceinline__ void march2d(const T *__restrict__ in, int rows, int cols, int r, int xs, int y0, int y1,
                                        T *ring, int mask, Outside outside, OutRow out_row) {
    const int t = threadIdx.x;
    const int W = kMarchW + 2 * r;
    const int xa = xs - r + t;
    const bool hasb = t < 2 * r;  // this thread also stages column xa + kMarchW
    const int xb = hasb ? xa + kMarchW : xa;
    auto load = [&](int yy, int x) -> T {
        const int yc = min(max(yy, 0), rows - 1), xc = min(max(x, 0), cols - 1);
        return in[(size_t)yc * cols + xc];
    };
    auto fix = [&](const T &v, int yy, int x) -> T {
        return ((unsigned)yy < (unsigned)rows && (unsigned)x < (unsigned)cols) ? v : outside(yy, x);
    };
    const int total = (y1 - y0) + 2 * r;  // rows to stage: y0-r .. y1-1+r
    T qa[kMarchK], qb[kMarchK];
#pragma unroll
    for (int k = 0; k < kMarchK; k++) {
        qa[k] = load(y0 - r + k, xa);
        qb[k] = load(y0 - r + k, xb);
    }
    for (int sk = 0; sk < total; sk += kMarchK) {  // uniform across the block
        const int a = y0 - r + sk;                  // first row staged at this step
#pragma unroll
        for (int k = 0; k < kMarchK; k++) {
            const int yy = a + k;
            T *rowp = ring + (yy & mask) * W;
            rowp[t] = fix(qa[k], yy, xa);
            if (hasb) rowp[t + kMarchW] = fix(qb[k], yy, xb);
            qa[k] = load(yy + kMarchK, xa);
            qb[k] = load(yy + kMarchK, xb);
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < kMarchK; k++) {
            const int yo = a - r + k;  // its window rows a-2r+k .. a+k are staged
            if (yo >= y0 && yo < y1) out_row(yo);
        }
    }
}

template <typename T, T (*f)(const NPLMatrix<T> &, int, int), int RC = 0>
__global__ __launch_bounds__(kThreads) void dm_stencil_kernel(T *out, const T *in, int rows, int cols, int r_arg, T nv,
                                                              int strips_x, int seg, int mask_arg) {
    const int r = RC > 0 ? RC : r_arg;
    const int mask = RC > 0 ? march2d_ring_rows(RC) - 1 : mask_arg;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    __builtin_assume(__builtin_amdgcn_workitem_id_x() < (unsigned)kThreads);  // lets window checks fold
    T *ring = reinterpret_cast<T *>(smem_raw);
    const int xs = (blockIdx.x % strips_x) * kMarchW;
    const int y0 = (blockIdx.x / strips_x) * seg;
    const int y1 = min(y0 + seg, rows);
    NPLMatrix<T> view;
    view.width = cols;
    view.height = rows;
    view.stencilSize = r;
    view.neutralValue = nv;
    view.data = in;
    view.tile = ring;
    const int x = xs + threadIdx.x;
    // window origin xs - r, relative to this thread's column (in-window
    // checks of reads within the radius then fold, see dc_stencil_kernel)
    view.tileX = x - ((int)threadIdx.x + r);
    view.tileW = kMarchW + 2 * r;
    view.tileH = 2 * r + 1;
    view.rowMask = mask;
    // The callbacks capture the view by value (a per-row copy): capturing it
    // by reference miscompiles for gfx950 once a large user function is
    // inlined ("Illegal instruction detected").
    march2d(
        in, rows, cols, r, xs, y0, y1, ring, mask, [=](int, int) { return nv; },
        [=](int yo) {
            NPLMatrix<T> v = view;
            v.tileY = yo - r;
            if (x < cols) out[(size_t)yo * cols + x] = f(v, x, yo);
        });
}

// Without LDS (window too large): every output reads through the view.
template <typename T, T (*f)(const NPLMatrix<T> &, int, int)>
__global__ __launch_bounds__(kThreads) void dm_stencil_global_kernel(T *out, const T *in, int rows, int cols, int r,
                                                                     T nv) {
    NPLMatrix<T> view;
    view.width = cols;
    view.height = rows;
    view.stencilSize = r;
    view.neutralValue = nv;
    view.data = in;
    MSL_GRID_STRIDE(k, (size_t)rows * cols) {
        const int y = (int)(k / (size_t)cols), x = (int)(k % (size_t)cols);
        out[k] = f(view, x, y);
    }
}

// Legacy mapStencilMM: result(row, col) = f(row, col, &view, ncol, nrow).
// kBorder: cells off the grid come from the border band (border_data) where
// it covers them; without one they are nv.
template <typename T, typename R, typename F, int RC = 0, bool kBorder = true>
__global__ __launch_bounds__(kThreads) void dm_stencil_mm_kernel(R *out, const T *in, int rows, int cols, int r_arg,
                                                                 T nv, const T *border_data, int border, int strips_x,
                                                                 int seg, int mask_arg, F f) {
    const int r = RC > 0 ? RC : r_arg;
    const int mask = RC > 0 ? march2d_ring_rows(RC) - 1 : mask_arg;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    __builtin_assume(__builtin_amdgcn_workitem_id_x() < (unsigned)kThreads);  // lets window checks fold
    T *ring = reinterpret_cast<T *>(smem_raw);
    const int xs = (blockIdx.x % strips_x) * kMarchW;
    const int y0 = (blockIdx.x / strips_x) * seg;
    const int y1 = min(y0 + seg, rows);
    PLMatrix<T> view;
    view.rows = rows;
    view.cols = cols;
    view.stencil_size = r;
    view.neutral_value = nv;
    view.data = in;
    view.border_data = border_data;
    view.border = border;
    view.tile = ring;
    const int col = xs + threadIdx.x;
    view.tileCol = col - ((int)threadIdx.x + r);  // xs - r, relative to the thread's column
    view.tileCols = kMarchW + 2 * r;
    view.tileRows = 2 * r + 1;
    view.rowMask = mask;
    march2d(
        in, rows, cols, r, xs, y0, y1, ring, mask,
        [=](int row, int c) {
            if constexpr (kBorder)
                return view.outside(row, c);
            else
                return nv;
        },
        [=, &f](int row) {
            PLMatrix<T> v = view;
            v.tileRow = row - r;
            if (col < cols) out[(size_t)row * cols + col] = f.F::operator()(row, col, &v, cols, rows);
        });
}

template <typename T, typename R, typename F>
__global__ __launch_bounds__(kThreads) void dm_stencil_mm_global_kernel(R *out, const T *in, int rows, int cols, int r,
                                                                        T nv, const T *border_data, int border, F f) {
    PLMatrix<T> view;
    view.rows = rows;
    view.cols = cols;
    view.stencil_size = r;
    view.neutral_value = nv;
    view.data = in;
    view.border_data = border_data;
    view.border = border;
    MSL_GRID_STRIDE(k, (size_t)rows * cols) {
        const int row = (int)(k / (size_t)cols), col = (int)(k % (size_t)cols);
        out[k] = f.F::operator()(row, col, &view, cols, rows);
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

// 3-D march: a 256-thread block owns a 64 x 16 column tile of the cube (four
// outputs per thread per plane) and a segment of planes.  At each step it
// stages kMarch3K planes of the window (tile plus r halo cells in x and y, nv
// off the cube) into an LDS ring of H >= 2r + 2*kMarch3K planes, the next
// step's planes being requested in registers meanwhile; after one barrier it
// computes the kMarch3K output planes whose 2r+1 window planes are staged.
constexpr int kM3X = 64;
// 16 rows per tile: four outputs per thread share one staged window plane
// of (64+2r) x (16+2r) cells (heat3D functor on 512^3: 3.8 TB/s with 8
// rows, 4.3 with 16; staging four planes per barrier instead of two lost
// a third, as the deeper LDS ring leaves fewer blocks per CU)
constexpr int kM3Y = 16;
constexpr int kMarch3K = 2;
constexpr int kM3Stage = 7;  // staged cells per thread per plane (r <= 3)

constexpr int march3d_ring_planes(int r) { return pow2_at_least(2 * r + 2 * kMarch3K); }

inline size_t march3d_lds_bytes(int r, size_t elem) {
    const size_t plane = (size_t)(kM3X + 2 * r) * (kM3Y + 2 * r);
    if (plane > (size_t)kM3Stage * kThreads) return (size_t)-1;
    return (size_t)march3d_ring_planes(r) * plane * elem;
}

struct March3D {
    int xtiles, ytiles, zseg, blocks;
};

// Tiles and z-segments of the 3-D march.  The block count aims at one full
// set of resident blocks: as many per CU as the LDS ring (lds bytes per
// block, 160 KiB per CU) and the waves (8 blocks of 4 waves) allow, rounded
// down to whole z-segment layers, so no partial second round of blocks
// trails the launch (measured on the 512^3 heat functor: 4 blocks per CU
// 3.0 TB/s, 6 of the 7 resident 3.75, 8 3.4).
inline March3D march3d_shape(int cols, int rows, int depth, size_t lds) {
    March3D m;
    m.xtiles = (cols + kM3X - 1) / kM3X;
    m.ytiles = (rows + kM3Y - 1) / kM3Y;
    const long cols2 = (long)m.xtiles * m.ytiles;
    long per_cu = lds > 0 ? (long)((160u << 10) / lds) : 8;
    if (per_cu > 8) per_cu = 8;
    if (per_cu < 1) per_cu = 1;
    const long target = (long)num_cus() * per_cu;
    long zsegs = target / cols2;
    if (zsegs < 1) zsegs = 1;
    int zseg = (int)((depth + zsegs - 1) / zsegs);
    if (zseg < 4) zseg = 4;
    m.zseg = zseg;
    m.blocks = m.xtiles * m.ytiles * ((depth + zseg - 1) / zseg);
    return m;
}

template <typename T, T (*f)(const PLCube<T> &, int, int, int), int RC = 0>
__global__ __launch_bounds__(kThreads) void dc_stencil_kernel(T *out, const T *in, int cols, int rows, int depth,
                                                              int r_arg, T nv, int xtiles, int ytiles, int zseg,
                                                              int mask_arg) {
    const int r = RC > 0 ? RC : r_arg;
    const int mask = RC > 0 ? march3d_ring_planes(RC) - 1 : mask_arg;
    // staged cells per thread per plane: just enough for the window plane
    // when the radius is known (3 at r = 1), else the maximum
    constexpr int NS = RC > 0 ? ((kM3X + 2 * RC) * (kM3Y + 2 * RC) + kThreads - 1) / kThreads : kM3Stage;
    static_assert(NS <= kM3Stage, "window plane larger than the staging slots");
    __builtin_assume(__builtin_amdgcn_workitem_id_x() < (unsigned)kThreads);  // lets window checks fold
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T *ring = reinterpret_cast<T *>(smem_raw);
    const int bx = blockIdx.x % xtiles;
    const int rest = blockIdx.x / xtiles;
    const int x0 = bx * kM3X, y0 = (rest % ytiles) * kM3Y;
    const int z0 = (rest / ytiles) * zseg, z1 = min(z0 + zseg, depth);
    const int WX = kM3X + 2 * r, WY = kM3Y + 2 * r, WE = WX * WY;
    const int t = threadIdx.x;
    const size_t plane_len = (size_t)rows * cols;
    // staged cells of this thread: e = t + k*256 in the window; their clamped
    // in-plane offsets and x-y validity do not change from plane to plane, so
    // a staging load is one uniform plane base plus a per-thread offset
    size_t soff[NS];
    bool sxy[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int e = min(t + k * kThreads, WE - 1);
        const int ex = x0 - r + e % WX;
        const int ey = y0 - r + e / WX;
        sxy[k] = (unsigned)ex < (unsigned)cols && (unsigned)ey < (unsigned)rows;
        soff[k] = (size_t)min(max(ey, 0), rows - 1) * cols + min(max(ex, 0), cols - 1);
    }
    auto load = [&](int k, int z) -> T {
        const int zc = min(max(z, 0), depth - 1);
        return in[(size_t)zc * plane_len + soff[k]];
    };
    PLCube<T> view;
    view.width = cols;
    view.height = rows;
    view.depth = depth;
    view.stencilSize = r;
    view.neutralValue = nv;
    view.data = in;
    view.tile = ring;
    const int ox = x0 + (t % kM3X), oy = y0 + t / kM3X;
    // window origin x0 - r, y0 - r, written relative to this thread's own
    // output cell: a read at (ox + dx, oy + dy) then has the window offset
    // (t % kM3X) + r + dx, whose range the compiler can see, so in-window
    // checks of reads within the radius fold away
    view.tileX = ox - ((t % kM3X) + r);
    view.tileY = oy - ((t / kM3X) + r);
    view.tileW = WX;
    view.tileH = WY;
    view.tileD = 2 * r + 1;
    view.planeMask = mask;
    // this thread's outputs: rows oy + 4j, column ox, in-plane offsets fixed
    size_t ooff[kM3Y / 4];
#pragma unroll
    for (int j = 0; j < kM3Y / 4; j++) ooff[j] = (size_t)(oy + 4 * j) * cols + ox;
    const int total = (z1 - z0) + 2 * r;
    T q[kMarch3K][NS];
#pragma unroll
    for (int d = 0; d < kMarch3K; d++)
#pragma unroll
        for (int k = 0; k < NS; k++) q[d][k] = load(k, z0 - r + d);
    for (int sk = 0; sk < total; sk += kMarch3K) {  // uniform across the block
        const int a = z0 - r + sk;                   // first plane staged at this step
#pragma unroll
        for (int d = 0; d < kMarch3K; d++) {
            const int zz = a + d;
            T *plane = ring + (zz & mask) * WE;
#pragma unroll
            for (int k = 0; k < NS; k++) {
                const int e = t + k * kThreads;
                const bool in_cube = sxy[k] && (unsigned)zz < (unsigned)depth;
                if (e < WE) plane[e] = in_cube ? q[d][k] : nv;
                q[d][k] = load(k, zz + kMarch3K);
            }
        }
        lds_barrier();
#pragma unroll
        for (int d = 0; d < kMarch3K; d++) {
            const int zo = a - r + d;  // its window planes a-2r+d .. a+d are staged
            if (zo >= z0 && zo < z1) {
                PLCube<T> v = view;
                v.tileZ = zo - r;
                T *oplane = out + (size_t)zo * plane_len;
#pragma unroll
                for (int j = 0; j < kM3Y / 4; j++) {
                    const int y = oy + 4 * j;
                    if (ox < cols && y < rows) oplane[ooff[j]] = f(v, ox, y, zo);
                }
            }
        }
    }
}

template <typename T, T (*f)(const PLCube<T> &, int, int, int)>
__global__ __launch_bounds__(kThreads) void dc_stencil_global_kernel(T *out, const T *in, int cols, int rows,
                                                                     int depth, int r, T nv) {
    PLCube<T> view;
    view.width = cols;
    view.height = rows;
    view.depth = depth;
    view.stencilSize = r;
    view.neutralValue = nv;
    view.data = in;
    MSL_GRID_STRIDE(k, (size_t)cols * rows * depth) {
        const size_t plane = (size_t)cols * rows;
        const int z = (int)(k / plane);
        const size_t rem = k - (size_t)z * plane;
        out[k] = f(view, (int)(rem % cols), (int)(rem / cols), z);
    }
}

}  // namespace detail
}  // namespace msl

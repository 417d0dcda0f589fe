// Elementwise skeletons: fills, map, zip, mapIndex.
//
// Reference kernels launch one element per thread in 1024-thread blocks with
// scalar 4-byte accesses (src/map_kernels.cu:36-43, src/zip_kernels.cu:37-43).
// Here every stream is a grid-stride loop over 16-byte vectors with four
// vectors in flight per thread, sized to fill all CUs once; each element is
// read once and written once (8 B/element for map, 12 B/element for zip).
#include "common.hpp"

namespace mslhip {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef int32_t v4i __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kUnroll = 4;      // 16-byte accesses in flight per lane, one input
constexpr int kZipUnroll = 2;   // two inputs: the same four loads in flight
                                // (8192^2 zip: 5.68 TB/s with 4 per input, 6.01 with 2)

// Plain loads and non-temporal stores: measured fastest on MI355X for
// streams (the non-temporal store keeps the output from evicting the input
// from the 256 MiB Infinity Cache; non-temporal loads were slower).
template <typename V>
__device__ __forceinline__ V ldg(const V *p) {
    return *p;
}
template <typename V>
__device__ __forceinline__ void stg(V *p, V v) {
    __builtin_nontemporal_store(v, p);
}

// ---- generic vectorised 1-in / 2-in / 0-in streams -------------------------
// Op provides:  S apply(S) / apply(S,S) for the scalar and V for vectors.
template <typename V, typename S, typename Op>
__global__ __launch_bounds__(kThreads) void stream1_kernel(S *__restrict__ out, const S *__restrict__ in, size_t n,
                                                           Op op) {
    const size_t n4 = n / 4;
    const V *in4 = reinterpret_cast<const V *>(in);
    V *out4 = reinterpret_cast<V *>(out);
    const size_t stride = (size_t)gridDim.x * kThreads;
    size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride) {
        V v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = ldg(in4 + i + u * stride);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) stg(out4 + i + u * stride, op.vec(v[u]));
    }
    for (; i < n4; i += stride) stg(out4 + i, op.vec(ldg(in4 + i)));
    // scalar tail (n % 4 elements)
    size_t t = n4 * 4 + (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (t < n && blockIdx.x == 0) out[t] = op.scalar(in[t]);
}

template <typename S, typename Op>
__global__ __launch_bounds__(kThreads) void stream1_scalar_kernel(S *__restrict__ out, const S *__restrict__ in,
                                                                  size_t n, Op op) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) out[i] = op.scalar(in[i]);
}

template <typename V, typename S, typename Op>
__global__ __launch_bounds__(kThreads) void stream2_kernel(S *__restrict__ out, const S *__restrict__ a,
                                                           const S *__restrict__ b, size_t n, Op op) {
    const size_t n4 = n / 4;
    const V *a4 = reinterpret_cast<const V *>(a);
    const V *b4 = reinterpret_cast<const V *>(b);
    V *out4 = reinterpret_cast<V *>(out);
    const size_t stride = (size_t)gridDim.x * kThreads;
    size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    for (; i + (kZipUnroll - 1) * stride < n4; i += kZipUnroll * stride) {
        V va[kZipUnroll], vb[kZipUnroll];
#pragma unroll
        for (int u = 0; u < kZipUnroll; u++) {
            va[u] = ldg(a4 + i + u * stride);
            vb[u] = ldg(b4 + i + u * stride);
        }
#pragma unroll
        for (int u = 0; u < kZipUnroll; u++) stg(out4 + i + u * stride, op.vec(va[u], vb[u]));
    }
    for (; i < n4; i += stride) stg(out4 + i, op.vec(ldg(a4 + i), ldg(b4 + i)));
    size_t t = n4 * 4 + (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (t < n && blockIdx.x == 0) out[t] = op.scalar(a[t], b[t]);
}

template <typename S, typename Op>
__global__ __launch_bounds__(kThreads) void stream2_scalar_kernel(S *__restrict__ out, const S *__restrict__ a,
                                                                  const S *__restrict__ b, size_t n, Op op) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) out[i] = op.scalar(a[i], b[i]);
}

// Generator streams: out[i] = op(i).
template <typename V, typename S, typename Op>
__global__ __launch_bounds__(kThreads) void gen_kernel(S *__restrict__ out, size_t n, Op op) {
    const size_t n4 = n / 4;
    V *out4 = reinterpret_cast<V *>(out);
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride) {
        V v;
        v.x = op(4 * i + 0);
        v.y = op(4 * i + 1);
        v.z = op(4 * i + 2);
        v.w = op(4 * i + 3);
        stg(out4 + i, v);
    }
    size_t t = n4 * 4 + (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (t < n && blockIdx.x == 0) out[t] = op(t);
}

template <typename S, typename Op>
__global__ __launch_bounds__(kThreads) void gen_scalar_kernel(S *__restrict__ out, size_t n, Op op) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) out[i] = op(i);
}

// ---- element functions ---------------------------------------------------
struct AffineF32 {
    float a, b;
    __device__ float scalar(float x) const { return add_rn(mul_rn(x, a), b); }
    __device__ v4f vec(v4f x) const {
        v4f r;
        r.x = scalar(x.x);
        r.y = scalar(x.y);
        r.z = scalar(x.z);
        r.w = scalar(x.w);
        return r;
    }
};

struct AffineU32 {
    uint32_t a, b;
    __device__ uint32_t scalar(uint32_t x) const { return x * a + b; }
    __device__ v4u vec(v4u x) const { return x * a + b; }
};

struct AddF32 {
    __device__ float scalar(float x, float y) const { return add_rn(x, y); }
    __device__ v4f vec(v4f x, v4f y) const {
        v4f r;
        r.x = add_rn(x.x, y.x);
        r.y = add_rn(x.y, y.y);
        r.z = add_rn(x.z, y.z);
        r.w = add_rn(x.w, y.w);
        return r;
    }
};

struct ConstF32 {
    float v;
    __device__ float operator()(size_t) const { return v; }
};
struct ConstI32 {
    int32_t v;
    __device__ int32_t operator()(size_t) const { return v; }
};
struct UniformF32 {
    uint64_t seed;
    float scale;
    __device__ float operator()(size_t i) const {
        uint32_t u = (uint32_t)(hash64(seed, i) >> 40);
        return mul_rn((float)u, scale);
    }
};
struct BernoulliI32 {
    uint64_t seed, threshold;
    __device__ int32_t operator()(size_t i) const { return hash64(seed, i) < threshold ? 1 : 0; }
};
struct BitsU32 {
    uint64_t seed;
    __device__ uint32_t operator()(size_t i) const { return (uint32_t)(hash64(seed, i) >> 32); }
};

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

template <typename V, typename S, typename Op>
int launch_gen(S *out, size_t n, Op op, const char *name) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (n == 0) return MSL_OK;
    MSL_REQUIRE(out != nullptr, std::string(name) + ": null pointer");
    int blocks = streaming_blocks((n + 3) / 4, kThreads, tune_int("MSL_STREAM_BLOCKS_PER_CU", 1));
    if (aligned16(out))
        gen_kernel<V, S, Op><<<blocks, kThreads, 0, ctx().stream>>>(out, n, op);
    else
        gen_scalar_kernel<S, Op><<<blocks, kThreads, 0, ctx().stream>>>(out, n, op);
    return check_launch(name);
}

template <typename V, typename S, typename Op>
int launch_map(S *out, const S *in, size_t n, Op op, const char *name) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (n == 0) return MSL_OK;
    MSL_REQUIRE(out && in, std::string(name) + ": null pointer");
    // 4 vectors in flight per thread; one 256-thread block per CU by default.
    int blocks = streaming_blocks((n + 15) / 16, kThreads, tune_int("MSL_STREAM_BLOCKS_PER_CU", 1));
    if (aligned16(out) && aligned16(in))
        stream1_kernel<V, S, Op><<<blocks, kThreads, 0, ctx().stream>>>(out, in, n, op);
    else
        stream1_scalar_kernel<S, Op><<<blocks, kThreads, 0, ctx().stream>>>(out, in, n, op);
    return check_launch(name);
}

template <typename V, typename S, typename Op>
int launch_zip(S *out, const S *a, const S *b, size_t n, Op op, const char *name) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (n == 0) return MSL_OK;
    MSL_REQUIRE(out && a && b, std::string(name) + ": null pointer");
    int blocks = streaming_blocks((n + 15) / 16, kThreads, tune_int("MSL_STREAM_BLOCKS_PER_CU", 1));
    if (aligned16(out) && aligned16(a) && aligned16(b))
        stream2_kernel<V, S, Op><<<blocks, kThreads, 0, ctx().stream>>>(out, a, b, n, op);
    else
        stream2_scalar_kernel<S, Op><<<blocks, kThreads, 0, ctx().stream>>>(out, a, b, n, op);
    return check_launch(name);
}

// mapIndex with f(row, col, v) = row*cols + col over a row-major matrix:
// for every flat index k = row*cols + col (row = k / cols, col = k % cols)
// the int value row*cols + col equals k, so the element is (float)k - the
// same int->float conversion (round to nearest even) the functor performs.
struct LinIndexF32 {
    __device__ float operator()(size_t k) const { return (float)(int32_t)k; }
};

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_fill_f32(float *d_out, size_t n, float value) {
    return launch_gen<v4f>(d_out, n, ConstF32{value}, "msl_fill_f32");
}
int msl_fill_i32(int32_t *d_out, size_t n, int32_t value) {
    return launch_gen<v4i>(d_out, n, ConstI32{value}, "msl_fill_i32");
}
int msl_fill_uniform_f32(float *d_out, size_t n, uint64_t seed, float scale) {
    return launch_gen<v4f>(d_out, n, UniformF32{seed, scale}, "msl_fill_uniform_f32");
}
int msl_fill_bernoulli_i32(int32_t *d_out, size_t n, uint64_t seed, uint64_t threshold) {
    return launch_gen<v4i>(d_out, n, BernoulliI32{seed, threshold}, "msl_fill_bernoulli_i32");
}
int msl_fill_u32(uint32_t *d_out, size_t n, uint64_t seed) {
    return launch_gen<v4u>(d_out, n, BitsU32{seed}, "msl_fill_u32");
}

int msl_map_affine_f32(float *d_out, const float *d_in, size_t n, float a, float b) {
    return launch_map<v4f>(d_out, d_in, n, AffineF32{a, b}, "msl_map_affine_f32");
}
int msl_map_affine_u32(uint32_t *d_out, const uint32_t *d_in, size_t n, uint32_t a, uint32_t b) {
    return launch_map<v4u>(d_out, d_in, n, AffineU32{a, b}, "msl_map_affine_u32");
}
int msl_zip_add_f32(float *d_out, const float *d_a, const float *d_b, size_t n) {
    return launch_zip<v4f>(d_out, d_a, d_b, n, AddF32{}, "msl_zip_add_f32");
}

int msl_mapindex_lin_f32(float *d_out, int rows, int cols) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(rows >= 0 && cols >= 0, "msl_mapindex_lin_f32: negative shape");
    MSL_REQUIRE((int64_t)rows * cols <= INT32_MAX, "msl_mapindex_lin_f32: rows*cols overflows int");
    if (rows == 0 || cols == 0) return MSL_OK;
    MSL_REQUIRE(d_out != nullptr, "msl_mapindex_lin_f32: null pointer");
    return launch_gen<v4f>(d_out, (size_t)rows * cols, LinIndexF32{}, "msl_mapindex_lin_f32");
}

}  // extern "C"

// Folds: compensated FP32 sum, wrapping uint32 sum, exact int32 sum, max.
//
// Reference: DS::fold (src/ds.cpp:739-912) copies the whole array to the host,
// cudaMallocs a partial buffer and runs a shared-memory tree whose second pass
// is skipped beyond 1024 blocks (src/fold_kernels.cu:298-403).  Here a fold is
// two launches and no host copy: stage 1 streams the array once with 16-byte
// loads (each lane accumulates, then a wave64 shuffle tree and a 4-wave LDS
// combine produce one partial per block); stage 2 is one block that combines
// the partials in a fixed order.  Results stay in device memory.
//
// Compensated sum: every lane keeps (s, c) and adds each element with Knuth's
// TwoSum (s' = s + x exactly split into s' and the rounding error, which is
// accumulated in c); partial pairs are merged with TwoSum as well.  Stage 2
// converts each (s, c) to FP64 and adds them in block order - the only FP64
// arithmetic in the fold.
#include "common.hpp"

namespace mslhip {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef int32_t v4i __attribute__((ext_vector_type(4)));

constexpr int kFoldThreads = 256;
constexpr int kFoldWaves = kFoldThreads / 64;

struct Pair {
    float s, c;
};

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = __fadd_rn(a, b);
    float bp = __fsub_rn(s, a);
    float ap = __fsub_rn(s, bp);
    e = __fadd_rn(__fsub_rn(a, ap), __fsub_rn(b, bp));
}

// ---- policies --------------------------------------------------------------
struct CompSumPolicy {
    using In = float;
    using V = v4f;
    using Part = Pair;
    __device__ static Part init() { return Pair{0.0f, 0.0f}; }
    __device__ static void push(Part &p, float x) {
        float s, e;
        two_sum(p.s, x, s, e);
        p.s = s;
        p.c = __fadd_rn(p.c, e);
    }
    __device__ static Part combine(Part a, Part b) {
        float s, e;
        two_sum(a.s, b.s, s, e);
        return Pair{s, __fadd_rn(__fadd_rn(a.c, b.c), e)};
    }
    __device__ static Part shfl_xor(Part p, int m) { return Pair{__shfl_xor(p.s, m), __shfl_xor(p.c, m)}; }
};

struct SumU32Policy {
    using In = uint32_t;
    using V = v4u;
    using Part = uint32_t;
    __device__ static Part init() { return 0u; }
    __device__ static void push(Part &p, uint32_t x) { p += x; }
    __device__ static Part combine(Part a, Part b) { return a + b; }
    __device__ static Part shfl_xor(Part p, int m) { return (uint32_t)__shfl_xor((int)p, m); }
};

struct SumI32Policy {
    using In = int32_t;
    using V = v4i;
    using Part = long long;
    __device__ static Part init() { return 0; }
    __device__ static void push(Part &p, int32_t x) { p += x; }
    __device__ static Part combine(Part a, Part b) { return a + b; }
    __device__ static Part shfl_xor(Part p, int m) { return __shfl_xor(p, m); }
};

struct MaxF32Policy {
    using In = float;
    using V = v4f;
    using Part = float;
    __device__ static Part init() { return -__builtin_huge_valf(); }
    __device__ static void push(Part &p, float x) { p = x > p ? x : p; }
    __device__ static Part combine(Part a, Part b) { return b > a ? b : a; }
    __device__ static Part shfl_xor(Part p, int m) { return __shfl_xor(p, m); }
};

template <typename P>
__device__ __forceinline__ typename P::Part wave_reduce(typename P::Part v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = P::combine(v, P::shfl_xor(v, m));
    return v;
}

template <typename P>
__device__ __forceinline__ typename P::Part block_reduce(typename P::Part v) {
    __shared__ typename P::Part lds[kFoldWaves];
    v = wave_reduce<P>(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    typename P::Part r = lds[0];
#pragma unroll
    for (int w = 1; w < kFoldWaves; w++) r = P::combine(r, lds[w]);
    return r;  // identical in every thread
}

// Stage 1: one partial per block.
template <typename P>
__global__ __launch_bounds__(kFoldThreads) void fold_stage1(const typename P::In *__restrict__ in, size_t n,
                                                            typename P::Part *__restrict__ partials, int vec_ok) {
    using V = typename P::V;
    typename P::Part acc = P::init();
    const size_t stride = (size_t)gridDim.x * kFoldThreads;
    size_t i = (size_t)blockIdx.x * kFoldThreads + threadIdx.x;
    size_t done = 0;
    if (vec_ok) {
        const size_t n4 = n / 4;
        const V *in4 = reinterpret_cast<const V *>(in);
        // Software pipeline: the next U vectors are requested before the
        // current U are accumulated, so each lane keeps up to 2U 16-byte
        // loads in flight.
        constexpr int U = 4;
        if (i + (U - 1) * stride < n4) {
            V v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = in4[i + u * stride];
            i += U * stride;
            for (; i + (U - 1) * stride < n4; i += U * stride) {
                V nv[U];
#pragma unroll
                for (int u = 0; u < U; u++) nv[u] = in4[i + u * stride];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    P::push(acc, v[u].x);
                    P::push(acc, v[u].y);
                    P::push(acc, v[u].z);
                    P::push(acc, v[u].w);
                    v[u] = nv[u];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                P::push(acc, v[u].x);
                P::push(acc, v[u].y);
                P::push(acc, v[u].z);
                P::push(acc, v[u].w);
            }
        }
        for (; i < n4; i += stride) {
            V v = in4[i];
            P::push(acc, v.x); P::push(acc, v.y); P::push(acc, v.z); P::push(acc, v.w);
        }
        done = n4 * 4;
        i = done + (size_t)blockIdx.x * kFoldThreads + threadIdx.x;
    }
    for (; i < n; i += stride) P::push(acc, in[i]);
    typename P::Part r = block_reduce<P>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// Stage 2 for the compensated sum: FP64 combine of (s, c) partials in a
// fixed order (thread t takes partials t, t+256, ...; then a fixed tree).
__global__ __launch_bounds__(kFoldThreads) void fold_sum_f32_stage2(const Pair *__restrict__ partials, int nparts,
                                                                     double *__restrict__ result) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < nparts; k += kFoldThreads) {
        Pair p = partials[k];
        acc += (double)p.s;
        acc += (double)p.c;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
    __shared__ double lds[kFoldWaves];
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = lds[0];
        for (int w = 1; w < kFoldWaves; w++) r += lds[w];
        result[0] = r;
    }
}

template <typename P>
__global__ __launch_bounds__(kFoldThreads) void fold_generic_stage2(const typename P::Part *__restrict__ partials,
                                                                     int nparts, typename P::Part *__restrict__ result) {
    typename P::Part acc = P::init();
    for (int k = threadIdx.x; k < nparts; k += kFoldThreads) acc = P::combine(acc, partials[k]);
    typename P::Part r = block_reduce<P>(acc);
    if (threadIdx.x == 0) result[0] = r;
}

inline int fold_blocks(size_t n) {
    size_t want = (n + (size_t)kFoldThreads * 16 - 1) / ((size_t)kFoldThreads * 16);
    size_t cap = (size_t)ctx().num_cus * tune_int("MSL_FOLD_BLOCKS_PER_CU", 2);
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

// Workspace layout: [0, 256) result scalars, then partials.
constexpr size_t kResultBytes = 256;

template <typename P, typename Stage2>
int run_fold(const typename P::In *in, size_t n, void *d_result, Stage2 stage2, const char *name) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(in != nullptr || n == 0, std::string(name) + ": null input");
    int blocks = fold_blocks(n);
    void *ws = nullptr;
    rc = workspace(kResultBytes + (size_t)blocks * sizeof(typename P::Part), &ws);
    if (rc) return rc;
    auto *partials = reinterpret_cast<typename P::Part *>((char *)ws + kResultBytes);
    if (!d_result) d_result = ws;
    int vec_ok = (((uintptr_t)in) & 15u) == 0;
    fold_stage1<P><<<blocks, kFoldThreads, 0, ctx().stream>>>(in, n, partials, vec_ok);
    rc = check_launch(name);
    if (rc) return rc;
    stage2(partials, blocks, d_result);
    return check_launch(name);
}

int scratch_result(void **p) { return workspace(kResultBytes, p); }

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_fold_sum_f32(const float *d_in, size_t n, double *d_result) {
    MSL_REQUIRE(d_result != nullptr, "msl_fold_sum_f32: null result pointer");
    return run_fold<CompSumPolicy>(d_in, n, d_result,
                                   [](Pair *parts, int np, void *res) {
                                       fold_sum_f32_stage2<<<1, kFoldThreads, 0, ctx().stream>>>(parts, np,
                                                                                                 (double *)res);
                                   },
                                   "msl_fold_sum_f32");
}

int msl_fold_sum_f32_host(const float *d_in, size_t n, double *h_sum64, float *h_sum32) {
    int rc = run_fold<CompSumPolicy>(d_in, n, nullptr,
                                     [](Pair *parts, int np, void *res) {
                                         fold_sum_f32_stage2<<<1, kFoldThreads, 0, ctx().stream>>>(parts, np,
                                                                                                   (double *)res);
                                     },
                                     "msl_fold_sum_f32_host");
    if (rc) return rc;
    void *ws = nullptr;
    rc = scratch_result(&ws);
    if (rc) return rc;
    double r = 0.0;
    rc = msl_memcpy_d2h(&r, ws, sizeof(double));
    if (rc) return rc;
    if (h_sum64) *h_sum64 = r;
    if (h_sum32) *h_sum32 = (float)r;
    return MSL_OK;
}

int msl_fold_sum_u32_host(const uint32_t *d_in, size_t n, uint32_t *h_sum) {
    MSL_REQUIRE(h_sum != nullptr, "msl_fold_sum_u32_host: null result pointer");
    int rc = run_fold<SumU32Policy>(d_in, n, nullptr,
                                    [](uint32_t *parts, int np, void *res) {
                                        fold_generic_stage2<SumU32Policy><<<1, kFoldThreads, 0, ctx().stream>>>(
                                            parts, np, (uint32_t *)res);
                                    },
                                    "msl_fold_sum_u32_host");
    if (rc) return rc;
    void *ws = nullptr;
    rc = scratch_result(&ws);
    if (rc) return rc;
    return msl_memcpy_d2h(h_sum, ws, sizeof(uint32_t));
}

int msl_fold_sum_i32_host(const int32_t *d_in, size_t n, int64_t *h_sum) {
    MSL_REQUIRE(h_sum != nullptr, "msl_fold_sum_i32_host: null result pointer");
    int rc = run_fold<SumI32Policy>(d_in, n, nullptr,
                                    [](long long *parts, int np, void *res) {
                                        fold_generic_stage2<SumI32Policy><<<1, kFoldThreads, 0, ctx().stream>>>(
                                            parts, np, (long long *)res);
                                    },
                                    "msl_fold_sum_i32_host");
    if (rc) return rc;
    void *ws = nullptr;
    rc = scratch_result(&ws);
    if (rc) return rc;
    long long r = 0;
    rc = msl_memcpy_d2h(&r, ws, sizeof(long long));
    if (rc) return rc;
    *h_sum = (int64_t)r;
    return MSL_OK;
}

int msl_fold_max_f32_host(const float *d_in, size_t n, float *h_max) {
    MSL_REQUIRE(h_max != nullptr, "msl_fold_max_f32_host: null result pointer");
    MSL_REQUIRE(n > 0, "msl_fold_max_f32_host: empty input");
    int rc = run_fold<MaxF32Policy>(d_in, n, nullptr,
                                    [](float *parts, int np, void *res) {
                                        fold_generic_stage2<MaxF32Policy><<<1, kFoldThreads, 0, ctx().stream>>>(
                                            parts, np, (float *)res);
                                    },
                                    "msl_fold_max_f32_host");
    if (rc) return rc;
    void *ws = nullptr;
    rc = scratch_result(&ws);
    if (rc) return rc;
    return msl_memcpy_d2h(h_max, ws, sizeof(float));
}

}  // extern "C"

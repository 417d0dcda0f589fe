// Folds: compensated FP32 sum, wrapping uint32 sum, exact int32 sum, max.
//
// Reference: DS::fold (src/ds.cpp:739-912) copies the whole array to the host,
// cudaMallocs a partial buffer and runs a shared-memory tree whose second pass
// is skipped beyond 1024 blocks (src/fold_kernels.cu:298-403).  Here a fold is
// two launches and no host copy: the array is streamed once with 16-byte
// loads (each lane accumulates, then a wave64 shuffle tree and a 4-wave LDS
// combine produce one partial per block) and one block combines the partials
// in a fixed order.  Results stay in device memory.
//
// Compensated sum: every lane keeps (s, c) and adds each element with Knuth's
// TwoSum (s' = s + x exactly split into s' and the rounding error, which is
// accumulated in c); partial pairs are merged with TwoSum as well.  The final
// combine converts each (s, c) to FP64 and adds them in block order - the
// only FP64 arithmetic in the fold.
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

// Stage 1: every block streams its grid-stride share with four 16-byte loads
// in flight per lane and reduces it to one partial.  The order of every
// combination is fixed, so results are deterministic for a given grid size.
template <typename P, bool kF64Final>
__global__ __launch_bounds__(kFoldThreads) void fold_kernel(const typename P::In *__restrict__ in, size_t n,
                                                            typename P::Part *__restrict__ partials, int vec_ok) {
    using V = typename P::V;
    typename P::Part acc = P::init();
    const size_t stride = (size_t)gridDim.x * kFoldThreads;
    size_t i = (size_t)blockIdx.x * kFoldThreads + threadIdx.x;
    if (vec_ok) {
        const size_t n4 = n / 4;
        const V *in4 = reinterpret_cast<const V *>(in);
        // 16-byte loads in flight per lane (8192^2: 3 5.78 TB/s, 4 5.67, 6
        // 5.33, 8 5.36); each lane's elements are folded in the same order
        // whatever U is
        constexpr int U = 3;
        for (; i + (U - 1) * stride < n4; i += U * stride) {
            V v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = in4[i + u * stride];
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
        i = n4 * 4 + (size_t)blockIdx.x * kFoldThreads + threadIdx.x;
    }
    for (; i < n; i += stride) P::push(acc, in[i]);
    const typename P::Part r = block_reduce<P>(acc);

    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// Final combine (second launch, one block): partials in block order.  A
// second launch is used instead of a last-block-done counter because the
// agent-scope release fence that counter needs writes back the whole L2 of
// each XCD and costs more than the extra launch.
template <typename P, bool kF64Final>
__global__ __launch_bounds__(kFoldThreads) void fold_final_kernel(const typename P::Part *__restrict__ partials,
                                                                  int nparts, void *__restrict__ result) {
    if constexpr (kF64Final) {
        // Compensated sum: FP64 combine of the (s, c) partials, the only FP64
        // arithmetic of the fold (thread t takes partials t, t+256, ...).
        double d = 0.0;
        for (int k = threadIdx.x; k < nparts; k += kFoldThreads) {
            const Pair p = partials[k];
            d += (double)p.s;
            d += (double)p.c;
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) d += __shfl_xor(d, m);
        __shared__ double lds[kFoldWaves];
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = d;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = lds[0];
            for (int w = 1; w < kFoldWaves; w++) t += lds[w];
            *reinterpret_cast<double *>(result) = t;
        }
    } else {
        typename P::Part a = P::init();
        for (int k = threadIdx.x; k < nparts; k += kFoldThreads) a = P::combine(a, partials[k]);
        const typename P::Part t = block_reduce<P>(a);
        if (threadIdx.x == 0) *reinterpret_cast<typename P::Part *>(result) = t;
    }
}

// Two blocks per CU: the read-only stream needs more loads in flight than
// one block's four 16-byte loads per lane (8192^2 floats with the input
// evicted from the Infinity Cache: 4.7 TB/s with one block per CU, 5.7 with
// two, 5.1 with four).
inline int fold_blocks(size_t n) {
    size_t want = (n + (size_t)kFoldThreads * 16 - 1) / ((size_t)kFoldThreads * 16);
    size_t cap = (size_t)ctx().num_cus * tune_int("MSL_FOLD_BLOCKS_PER_CU", 2);
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

// Workspace layout: [0, 256) result scalars, partials from byte 512.
constexpr size_t kResultBytes = 256;
constexpr size_t kPartialsOffset = 512;

template <typename P, bool kF64Final>
int run_fold(const typename P::In *in, size_t n, void *d_result, const char *name) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(in != nullptr || n == 0, std::string(name) + ": null input");
    const int blocks = fold_blocks(n);
    void *ws = nullptr;
    rc = workspace(kPartialsOffset + (size_t)blocks * sizeof(typename P::Part), &ws);
    if (rc) return rc;
    auto *partials = reinterpret_cast<typename P::Part *>((char *)ws + kPartialsOffset);
    if (!d_result) d_result = ws;
    const int vec_ok = (((uintptr_t)in) & 15u) == 0;
    fold_kernel<P, kF64Final><<<blocks, kFoldThreads, 0, ctx().stream>>>(in, n, partials, vec_ok);
    rc = check_launch(name);
    if (rc) return rc;
    fold_final_kernel<P, kF64Final><<<1, kFoldThreads, 0, ctx().stream>>>(partials, blocks, d_result);
    return check_launch(name);
}

int scratch_result(void **p) { return workspace(kResultBytes, p); }

// ---- allocation-free compensated fold ---------------------------------------
// msl_fold_sum_f32_noalloc keeps its block partials in this file-scope device
// array instead of the context workspace, so it never allocates (callers under
// HIP graph capture or in loops that must not call hipMalloc).  Each GPU holds
// its own copy for the life of the process; in host code the array's name is
// not a device address, so the address is looked up once per device with
// hipGetSymbolAddress (noalloc_setup, called by msl_init with the device
// current) and only that pointer is passed to kernels.  Stage 1 writes every
// partial it later reads, so nothing needs clearing between calls.  Calls that
// overlap on one GPU (several host threads or streams) share the array and
// race.
constexpr int kNoAllocMaxPartials = 4096;
constexpr size_t kScratchBytes = (size_t)kNoAllocMaxPartials * sizeof(Pair);  // 32 KiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(MSL_E_INVALID, "msl_init: device index beyond 63");
    if (g_scratch_ptr[device]) return MSL_OK;
    void *p = nullptr;
    MSL_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
    return MSL_OK;
}

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_fold_sum_f32(const float *d_in, size_t n, double *d_result) {
    MSL_REQUIRE(d_result != nullptr, "msl_fold_sum_f32: null result pointer");
    return run_fold<CompSumPolicy, true>(d_in, n, d_result, "msl_fold_sum_f32");
}

int msl_fold_sum_f32_noalloc(const float *d_in, size_t n, double *d_result) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(d_result != nullptr, "msl_fold_sum_f32_noalloc: null result pointer");
    MSL_REQUIRE(d_in != nullptr || n == 0, "msl_fold_sum_f32_noalloc: null input");
    const int blocks = fold_blocks(n);
    if (blocks > kNoAllocMaxPartials)
        return fail(MSL_E_INVALID, "msl_fold_sum_f32_noalloc: " + std::to_string(blocks) +
                                       " block partials exceed the 4096 the allocation-free scratch holds "
                                       "(lower MSL_FOLD_BLOCKS_PER_CU)");
    Pair *partials = reinterpret_cast<Pair *>(g_scratch_ptr[ctx().device]);
    MSL_REQUIRE(partials != nullptr, "msl_fold_sum_f32_noalloc: scratch not set up on this device");
    const int vec_ok = (((uintptr_t)d_in) & 15u) == 0;
    fold_kernel<CompSumPolicy, true><<<blocks, kFoldThreads, 0, ctx().stream>>>(d_in, n, partials, vec_ok);
    rc = check_launch("msl_fold_sum_f32_noalloc");
    if (rc) return rc;
    fold_final_kernel<CompSumPolicy, true><<<1, kFoldThreads, 0, ctx().stream>>>(partials, blocks, d_result);
    return check_launch("msl_fold_sum_f32_noalloc");
}

int msl_fold_sum_f32_host(const float *d_in, size_t n, double *h_sum64, float *h_sum32) {
    int rc = run_fold<CompSumPolicy, true>(d_in, n, nullptr, "msl_fold_sum_f32_host");
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
    int rc = run_fold<SumU32Policy, false>(d_in, n, nullptr, "msl_fold_sum_u32_host");
    if (rc) return rc;
    void *ws = nullptr;
    rc = scratch_result(&ws);
    if (rc) return rc;
    return msl_memcpy_d2h(h_sum, ws, sizeof(uint32_t));
}

int msl_fold_sum_i32_host(const int32_t *d_in, size_t n, int64_t *h_sum) {
    MSL_REQUIRE(h_sum != nullptr, "msl_fold_sum_i32_host: null result pointer");
    int rc = run_fold<SumI32Policy, false>(d_in, n, nullptr, "msl_fold_sum_i32_host");
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
    int rc = run_fold<MaxF32Policy, false>(d_in, n, nullptr, "msl_fold_max_f32_host");
    if (rc) return rc;
    void *ws = nullptr;
    rc = scratch_result(&ws);
    if (rc) return rc;
    return msl_memcpy_d2h(h_max, ws, sizeof(float));
}

}  // extern "C"

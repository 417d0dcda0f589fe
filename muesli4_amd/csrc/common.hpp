// Internal helpers shared by the translation units of libmsl_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "msl_hip.h"

namespace mslhip {

// ---- error state ---------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define MSL_HIP_TRY(expr)                                   \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return ::mslhip::hip_fail(_e, #expr); \
    } while (0)

#define MSL_REQUIRE(cond, msg)                                          \
    do {                                                                \
        if (!(cond)) return ::mslhip::fail(MSL_E_INVALID, (msg));       \
    } while (0)

// ---- context --------------------------------------------------------------
struct Context {
    bool ready = false;
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool capturing = false;
};

Context &ctx();
// Returns MSL_OK when the context is initialised (initialising device 0 on
// first use), else an error code.
int ensure_ready();
// Launch-error check after a kernel launch.
int check_launch(const char *what);

// Scratch workspace owned by the context (grown outside graph capture only).
int workspace(size_t bytes, void **ptr);
// Looks up, with `device` current, the device address of the fixed scratch
// array of the allocation-free fold (fold.hip); called once per device by
// msl_init.
int noalloc_setup(int device);

// Integer tuning knob from the environment, else dflt.  The result of the
// first call is cached per name, dflt included: callers whose default
// varies pass a fixed sentinel (0) and apply their default themselves.
int tune_int(const char *name, int dflt);

// Number of workgroups that fills the device for a streaming kernel.
inline int streaming_blocks(size_t work_items, int threads, int per_cu) {
    size_t want = (work_items + threads - 1) / threads;
    size_t cap = (size_t)ctx().num_cus * per_cu;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

// Store policy of the stencil marches.  A grid of up to 128 MiB ping-pongs
// with its partner (256 MiB together) inside the 256 MiB Infinity Cache
// well enough that ordinary stores win: the next step finds its input there.
// Larger grids use non-temporal stores, which do not allocate in the
// Infinity Cache and so leave it to the input being read (measured on
// 8192^2 and 512^3 ping-pongs: ordinary stores lose 2-7 %).
constexpr size_t kCachedStoreMaxBytes = (size_t)128 << 20;
inline bool stream_stores(size_t grid_bytes) { return grid_bytes > kCachedStoreMaxBytes; }

// ---- device-side helpers --------------------------------------------------
__host__ __device__ inline uint64_t hash64(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// IEEE single operations that the compiler may not contract into an FMA.
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }

// Knuth's TwoSum: s = RN(a + b) and the exact rounding error e, so that
// a + b == s + e in real arithmetic.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = __fadd_rn(a, b);
    float bp = __fsub_rn(s, a);
    float ap = __fsub_rn(s, bp);
    e = __fadd_rn(__fsub_rn(a, ap), __fsub_rn(b, bp));
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// waits for every outstanding global load of the wave, which would drain the
// software prefetch queues of the stencil marches at each step; the
// "local"-only fences make the barrier wait for LDS operations alone.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders one wave's own LDS accesses (RAW/WAR between its lanes).  A wave's
// LDS instructions execute in issue order, so this is a compiler-level
// barrier only: no wait on outstanding global loads, no workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Correctly rounded x / d for the divisors d = (2r+1)^2, r = 1..16, without
// the IEEE division sequence: q0 = x*rcp (rcp = RN(1/d)), the remainder
// x - q0*d is exact with one FMA, and one FMA correction gives RN(x/d).
// When the remainder is 0 (q0 exact; covers -0) or NaN (x infinite or NaN)
// q0 itself is the answer.  Checked bit-for-bit against IEEE division for
// all 2^32 inputs of every such d by msl_selftest_div_odd_square.
__device__ __forceinline__ float div_odd_square(float x, float d, float rcp) {
    float q0 = __fmul_rn(x, rcp);
    float r = __fmaf_rn(-q0, d, x);
    float q1 = __fmaf_rn(r, rcp, q0);
    return (r != 0.0f && r == r) ? q1 : q0;
}

// Correctly rounded (0 + x) / d for the same divisors: the finish of a sum
// that the caller formed without its leading "0 +" (which only turns a -0
// sum into +0).  The sequence of div_odd_square, with the remainder clamped
// to a finite value instead of selecting q0: for infinite x the correction
// then adds a finite term to q0 = +-inf, for NaN x it stays NaN, for x = -0
// the remainder +0 gives +0 + -0 = +0, and for an exact quotient q0 + 0 =
// q0.  Checked bit-for-bit against IEEE division of 0 + x on all 2^32
// inputs by msl_selftest_div_odd_square.
__device__ __forceinline__ float div_plus0(float x, float d, float rcp) {
    const float q0 = __fmul_rn(x, rcp);
    const float r = __fmaf_rn(-q0, d, x);
    const float rc = __builtin_amdgcn_fmed3f(r, -3.0e38f, 3.0e38f);  // NaN -> -3e38
    return __fmaf_rn(rc, rcp, q0);
}

// Correctly rounded x / 6 (the heat functor's division) by the sequence of
// div_odd_square with rcp = RN(1/6), plus one case that sequence misses: a
// quotient exactly halfway between two subnormals (x = (6k+3) * 2^-149, where
// the remainder r = x - 6*q0 is +-3 * 2^-149 and the inexact reciprocal
// always rounds away from q0), which must round to the even one of q0 and
// q1 = q0 +- 2^-149.  Branch-free; checked bit-for-bit against IEEE
// division on all 2^32 inputs by msl_selftest_div_six.
__device__ __forceinline__ float div6(float x) {
    constexpr float kRcp6 = 1.0f / 6.0f;
    const float q0 = __fmul_rn(x, kRcp6);
    const float r = __fmaf_rn(-q0, 6.0f, x);
    const float q1 = __fmaf_rn(r, kRcp6, q0);
    const bool keep_even_q0 = fabsf(r) == 0x1.8p-148f && (__float_as_uint(q0) & 1u) == 0u;
    return (r != 0.0f && r == r && !keep_even_q0) ? q1 : q0;
}

// Correctly rounded (0 + x) / 6 for a sum formed without its leading "0 +":
// div6 with the remainder clamped to a finite value as in div_plus0.
// Checked on all 2^32 inputs by msl_selftest_div_six.
__device__ __forceinline__ float div6_plus0(float x) {
    constexpr float kRcp6 = 1.0f / 6.0f;
    const float q0 = __fmul_rn(x, kRcp6);
    const float r = __fmaf_rn(-q0, 6.0f, x);
    const bool keep_even_q0 = fabsf(r) == 0x1.8p-148f && (__float_as_uint(q0) & 1u) == 0u;
    const float q = __fmaf_rn(__builtin_amdgcn_fmed3f(r, -3.0e38f, 3.0e38f), kRcp6, q0);
    return keep_even_q0 ? q0 : q;
}

// Remap blockIdx so that each XCD (blocks are dealt round-robin to the 8
// XCDs) receives a contiguous range of logical tiles; neighbouring tiles
// then share one L2.  Identity when the grid does not split evenly.
__device__ __forceinline__ unsigned xcd_tile(unsigned bid, unsigned nblocks) {
    constexpr unsigned kXcd = 8;
    if (nblocks % kXcd != 0) return bid;
    return (bid % kXcd) * (nblocks / kXcd) + bid / kXcd;
}

}  // namespace mslhip

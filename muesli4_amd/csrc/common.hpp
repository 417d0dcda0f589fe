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

// Number of workgroups that fills the device for a streaming kernel.
inline int streaming_blocks(size_t work_items, int threads, int per_cu) {
    size_t want = (work_items + threads - 1) / threads;
    size_t cap = (size_t)ctx().num_cus * per_cu;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

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

// Remap blockIdx so that each XCD (blocks are dealt round-robin to the 8
// XCDs) receives a contiguous range of logical tiles; neighbouring tiles
// then share one L2.  Identity when the grid does not split evenly.
__device__ __forceinline__ unsigned xcd_tile(unsigned bid, unsigned nblocks) {
    constexpr unsigned kXcd = 8;
    if (nblocks % kXcd != 0) return bid;
    return (bid % kXcd) * (nblocks / kXcd) + bid / kXcd;
}

}  // namespace mslhip

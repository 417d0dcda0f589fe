// Runtime: context, stream, error state, residency, timing and graph capture.
//
// Replaces the device side of msl::initSkeletons / terminateSkeletons /
// syncStreams (reference src/muesli.cpp:63-158, 213-223): one process, one
// device, one stream; skeletons are stream-ordered instead of synchronising
// after every call.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace mslhip {

static thread_local std::string g_last_error;
static Context g_ctx;
static std::mutex g_init_mutex;
static void *g_ws = nullptr;
// Pinned, device-accessible host buffer for small device-to-host copies: a
// kernel writes the bytes there, which avoids the copy engine's latency
// (a scalar fold result otherwise costs about as much as the fold).
constexpr size_t kSmallCopyBytes = 64 << 10;
static char *g_small = nullptr;
static size_t g_ws_bytes = 0;
static std::vector<void *> g_ws_retired;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    set_error(m);
    return e == hipErrorOutOfMemory ? MSL_E_NOMEM : MSL_E_HIP;
}

Context &ctx() { return g_ctx; }

int ensure_ready() {
    if (g_ctx.ready) return MSL_OK;
    return msl_init(0);
}

int tune_int(const char *name, int dflt) {
    static std::mutex mu;
    static std::map<std::string, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(name);
    if (it != cache.end()) return it->second;
    int v = dflt;
    if (const char *e = std::getenv(name)) {
        char *end = nullptr;
        long x = std::strtol(e, &end, 10);
        if (end != e && x > 0 && x < (1 << 20)) v = (int)x;
    }
    cache[name] = v;
    return v;
}

__global__ void small_copy_kernel(char *__restrict__ dst, const char *__restrict__ src, size_t bytes) {
    for (size_t i = threadIdx.x; i < bytes; i += blockDim.x) dst[i] = src[i];
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return MSL_OK;
}

int workspace(size_t bytes, void **ptr) {
    if (bytes <= g_ws_bytes && g_ws) {
        *ptr = g_ws;
        return MSL_OK;
    }
    if (g_ctx.capturing) return fail(MSL_E_STATE, "workspace must grow outside graph capture; call once before capturing");
    // A smaller workspace is retired, not freed: graphs captured earlier may
    // still reference it.  Retired blocks are released by msl_finalize.
    if (g_ws) {
        g_ws_retired.push_back(g_ws);
        g_ws = nullptr;
        g_ws_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    MSL_HIP_TRY(hipMalloc(&g_ws, want));
    g_ws_bytes = want;
    *ptr = g_ws;
    return MSL_OK;
}

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_init(int device) {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (g_ctx.ready) {
        if (g_ctx.device != device) return fail(MSL_E_STATE, "msl_init: already initialised on another device");
        return MSL_OK;
    }
    int count = 0;
    MSL_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(MSL_E_INVALID, "msl_init: no such device");
    MSL_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    MSL_HIP_TRY(hipGetDeviceProperties(&prop, device));
    // the allocation-free fold's scratch address on this device
    int rc = noalloc_setup(device);
    if (rc) return rc;
    g_ctx.device = device;
    g_ctx.num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    // Default-priority, non-blocking stream (one per device as in the reference).
    MSL_HIP_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    MSL_HIP_TRY(hipEventCreate(&g_ctx.t0));
    MSL_HIP_TRY(hipEventCreate(&g_ctx.t1));
    MSL_HIP_TRY(hipHostMalloc((void **)&g_small, kSmallCopyBytes, hipHostMallocDefault));
    g_ctx.ready = true;
    return MSL_OK;
}

int msl_finalize(void) {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (!g_ctx.ready) return MSL_OK;
    MSL_HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    if (g_ws) {
        MSL_HIP_TRY(hipFree(g_ws));
        g_ws = nullptr;
        g_ws_bytes = 0;
    }
    for (void *p : g_ws_retired) hipFree(p);
    g_ws_retired.clear();
    if (g_small) {
        (void)hipHostFree(g_small);
        g_small = nullptr;
    }
    hipEventDestroy(g_ctx.t0);
    hipEventDestroy(g_ctx.t1);
    hipStreamDestroy(g_ctx.stream);
    g_ctx = Context();
    return MSL_OK;
}

const char *msl_last_error(void) { return g_last_error.c_str(); }

void *msl_stream(void) {
    if (ensure_ready() != MSL_OK) return nullptr;
    return (void *)g_ctx.stream;
}

int msl_synchronize(void) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return MSL_OK;
}

int msl_device_info(int *num_cus, int *warp_size, size_t *total_mem, char *name, int name_len) {
    int rc = ensure_ready();
    if (rc) return rc;
    hipDeviceProp_t prop;
    MSL_HIP_TRY(hipGetDeviceProperties(&prop, g_ctx.device));
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (warp_size) *warp_size = prop.warpSize;
    if (total_mem) *total_mem = prop.totalGlobalMem;
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", prop.gcnArchName);
    }
    return MSL_OK;
}

int msl_malloc(void **d_ptr, size_t bytes) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(d_ptr != nullptr, "msl_malloc: null out-pointer");
    if (g_ctx.capturing) return fail(MSL_E_STATE, "msl_malloc: not allowed during graph capture");
    *d_ptr = nullptr;
    if (bytes == 0) return MSL_OK;
    MSL_HIP_TRY(hipMalloc(d_ptr, bytes));
    return MSL_OK;
}

int msl_free(void *d_ptr) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (!d_ptr) return MSL_OK;
    if (g_ctx.capturing) return fail(MSL_E_STATE, "msl_free: not allowed during graph capture");
    MSL_HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    MSL_HIP_TRY(hipFree(d_ptr));
    return MSL_OK;
}

int msl_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (bytes == 0) return MSL_OK;
    MSL_REQUIRE(d_dst && h_src, "msl_memcpy_h2d: null pointer");
    MSL_HIP_TRY(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, g_ctx.stream));
    MSL_HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return MSL_OK;
}

int msl_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (bytes == 0) return MSL_OK;
    MSL_REQUIRE(h_dst && d_src, "msl_memcpy_d2h: null pointer");
    if (bytes <= kSmallCopyBytes && g_small && !g_ctx.capturing) {
        small_copy_kernel<<<1, 256, 0, g_ctx.stream>>>(g_small, (const char *)d_src, bytes);
        rc = check_launch("msl_memcpy_d2h");
        if (rc) return rc;
        MSL_HIP_TRY(hipStreamSynchronize(g_ctx.stream));
        std::memcpy(h_dst, g_small, bytes);
        return MSL_OK;
    }
    MSL_HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
    MSL_HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return MSL_OK;
}

int msl_result_slot(void **h_slot, size_t *h_bytes) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(h_slot != nullptr, "msl_result_slot: null out-pointer");
    MSL_REQUIRE(g_small != nullptr, "msl_result_slot: no pinned slot");
    *h_slot = g_small;
    if (h_bytes) *h_bytes = kSmallCopyBytes;
    return MSL_OK;
}

int msl_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (bytes == 0) return MSL_OK;
    MSL_REQUIRE(d_dst && d_src, "msl_memcpy_d2d: null pointer");
    MSL_HIP_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, g_ctx.stream));
    return MSL_OK;
}

int msl_timer_start(void) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_HIP_TRY(hipEventRecord(g_ctx.t0, g_ctx.stream));
    return MSL_OK;
}

int msl_timer_stop(float *h_ms) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(h_ms != nullptr, "msl_timer_stop: null out-pointer");
    MSL_HIP_TRY(hipEventRecord(g_ctx.t1, g_ctx.stream));
    MSL_HIP_TRY(hipEventSynchronize(g_ctx.t1));
    MSL_HIP_TRY(hipEventElapsedTime(h_ms, g_ctx.t0, g_ctx.t1));
    return MSL_OK;
}

int msl_graph_begin(void) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (g_ctx.capturing) return fail(MSL_E_STATE, "msl_graph_begin: capture already active");
    MSL_HIP_TRY(hipStreamBeginCapture(g_ctx.stream, hipStreamCaptureModeThreadLocal));
    g_ctx.capturing = true;
    return MSL_OK;
}

int msl_graph_end(void **graph) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(graph != nullptr, "msl_graph_end: null out-pointer");
    if (!g_ctx.capturing) return fail(MSL_E_STATE, "msl_graph_end: no capture active");
    hipGraph_t g = nullptr;
    g_ctx.capturing = false;
    MSL_HIP_TRY(hipStreamEndCapture(g_ctx.stream, &g));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph = (void *)exec;
    return MSL_OK;
}

int msl_graph_launch(void *graph) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(graph != nullptr, "msl_graph_launch: null graph");
    MSL_HIP_TRY(hipGraphLaunch((hipGraphExec_t)graph, g_ctx.stream));
    return MSL_OK;
}

int msl_graph_destroy(void *graph) {
    if (!graph) return MSL_OK;
    MSL_HIP_TRY(hipGraphExecDestroy((hipGraphExec_t)graph));
    return MSL_OK;
}

}  // extern "C"

// 3-D 7-point heat step on an nx x ny x nz cube (x fastest).
//
// Reference: DC::mapStencil<heat3D> runs one thread per cell with seven
// bounds-checked reads through PLCube::operator() and no reuse
// (include/detail/map_kernels.cuh:140-149, include/plcube.h:105-119,
// functor examples/heatdiffusion.cpp:55-65).
//
// Here one wavefront owns a 64 x 8 column of the cube and marches a segment
// of z-planes with all neighbours in registers (heat7_march): each cell is
// read from HBM once and written once (8 bytes per cell per step), in the
// reference functor's operation order.  The optional fused residual folds
// max |v - in| into one device scalar with an integer atomic max (the bit
// patterns of non-negative floats order like unsigned integers).
#include <cstdlib>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "heat.hpp"

namespace mslhip {

// One wave per (64 x kTY column tile, z-segment); four independent waves per
// workgroup, no LDS and no barrier.
template <bool kResidual, bool NT>
__global__ __launch_bounds__(256) void heat7_kernel(float *__restrict__ out, const float *__restrict__ in, int nx,
                                                    int ny, int nz, float nv, int zseg, int xtiles, int ytiles,
                                                    int nitems, float *__restrict__ maxdiff) {
    const int lane = threadIdx.x & 63;
    const int item = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
    float dmax = 0.0f;
    if (item < nitems) {
        const int xt = item % xtiles;
        const int rest = item / xtiles;
        const int yt = rest % ytiles;
        const int zs = rest / ytiles;
        const int z0 = zs * zseg;
        heat7_march<kResidual, false, NT>(out, in, nx, ny, nz, nv, xt * kTX, yt * kTY, z0, min(z0 + zseg, nz), &dmax,
                                          nullptr);
    }
    if (kResidual) {
        __shared__ float wmax[4];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const float o = __shfl_xor(dmax, m);
            dmax = o > dmax ? o : dmax;
        }
        if (lane == 0) wmax[threadIdx.x >> 6] = dmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            const float m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
            atomicMax(reinterpret_cast<unsigned int *>(maxdiff), __float_as_uint(m));
        }
    }
}

// Two generations per launch (heat7x2_march); tiles of kT2X x kT2Y columns.
template <bool kResidual, bool NT>
__global__ __launch_bounds__(256) void heat7x2_kernel(float *__restrict__ out, const float *__restrict__ in, int nx,
                                                      int ny, int nz, float nv, int zseg, int xtiles, int ytiles,
                                                      int nitems, float *__restrict__ maxdiff) {
    const int lane = threadIdx.x & 63;
    const int item = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
    float dmax = 0.0f;
    if (item < nitems) {
        const int xt = item % xtiles;
        const int rest = item / xtiles;
        const int yt = rest % ytiles;
        const int zs = rest / ytiles;
        const int z0 = zs * zseg;
        heat7x2_march<kResidual, NT>(out, in, nx, ny, nz, nv, xt * kT2X, yt * kT2Y, z0, min(z0 + zseg, nz), &dmax);
    }
    if (kResidual) {
        __shared__ float wmax[4];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const float o = __shfl_xor(dmax, m);
            dmax = o > dmax ? o : dmax;
        }
        if (lane == 0) wmax[threadIdx.x >> 6] = dmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            const float m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
            atomicMax(reinterpret_cast<unsigned int *>(maxdiff), __float_as_uint(m));
        }
    }
}

struct HeatShape {
    int zseg, xtiles, ytiles, zsegs, nitems;
};

inline HeatShape heat_shape(int nx, int ny, int nz) {
    HeatShape s;
    s.xtiles = (nx + kTX - 1) / kTX;
    s.ytiles = (ny + kTY - 1) / kTY;
    const long cols = (long)s.xtiles * s.ytiles;
    const long target = (long)ctx().num_cus * 32;  // waves wanted
    long zsegs = (target + cols - 1) / cols;
    if (zsegs < 1) zsegs = 1;
    int zseg = (int)((nz + zsegs - 1) / zsegs);
    zseg = (zseg + kP - 1) / kP * kP;  // whole blocks of the plane ring
    if (zseg < 16) zseg = 16;
    if (int z = tune_int("MSL_HEAT_ZSEG", 0)) zseg = z;
    s.zseg = zseg;
    s.zsegs = (nz + zseg - 1) / zseg;
    s.nitems = s.xtiles * s.ytiles * s.zsegs;
    return s;
}

// Cubes above this size run their iterations two generations per launch:
// the pass reads and writes half the bytes per generation but computes 1.3x
// the cells (tile halos), which pays where HBM or the Infinity Cache bounds
// the single step (measured: 272^3 4.4 -> 5.4 TB/s, 320^3 5.0 -> 5.6 TB/s,
// 512^3 4.9 -> 6.5 TB/s) and not on smaller cubes (256^3 5.6 -> 4.8 TB/s).
constexpr size_t kHeatPairMinBytes = (size_t)64 << 20;

// Two generations in one launch; maxdiff (if any) gets the second one's
// max |change|.
int heat_step2(float *out, const float *in, int nx, int ny, int nz, float nv, float *maxdiff) {
    HeatShape s = heat_shape(nx, ny, nz);
    s.xtiles = (nx + kT2X - 1) / kT2X;
    s.ytiles = (ny + kT2Y - 1) / kT2Y;
    s.nitems = s.xtiles * s.ytiles * s.zsegs;
    const bool nt = stream_stores((size_t)nx * ny * nz * sizeof(float));
    const int blocks = (s.nitems + 3) / 4;
    if (maxdiff && nt)
        heat7x2_kernel<true, true><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles,
                                                                     s.ytiles, s.nitems, maxdiff);
    else if (maxdiff)
        heat7x2_kernel<true, false><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles,
                                                                      s.ytiles, s.nitems, maxdiff);
    else if (nt)
        heat7x2_kernel<false, true><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles,
                                                                      s.ytiles, s.nitems, nullptr);
    else
        heat7x2_kernel<false, false><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles,
                                                                       s.ytiles, s.nitems, nullptr);
    return check_launch("msl_heat7_iterate_f32");
}

int heat_step(float *out, const float *in, int nx, int ny, int nz, float nv, float *maxdiff) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(nx >= 0 && ny >= 0 && nz >= 0, "msl_heat7_f32: negative shape");
    MSL_REQUIRE((int64_t)nx * ny * nz <= ((int64_t)1 << 40), "msl_heat7_f32: cube too large");
    MSL_REQUIRE((int64_t)nx * ny < ((int64_t)1 << 30), "msl_heat7_f32: plane of 2^30 cells or more");
    if (nx == 0 || ny == 0 || nz == 0) return MSL_OK;
    MSL_REQUIRE(out && in && out != in, "msl_heat7_f32: null or aliased buffers");
    HeatShape s = heat_shape(nx, ny, nz);
    const bool nt = stream_stores((size_t)nx * ny * nz * sizeof(float));
    const int blocks = (s.nitems + 3) / 4;
    if (maxdiff && nt)
        heat7_kernel<true, true><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles, s.ytiles,
                                                                   s.nitems, maxdiff);
    else if (maxdiff)
        heat7_kernel<true, false><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles, s.ytiles,
                                                                    s.nitems, maxdiff);
    else if (nt)
        heat7_kernel<false, true><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles, s.ytiles,
                                                                    s.nitems, nullptr);
    else
        heat7_kernel<false, false><<<blocks, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles,
                                                                     s.ytiles, s.nitems, nullptr);
    return check_launch("msl_heat7_f32");
}

// Small fill used instead of hipMemsetAsync so that captured graphs contain
// kernel nodes only.
__global__ void zero_f32_kernel(float *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Generations per launch of msl_heat7_iterate_f32, in launch order.  Two
// generations per launch on cubes above kHeatPairMinBytes (MSL_HEAT_GENS =
// 1 or 2 forces one or two, read on every call so a process can compare
// both).  A residual generation must be the last of its launch, and the
// number of launches must have the parity of iters (the result ends in b
// exactly when iters is odd, as with one launch per generation).
static std::vector<int> heat_iterate_plan(int nx, int ny, int nz, int iters, int residual_every) {
    const char *ge = getenv("MSL_HEAT_GENS");
    const int gens_env = ge ? atoi(ge) : 0;
    const bool pairs = gens_env == 2 || (gens_env != 1 && (size_t)nx * ny * nz * sizeof(float) > kHeatPairMinBytes);
    auto is_res = [&](int it) { return residual_every > 0 && ((it + 1) % residual_every == 0); };
    std::vector<int> plan;
    for (int it = 0; it < iters; it += plan.back())
        plan.push_back((pairs && it + 1 < iters && !is_res(it)) ? 2 : 1);
    if ((int)(plan.size() & 1) != (iters & 1)) {
        for (size_t j = plan.size(); j-- > 0;)
            if (plan[j] == 2) {
                plan[j] = 1;
                plan.insert(plan.begin() + j + 1, 1);
                break;
            }
    }
    return plan;
}

struct HeatGraph {
    hipGraphExec_t exec = nullptr;
    float *d_res = nullptr;  // residual slots owned by this graph
};
struct HeatGraphs {
    std::mutex mu;
    std::map<std::string, HeatGraph> graphs;
};
static HeatGraphs g_heat_graphs;

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_heat7_f32(float *d_out, const float *d_in, int nx, int ny, int nz, float nv) {
    return heat_step(d_out, d_in, nx, ny, nz, nv, nullptr);
}

int msl_heat7_residual_f32(float *d_out, const float *d_in, int nx, int ny, int nz, float nv, float *d_maxdiff) {
    MSL_REQUIRE(d_maxdiff != nullptr, "msl_heat7_residual_f32: null residual pointer");
    return heat_step(d_out, d_in, nx, ny, nz, nv, d_maxdiff);
}

int msl_heat7_iterate_f32(float *d_a, float *d_b, int nx, int ny, int nz, float nv, int iters, int residual_every,
                          float *h_residuals, int *h_result_in_b) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(iters >= 0 && residual_every >= 0, "msl_heat7_iterate_f32: negative count");
    MSL_REQUIRE(nx >= 0 && ny >= 0 && nz >= 0, "msl_heat7_iterate_f32: negative shape");
    MSL_REQUIRE((int64_t)nx * ny < ((int64_t)1 << 30) && (int64_t)nx * ny * nz <= ((int64_t)1 << 40),
                "msl_heat7_iterate_f32: cube too large");
    const int nres = residual_every > 0 ? iters / residual_every : 0;
    MSL_REQUIRE(nres == 0 || h_residuals != nullptr, "msl_heat7_iterate_f32: residual output required");
    if (h_result_in_b) *h_result_in_b = iters & 1;
    if (iters == 0 || nx == 0 || ny == 0 || nz == 0) return MSL_OK;
    const std::vector<int> plan = heat_iterate_plan(nx, ny, nz, iters, residual_every);
    const bool pairs = std::find(plan.begin(), plan.end(), 2) != plan.end();
    char keybuf[256];
    snprintf(keybuf, sizeof keybuf, "heat7|%p|%p|%d|%d|%d|%08x|%d|%d|%d", (void *)d_a, (void *)d_b, nx, ny, nz,
             __builtin_bit_cast(unsigned, nv), iters, residual_every, pairs ? 2 : 1);
    std::string key(keybuf);
    HeatGraph hg;
    {
        std::lock_guard<std::mutex> lock(g_heat_graphs.mu);
        auto it = g_heat_graphs.graphs.find(key);
        if (it != g_heat_graphs.graphs.end()) hg = it->second;
    }
    if (!hg.exec) {
        if (nres > 0) MSL_HIP_TRY(hipMalloc(&hg.d_res, (size_t)nres * sizeof(float)));
        float *d_res = hg.d_res;
        rc = msl_graph_begin();
        if (rc) {
            if (hg.d_res) hipFree(hg.d_res);
            return rc;
        }
        int brc = MSL_OK;
        if (nres > 0) {
            zero_f32_kernel<<<(nres + 255) / 256, 256, 0, ctx().stream>>>(d_res, nres);
            brc = check_launch("zero residuals");
        }
        auto is_res = [&](int it) { return residual_every > 0 && ((it + 1) % residual_every == 0); };
        int k = 0, it = 0;
        for (size_t j = 0; j < plan.size() && brc == MSL_OK; j++) {
            float *src = (j & 1) ? d_b : d_a;
            float *dst = (j & 1) ? d_a : d_b;
            it += plan[j];
            const bool res = is_res(it - 1);
            brc = plan[j] == 2 ? heat_step2(dst, src, nx, ny, nz, nv, res ? d_res + (k++) : nullptr)
                               : heat_step(dst, src, nx, ny, nz, nv, res ? d_res + (k++) : nullptr);
        }
        void *g = nullptr;
        int erc = msl_graph_end(&g);
        if (brc || erc) {
            if (g) msl_graph_destroy(g);
            if (hg.d_res) hipFree(hg.d_res);
            return brc ? brc : erc;
        }
        hg.exec = (hipGraphExec_t)g;
        std::lock_guard<std::mutex> lock(g_heat_graphs.mu);
        if (g_heat_graphs.graphs.size() > 32) {
            hipStreamSynchronize(ctx().stream);
            for (auto &kv : g_heat_graphs.graphs) {
                hipGraphExecDestroy(kv.second.exec);
                if (kv.second.d_res) hipFree(kv.second.d_res);
            }
            g_heat_graphs.graphs.clear();
        }
        g_heat_graphs.graphs[key] = hg;
    }
    hipGraphExec_t exec = hg.exec;
    float *d_res = hg.d_res;
    MSL_HIP_TRY(hipGraphLaunch(exec, ctx().stream));
    if (nres > 0) {
        MSL_HIP_TRY(hipMemcpyAsync(h_residuals, d_res, (size_t)nres * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx().stream));
        MSL_HIP_TRY(hipStreamSynchronize(ctx().stream));
    }
    return MSL_OK;
}

int msl_heat7_iterate_plan(int nx, int ny, int nz, int iters, int residual_every, int *h_gens, int capacity,
                           int *h_launches) {
    MSL_REQUIRE(iters >= 0 && residual_every >= 0 && nx >= 0 && ny >= 0 && nz >= 0,
                "msl_heat7_iterate_plan: negative argument");
    MSL_REQUIRE(h_launches != nullptr && capacity >= 0, "msl_heat7_iterate_plan: bad output arguments");
    std::vector<int> plan;
    if (iters > 0 && nx > 0 && ny > 0 && nz > 0) plan = heat_iterate_plan(nx, ny, nz, iters, residual_every);
    *h_launches = (int)plan.size();
    MSL_REQUIRE(h_gens != nullptr || plan.empty() || capacity == 0, "msl_heat7_iterate_plan: null output");
    for (size_t k = 0; k < plan.size() && (int)k < capacity; k++) h_gens[k] = plan[k];
    return MSL_OK;
}

}  // extern "C"

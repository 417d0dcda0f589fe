// 3-D 7-point heat step on an nx x ny x nz cube (x fastest).
//
// Reference: DC::mapStencil<heat3D> runs one thread per cell with seven
// bounds-checked reads through PLCube::operator() and no reuse
// (include/detail/map_kernels.cuh:140-149, include/plcube.h:105-119,
// functor examples/heatdiffusion.cpp:55-65).
//
// Here a 256-thread workgroup owns a 64 x (4*RPT) column of the cube and
// marches a segment of z-planes (2.5-D blocking): each thread keeps the z-1,
// z and z+1 values of its RPT cells in registers (plane z+1 prefetched one
// step ahead), the current plane plus a one-cell x/y halo is staged in LDS
// (double-buffered, one barrier per plane), and the cell is
//   v = 0 + in(x-1) + in(x+1) + in(y-1) + in(y+1) + in(z-1) + in(z+1); v / 6
// in exactly that order.  Each cell is read from HBM once and written once
// (8 bytes per cell per step); the optional fused residual folds
// max |v - in| into one device scalar with an integer atomic max (the bit
// patterns of non-negative floats order like unsigned integers).
#include <map>
#include <mutex>
#include <string>

#include "common.hpp"

namespace mslhip {

constexpr int kTX = 64;  // x extent of a tile (one wave row)
constexpr int kWY = 4;   // waves per workgroup (y)
constexpr int kRPT = 4;  // y rows per thread
constexpr int kTY = kWY * kRPT;
constexpr int kP = 8;    // plane ring: z-1 .. z+6 (planes z+2 .. z+6 in flight)

// Branch-free cell read: the address is clamped into the cube and the value
// replaced by nv when the cell lies outside, so loads are never predicated
// and the compiler's memory-counter waits stay exact across the march.
__device__ __forceinline__ float ld_cell(const float *__restrict__ in, int x, int y, int z, int nx, int ny, int nz,
                                         float nv) {
    const int xc = min(max(x, 0), nx - 1);
    const int yc = min(max(y, 0), ny - 1);
    const int zc = min(max(z, 0), nz - 1);
    const float v = in[((size_t)zc * ny + yc) * nx + xc];
    return (x == xc && y == yc && z == zc) ? v : nv;
}

template <bool kResidual>
__global__ __launch_bounds__(256) void heat7_kernel(float *__restrict__ out, const float *__restrict__ in, int nx,
                                                    int ny, int nz, float nv, int zseg, int xtiles, int ytiles,
                                                    int nitems, float *__restrict__ maxdiff) {
    __shared__ float tile[2][kTY + 2][kTX + 2];
    __shared__ float wmax[kWY];
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;  // wave index
    const int item = (int)xcd_tile(blockIdx.x, gridDim.x);
    const int xt = item % xtiles;
    const int rest = item / xtiles;
    const int yt = rest % ytiles;
    const int zs = rest / ytiles;
    const int x = xt * kTX + tx;
    const int ybase = yt * kTY;
    const int z0 = zs * zseg;
    const int z1 = min(z0 + zseg, nz);

    // Halo duty for plane loads: wave 0 loads row ybase-1, wave 1 row
    // ybase+kTY, wave 2 lanes 0..kTY-1 the column x=-1 side, lanes
    // 32..32+kTY-1 the x=+kTX side.  Other lanes load a harmless duplicate
    // of their own cell and skip the LDS write.
    int hx = x, hy = ybase + ty, hsx = 0, hsy = 0;
    bool hal = false;
    if (ty == 0) {
        hal = true; hy = ybase - 1; hsx = tx + 1; hsy = 0;
    } else if (ty == 1) {
        hal = true; hy = ybase + kTY; hsx = tx + 1; hsy = kTY + 1;
    } else if (ty == 2 && tx < kTY) {
        hal = true; hx = xt * kTX - 1; hy = ybase + tx; hsx = 0; hsy = tx + 1;
    } else if (ty == 2 && tx >= 32 && tx < 32 + kTY) {
        hal = true; hx = xt * kTX + kTX; hy = ybase + (tx - 32); hsx = kTX + 1; hsy = tx - 32 + 1;
    }

    // Ring of planes per cell: slot(zz) = (zz - z0 + 1) & (kP-1).
    float q[kRPT][kP];
    float hq[kP];
#pragma unroll
    for (int k = 0; k < kP - 1; k++) {
        const int zz = z0 - 1 + k;
#pragma unroll
        for (int r = 0; r < kRPT; r++) q[r][k] = ld_cell(in, x, ybase + ty + r * kWY, zz, nx, ny, nz, nv);
        hq[k] = ld_cell(in, hx, hy, zz, nx, ny, nz, nv);
    }
    float dmax = 0.0f;
    // Whole blocks of kP planes; planes past z1 are computed but not stored.
    for (int zb = z0; zb < z1; zb += kP) {
#pragma unroll
        for (int u = 0; u < kP; u++) {
            const int z = zb + u;
            const int sb = u & (kP - 1);        // z-1
            const int sc = (u + 1) & (kP - 1);  // z
            const int sa = (u + 2) & (kP - 1);  // z+1
            float (*tl)[kTX + 2] = tile[u & 1];
            // stage plane z
#pragma unroll
            for (int r = 0; r < kRPT; r++) tl[ty + r * kWY + 1][tx + 1] = q[r][sc];
            if (hal) tl[hsy][hsx] = hq[sc];
            // prefetch plane z+6 into the slot of plane z+6 (= slot of z-2)
            {
                const int sn = (u + 7) & (kP - 1);
                const int zz = z + 6;
#pragma unroll
                for (int r = 0; r < kRPT; r++) q[r][sn] = ld_cell(in, x, ybase + ty + r * kWY, zz, nx, ny, nz, nv);
                hq[sn] = ld_cell(in, hx, hy, zz, nx, ny, nz, nv);
            }
            lds_barrier();
#pragma unroll
            for (int r = 0; r < kRPT; r++) {
                const int ly = ty + r * kWY + 1;
                const int y = ybase + ty + r * kWY;
                float v = 0.0f;
                v = __fadd_rn(v, tl[ly][tx]);          // x-1
                v = __fadd_rn(v, tl[ly][tx + 2]);      // x+1
                v = __fadd_rn(v, tl[ly - 1][tx + 1]);  // y-1
                v = __fadd_rn(v, tl[ly + 1][tx + 1]);  // y+1
                v = __fadd_rn(v, q[r][sb]);            // z-1
                v = __fadd_rn(v, q[r][sa]);            // z+1
                v = __fdiv_rn(v, 6.0f);
                if (z < z1 && x < nx && y < ny) {
                    __builtin_nontemporal_store(v, out + ((size_t)z * ny + y) * nx + x);
                    if (kResidual) {
                        float d = fabsf(__fsub_rn(v, q[r][sc]));
                        dmax = d > dmax ? d : dmax;
                    }
                }
            }
        }
    }
    if (kResidual) {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            float o = __shfl_xor(dmax, m);
            dmax = o > dmax ? o : dmax;
        }
        if (tx == 0) wmax[ty] = dmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = wmax[0];
            for (int w = 1; w < kWY; w++) m = wmax[w] > m ? wmax[w] : m;
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
    long cols = (long)s.xtiles * s.ytiles;
    long target = (long)ctx().num_cus * 8;  // workgroups wanted
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

int heat_step(float *out, const float *in, int nx, int ny, int nz, float nv, float *maxdiff) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(nx >= 0 && ny >= 0 && nz >= 0, "msl_heat7_f32: negative shape");
    MSL_REQUIRE((int64_t)nx * ny * nz <= ((int64_t)1 << 40), "msl_heat7_f32: cube too large");
    if (nx == 0 || ny == 0 || nz == 0) return MSL_OK;
    MSL_REQUIRE(out && in && out != in, "msl_heat7_f32: null or aliased buffers");
    HeatShape s = heat_shape(nx, ny, nz);
    if (maxdiff)
        heat7_kernel<true><<<s.nitems, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles, s.ytiles,
                                                                s.nitems, maxdiff);
    else
        heat7_kernel<false><<<s.nitems, 256, 0, ctx().stream>>>(out, in, nx, ny, nz, nv, s.zseg, s.xtiles, s.ytiles,
                                                                 s.nitems, nullptr);
    return check_launch("msl_heat7_f32");
}

// Small fill used instead of hipMemsetAsync so that captured graphs contain
// kernel nodes only.
__global__ void zero_f32_kernel(float *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
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
    const int nres = residual_every > 0 ? iters / residual_every : 0;
    MSL_REQUIRE(nres == 0 || h_residuals != nullptr, "msl_heat7_iterate_f32: residual output required");
    if (h_result_in_b) *h_result_in_b = iters & 1;
    if (iters == 0 || nx == 0 || ny == 0 || nz == 0) return MSL_OK;
    char keybuf[256];
    snprintf(keybuf, sizeof keybuf, "heat7|%p|%p|%d|%d|%d|%08x|%d|%d", (void *)d_a, (void *)d_b, nx, ny, nz,
             __builtin_bit_cast(unsigned, nv), iters, residual_every);
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
        int k = 0;
        for (int it = 0; it < iters && brc == MSL_OK; it++) {
            float *src = (it & 1) ? d_b : d_a;
            float *dst = (it & 1) ? d_a : d_b;
            bool res = residual_every > 0 && ((it + 1) % residual_every == 0);
            brc = heat_step(dst, src, nx, ny, nz, nv, res ? d_res + (k++) : nullptr);
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

}  // extern "C"

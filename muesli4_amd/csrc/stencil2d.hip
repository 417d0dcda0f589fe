// 2-D stencils on a rows x cols grid: mean blur of radius R and Game of Life.
//
// Reference: DM::mapStencil<f> launches one thread per output that makes
// (2r+1)^2 bounds-checked global reads through NPLMatrix::operator()
// (include/detail/map_kernels.cuh:159-168, include/nplmatrix.h:123-139).
//
// Here one wavefront owns a strip of 256 columns (64 lanes x 4 columns) and a
// segment of rows, and marches down the segment:
//   1. the input row y+R (strip plus R halo columns each side) is loaded with
//      16-byte loads, rows prefetched D ahead in registers, and staged in a
//      per-wave LDS row;
//   2. each lane reads its 4+2R window back with 16-byte LDS reads and forms
//      4 horizontal row sums (dx ascending from 0.0f);
//   3. row sums live in a register ring of 2R+1 entries (the loop is unrolled
//      over the ring period so every ring index is static) and each output is
//      the ordered sum of its 2R+1 row sums (dy ascending from 0.0f), divided
//      by (2R+1)^2, correctly rounded (div_odd_square: multiply plus one FMA
//      correction, verified bit-identical to IEEE division on all inputs).
// This keeps the exact operation order of the scalar definition (bit-exact)
// while doing 2(2R+1) adds per output instead of (2R+1)^2, reading each
// input element once from HBM (halo rows/columns come from L2) and writing
// each output once: 8 bytes per cell per step.
#include <map>
#include <mutex>
#include <string>

#include "common.hpp"

namespace mslhip {

constexpr int kStripW = 256;  // columns per wave strip

template <typename T>
struct Vec4 {
    T v[4];
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef int32_t v4i __attribute__((ext_vector_type(4)));
template <typename T>
struct NativeVec;
template <>
struct NativeVec<float> {
    typedef v4f type;
};
template <>
struct NativeVec<int32_t> {
    typedef v4i type;
};

// Branch-free read of the 4 columns [c, c+4) of row yy: the row and the
// columns are clamped into the grid so the load is always issued at a valid
// address, and values outside the grid are replaced by nv afterwards.  With
// VEC (cols % 4 == 0, 16-byte aligned rows, c % 4 == 0) a quad is either
// wholly inside or wholly outside the grid and is read with one 16-byte load.
template <typename T, bool VEC>
__device__ __forceinline__ Vec4<T> read_quad(const T *__restrict__ in, int yy, int c, int rows, int cols, T nv) {
    typedef typename NativeVec<T>::type V;
    const int yc = min(max(yy, 0), rows - 1);
    const bool rowok = yy == yc;
    const T *rowp = in + (size_t)yc * cols;
    Vec4<T> q;
    if constexpr (VEC) {
        const int cc = min(max(c, 0), cols - 4);
        const bool ok = rowok && cc == c;
        const V v = *reinterpret_cast<const V *>(rowp + cc);
        q.v[0] = ok ? v.x : nv;
        q.v[1] = ok ? v.y : nv;
        q.v[2] = ok ? v.z : nv;
        q.v[3] = ok ? v.w : nv;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ck = c + k;
            const int cc = min(max(ck, 0), cols - 1);
            const T v = rowp[cc];
            q.v[k] = (rowok && cc == ck) ? v : nv;
        }
    }
    return q;
}

// ---- per-stencil policies -------------------------------------------------
template <int R_>
struct MeanPolicy {
    static constexpr int R = R_;
    static constexpr bool kCenter = false;
    typedef float T;
    float denom;  // (2R+1)^2
    float rcp;    // RN(1 / denom)
    // Row sum of window w (index RP-R+c+d holds column x0+c+d-R).
    template <int RP>
    __device__ __forceinline__ float rowsum(const float *w, int c) const {
        float s = 0.0f;
#pragma unroll
        for (int d = 0; d <= 2 * R; d++) s = __fadd_rn(s, w[RP - R + c + d]);
        return s;
    }
    // Correctly rounded colsum / denom (identical to IEEE division).
    __device__ __forceinline__ float finish(float colsum, float /*center*/) const {
        return div_odd_square(colsum, denom, rcp);
    }
    __device__ __forceinline__ float colsum_init() const { return 0.0f; }
    __device__ __forceinline__ float acc(float a, float b) const { return __fadd_rn(a, b); }
};

struct LifePolicy {
    static constexpr int R = 1;
    static constexpr bool kCenter = true;
    typedef int32_t T;
    float denom;  // unused
    template <int RP>
    __device__ __forceinline__ int32_t rowsum(const int32_t *w, int c) const {
        return w[RP - 1 + c] + w[RP + c] + w[RP + 1 + c];
    }
    // B3/S23 as in examples/gameoflife.cpp:18-43 (sum of 8 neighbours).
    __device__ __forceinline__ int32_t finish(int32_t box, int32_t live) const {
        int32_t sum = box - live;
        int32_t next = 0;
        if (live == 1 && (sum == 2 || sum == 3)) next = 1;
        if (live == 0 && sum == 3) next = 1;
        return next;
    }
    __device__ __forceinline__ int32_t colsum_init() const { return 0; }
    __device__ __forceinline__ int32_t acc(int32_t a, int32_t b) const { return a + b; }
};

template <int R>
struct StripConsts {
    static constexpr int RP = (R + 3) / 4 * 4;     // halo padded to whole quads
    static constexpr int LDSW = kStripW + 2 * RP;  // elements per staged row
    static constexpr int NQ = (2 * RP) / 4 + 1;    // 16-byte window reads per lane
    static constexpr int PER = 2 * R + 1;          // ring period (= loop unroll)
    // Rows requested ahead of use: a divisor of PER (so queue slots stay
    // static), larger where the per-row work is too short to hide latency.
    static constexpr int D = PER <= 7 ? PER : (R > 10 ? 1 : (PER % 5 == 0 ? 5 : (PER % 3 == 0 ? 3 : 1)));
};

// March one wave down its (strip, segment).  One code path for every tile:
// reads are clamped (read_quad), stores are masked per lane, and the loop is
// unrolled over exactly one ring period so the code stays small (the
// instruction cache is shared by two CUs) while every register index is a
// compile-time constant.  Row t+D is requested when row t is staged.
template <typename Pol, bool VEC>
__device__ __forceinline__ void strip_march(typename Pol::T *__restrict__ out, const typename Pol::T *__restrict__ in,
                                            int rows, int cols, typename Pol::T nv, int xs, int y0, int yend,
                                            const Pol &pol, typename Pol::T (*lds)[StripConsts<Pol::R>::LDSW]) {
    typedef typename Pol::T T;
    typedef typename NativeVec<T>::type V;
    constexpr int R = Pol::R;
    constexpr int RP = StripConsts<R>::RP;
    constexpr int NQ = StripConsts<R>::NQ;
    constexpr int PER = StripConsts<R>::PER;

    const int lane = threadIdx.x & 63;
    const int x0 = xs + 4 * lane;
    const int total = (yend - y0) + 2 * R;  // input rows y0-R .. yend-1+R
    // Halo quads: left quads for lanes [0, RP/4), right quads for lanes
    // [RP/4, RP/2); every other lane re-reads and re-stages its own centre
    // quad, so no lane diverges.
    const bool has_halo = lane < RP / 2;
    const int hx = has_halo ? (lane < RP / 4 ? xs - RP + 4 * lane : xs + kStripW + 4 * (lane - RP / 4)) : x0;
    const int hslot = has_halo ? (lane < RP / 4 ? 4 * lane : RP + kStripW + 4 * (lane - RP / 4)) : RP + 4 * lane;
    const bool lane_in = x0 < cols;

    constexpr int D = StripConsts<R>::D;
    Vec4<T> ring[PER];  // row sums
    Vec4<T> cring[Pol::kCenter ? PER : 1];
    Vec4<T> qc[D], qh[D];  // requested rows t .. t+D-1 (centre and halo quads)
#pragma unroll
    for (int d = 0; d < D; d++) {
        qc[d] = read_quad<T, VEC>(in, y0 - R + d, x0, rows, cols, nv);
        qh[d] = read_quad<T, VEC>(in, y0 - R + d, hx, rows, cols, nv);
    }

    for (int tb = 0; tb < total; tb += PER) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int t = tb + u;  // ring slot of row t is u (tb is a multiple of PER)
            const int qs = u % D;  // queue slot of row t
            T *buf = lds[t & 1];
            *reinterpret_cast<V *>(buf + RP + 4 * lane) = V{qc[qs].v[0], qc[qs].v[1], qc[qs].v[2], qc[qs].v[3]};
            *reinterpret_cast<V *>(buf + hslot) = V{qh[qs].v[0], qh[qs].v[1], qh[qs].v[2], qh[qs].v[3]};
            // request row t+D into the slot just staged
            qc[qs] = read_quad<T, VEC>(in, y0 - R + t + D, x0, rows, cols, nv);
            qh[qs] = read_quad<T, VEC>(in, y0 - R + t + D, hx, rows, cols, nv);
            wave_lds_sync();
            T w[4 * NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const V v = *reinterpret_cast<const V *>(buf + 4 * lane + 4 * q);
                w[4 * q + 0] = v.x;
                w[4 * q + 1] = v.y;
                w[4 * q + 2] = v.z;
                w[4 * q + 3] = v.w;
            }
#pragma unroll
            for (int c = 0; c < 4; c++) ring[u].v[c] = pol.template rowsum<RP>(w, c);
            if constexpr (Pol::kCenter) {
#pragma unroll
                for (int c = 0; c < 4; c++) cring[u].v[c] = w[RP + c];
            }
            const int y = y0 - 2 * R + t;  // output row completed by row t
            if (t >= 2 * R && y < yend) {
                Vec4<T> o;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    T sum = pol.colsum_init();
#pragma unroll
                    for (int k = 0; k < PER; k++) sum = pol.acc(sum, ring[(u + 1 + k) % PER].v[c]);
                    T centre = T(0);
                    if constexpr (Pol::kCenter) centre = cring[(u + 1 + R) % PER].v[c];
                    o.v[c] = pol.finish(sum, centre);
                }
                T *orow = out + (size_t)y * cols;
                if constexpr (VEC) {
                    if (lane_in)
                        __builtin_nontemporal_store(V{o.v[0], o.v[1], o.v[2], o.v[3]},
                                                    reinterpret_cast<V *>(orow + x0));
                } else {
#pragma unroll
                    for (int c = 0; c < 4; c++)
                        if (x0 + c < cols) orow[x0 + c] = o.v[c];
                }
            }
        }
    }
}

// One wave per (strip, segment) item, four independent waves per 256-thread
// workgroup; each wave stages rows in its own LDS region, so only
// wave-level ordering is needed (no workgroup barrier).
template <typename Pol, bool VEC>
__global__ __launch_bounds__(256) void strip_stencil_kernel(typename Pol::T *__restrict__ out,
                                                            const typename Pol::T *__restrict__ in, int rows, int cols,
                                                            typename Pol::T nv, int seg, int strips_x, int nitems,
                                                            Pol pol) {
    typedef typename Pol::T T;
    constexpr int R = Pol::R;
    __shared__ __attribute__((aligned(16))) T lds[4][2][StripConsts<R>::LDSW];

    const int wave = threadIdx.x >> 6;
    const int item = (int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave;
    if (item >= nitems) return;
    const int sx = item % strips_x;
    const int sy = item / strips_x;
    const int y0 = sy * seg;
    strip_march<Pol, VEC>(out, in, rows, cols, nv, sx * kStripW, y0, min(y0 + seg, rows), pol, lds[wave]);
}

struct StripShape {
    int seg, strips_x, segs, nitems;
};

// Segment length: about 16 waves per CU.
inline StripShape strip_shape(int rows, int cols, int seg) {
    StripShape s;
    s.strips_x = (cols + kStripW - 1) / kStripW;
    if (seg <= 0) seg = tune_int("MSL_STRIP_SEG", 0);
    if (seg <= 0) {
        long target = (long)ctx().num_cus * 16;
        long segs = (target + s.strips_x - 1) / s.strips_x;
        seg = (int)((rows + segs - 1) / segs);
        if (seg < 8) seg = 8;
    }
    s.seg = seg;
    s.segs = (rows + seg - 1) / seg;
    s.nitems = s.strips_x * s.segs;
    return s;
}

template <typename Pol>
int launch_strip(typename Pol::T *out, const typename Pol::T *in, int rows, int cols, typename Pol::T nv, Pol pol,
                 int seg, const char *name) {
    StripShape s = strip_shape(rows, cols, seg);
    const bool vec = ((cols & 3) == 0) && ((((uintptr_t)in) & 15u) == 0) && ((((uintptr_t)out) & 15u) == 0);
    const int blocks = (s.nitems + 3) / 4;
    if (vec)
        strip_stencil_kernel<Pol, true><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg, s.strips_x,
                                                                           s.nitems, pol);
    else
        strip_stencil_kernel<Pol, false><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg, s.strips_x,
                                                                            s.nitems, pol);
    return check_launch(name);
}

template <int R>
int launch_mean_r(float *out, const float *in, int rows, int cols, float nv, int seg) {
    MeanPolicy<R> pol;
    pol.denom = (float)((2 * R + 1) * (2 * R + 1));
    pol.rcp = 1.0f / pol.denom;
    return launch_strip(out, in, rows, cols, nv, pol, seg, "msl_stencil2d_mean_f32");
}

typedef int (*MeanLauncher)(float *, const float *, int, int, float, int);
static const MeanLauncher kMeanLaunchers[] = {
    nullptr,           launch_mean_r<1>,  launch_mean_r<2>,  launch_mean_r<3>,  launch_mean_r<4>,
    launch_mean_r<5>,  launch_mean_r<6>,  launch_mean_r<7>,  launch_mean_r<8>,  launch_mean_r<9>,
    launch_mean_r<10>, launch_mean_r<11>, launch_mean_r<12>, launch_mean_r<13>, launch_mean_r<14>,
    launch_mean_r<15>, launch_mean_r<16>};
constexpr int kMaxMeanR = 16;

int mean_step(float *out, const float *in, int rows, int cols, int r, float nv, int seg) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(rows >= 0 && cols >= 0, "msl_stencil2d_mean_f32: negative shape");
    MSL_REQUIRE(r >= 1 && r <= kMaxMeanR, "msl_stencil2d_mean_f32: radius must be in [1, 16]");
    MSL_REQUIRE((int64_t)rows * cols <= INT32_MAX, "msl_stencil2d_mean_f32: grid too large");
    if (rows == 0 || cols == 0) return MSL_OK;
    MSL_REQUIRE(out && in && out != in, "msl_stencil2d_mean_f32: null or aliased buffers");
    return kMeanLaunchers[r](out, in, rows, cols, nv, seg);
}

int gol_step(int32_t *out, const int32_t *in, int rows, int cols) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(rows >= 0 && cols >= 0, "msl_gol_i32: negative shape");
    MSL_REQUIRE((int64_t)rows * cols <= INT32_MAX, "msl_gol_i32: grid too large");
    if (rows == 0 || cols == 0) return MSL_OK;
    MSL_REQUIRE(out && in && out != in, "msl_gol_i32: null or aliased buffers");
    LifePolicy pol;
    pol.denom = 1.0f;
    return launch_strip(out, in, rows, cols, (int32_t)0, pol, 0, "msl_gol_i32");
}

// ---- exhaustive check of div_odd_square against IEEE division -------------
__global__ void zero_u64_kernel(unsigned long long *p) { *p = 0ull; }

// Every 32-bit pattern x: compare div_odd_square(x) with x / d bit for bit
// (NaN against NaN counts as equal); one atomic per block.
__global__ __launch_bounds__(256) void div_check_kernel(float d, float rcp, unsigned long long *count) {
    __shared__ unsigned int part[256];
    unsigned int bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < (1ull << 32); i += stride) {
        const float x = __uint_as_float((unsigned int)i);
        const float ref = __fdiv_rn(x, d);
        const float got = div_odd_square(x, d, rcp);
        const bool same = (__float_as_uint(ref) == __float_as_uint(got)) || (ref != ref && got != got);
        bad += same ? 0u : 1u;
    }
    part[threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part[0]) atomicAdd(count, (unsigned long long)part[0]);
}

// ---- graph cache for iteration loops ---------------------------------------
struct GraphCache {
    std::mutex mu;
    std::map<std::string, hipGraphExec_t> graphs;
};
static GraphCache g_cache;

template <typename Body>
int run_cached_graph(const std::string &key, Body body) {
    hipGraphExec_t exec = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_cache.mu);
        auto it = g_cache.graphs.find(key);
        if (it != g_cache.graphs.end()) exec = it->second;
    }
    if (!exec) {
        int rc = msl_graph_begin();
        if (rc) return rc;
        int brc = body();
        void *g = nullptr;
        int erc = msl_graph_end(&g);
        if (brc) {
            if (g) msl_graph_destroy(g);
            return brc;
        }
        if (erc) return erc;
        exec = (hipGraphExec_t)g;
        std::lock_guard<std::mutex> lock(g_cache.mu);
        if (g_cache.graphs.size() > 64) {
            for (auto &kv : g_cache.graphs) hipGraphExecDestroy(kv.second);
            g_cache.graphs.clear();
        }
        g_cache.graphs[key] = exec;
    }
    MSL_HIP_TRY(hipGraphLaunch(exec, ctx().stream));
    return MSL_OK;
}

std::string graph_key(const char *kind, const void *a, const void *b, long p0, long p1, long p2, long p3, long p4) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s|%p|%p|%ld|%ld|%ld|%ld|%ld", kind, a, b, p0, p1, p2, p3, p4);
    return std::string(buf);
}

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_stencil2d_mean_f32(float *d_out, const float *d_in, int rows, int cols, int r, float nv) {
    return mean_step(d_out, d_in, rows, cols, r, nv, 0);
}

int msl_stencil2d_mean_iterate_f32(float *d_a, float *d_b, int rows, int cols, int r, float nv, int iters,
                                   int *h_result_in_b) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(iters >= 0, "msl_stencil2d_mean_iterate_f32: negative iteration count");
    MSL_REQUIRE(r >= 1 && r <= kMaxMeanR, "msl_stencil2d_mean_iterate_f32: radius must be in [1, 16]");
    MSL_REQUIRE(rows >= 0 && cols >= 0, "msl_stencil2d_mean_iterate_f32: negative shape");
    if (h_result_in_b) *h_result_in_b = iters & 1;
    if (iters == 0 || rows == 0 || cols == 0) return MSL_OK;
    union {
        float f;
        int32_t i;
    } nvb;
    nvb.f = nv;
    std::string key = graph_key("mean2d", d_a, d_b, rows, cols, r, nvb.i, iters);
    return run_cached_graph(key, [&]() {
        for (int it = 0; it < iters; it++) {
            float *src = (it & 1) ? d_b : d_a;
            float *dst = (it & 1) ? d_a : d_b;
            int brc = mean_step(dst, src, rows, cols, r, nv, 0);
            if (brc) return brc;
        }
        return MSL_OK;
    });
}

int msl_selftest_div_odd_square(int r, uint64_t *h_mismatches) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(r >= 1 && r <= kMaxMeanR, "msl_selftest_div_odd_square: r must be in [1, 16]");
    MSL_REQUIRE(h_mismatches != nullptr, "msl_selftest_div_odd_square: null out-pointer");
    unsigned long long *d_count = nullptr;
    MSL_HIP_TRY(hipMalloc(&d_count, sizeof(unsigned long long)));
    zero_u64_kernel<<<1, 1, 0, ctx().stream>>>(d_count);
    const float d = (float)((2 * r + 1) * (2 * r + 1));
    div_check_kernel<<<ctx().num_cus * 8, 256, 0, ctx().stream>>>(d, 1.0f / d, d_count);
    rc = check_launch("msl_selftest_div_odd_square");
    unsigned long long h = 0;
    hipError_t e = hipMemcpyAsync(&h, d_count, sizeof h, hipMemcpyDeviceToHost, ctx().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx().stream);
    hipFree(d_count);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "msl_selftest_div_odd_square");
    *h_mismatches = h;
    return MSL_OK;
}

int msl_gol_i32(int32_t *d_out, const int32_t *d_in, int rows, int cols) { return gol_step(d_out, d_in, rows, cols); }

int msl_gol_iterate_i32(int32_t *d_a, int32_t *d_b, int rows, int cols, int iters, int *h_result_in_b) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(iters >= 0, "msl_gol_iterate_i32: negative iteration count");
    MSL_REQUIRE(rows >= 0 && cols >= 0, "msl_gol_iterate_i32: negative shape");
    if (h_result_in_b) *h_result_in_b = iters & 1;
    if (iters == 0 || rows == 0 || cols == 0) return MSL_OK;
    std::string key = graph_key("gol", d_a, d_b, rows, cols, iters, 0, 0);
    return run_cached_graph(key, [&]() {
        for (int it = 0; it < iters; it++) {
            int32_t *src = (it & 1) ? d_b : d_a;
            int32_t *dst = (it & 1) ? d_a : d_b;
            int brc = gol_step(dst, src, rows, cols);
            if (brc) return brc;
        }
        return MSL_OK;
    });
}

}  // extern "C"

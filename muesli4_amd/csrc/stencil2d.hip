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
//      by (2R+1)^2 with IEEE division.
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

// Load 4 consecutive columns [c, c+4) of row pointer p (row already known to
// be inside the grid); columns outside [0, cols) read as nv.
template <typename T>
__device__ __forceinline__ Vec4<T> load_quad(const T *__restrict__ p, int c, int cols, T nv, bool vec) {
    Vec4<T> q;
    if (vec && c >= 0 && c + 3 < cols) {
        typename NativeVec<T>::type v = *reinterpret_cast<const typename NativeVec<T>::type *>(p + c);
        q.v[0] = v.x;
        q.v[1] = v.y;
        q.v[2] = v.z;
        q.v[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) q.v[k] = (c + k >= 0 && c + k < cols) ? p[c + k] : nv;
    }
    return q;
}

// ---- per-stencil policies -------------------------------------------------
template <int R_>
struct MeanPolicy {
    static constexpr int R = R_;
    static constexpr bool kCenter = false;
    typedef float T;
    float denom;
    // Row sum of window w (index RP-R+c+d holds column x0+c+d-R).
    template <int RP>
    __device__ __forceinline__ float rowsum(const float *w, int c) const {
        float s = 0.0f;
#pragma unroll
        for (int d = 0; d <= 2 * R; d++) s = __fadd_rn(s, w[RP - R + c + d]);
        return s;
    }
    __device__ __forceinline__ float finish(float colsum, float /*center*/) const { return __fdiv_rn(colsum, denom); }
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
struct PrefetchDepth {
    static constexpr int value = R <= 2 ? 8 : (R <= 5 ? 4 : 2);
};

// One wave per (strip, segment) item; 64-thread workgroups so that the LDS
// hand-off between lanes is ordered by __syncthreads() on a single wave.
template <typename Pol>
__global__ __launch_bounds__(64) void strip_stencil_kernel(typename Pol::T *__restrict__ out,
                                                           const typename Pol::T *__restrict__ in, int rows, int cols,
                                                           typename Pol::T nv, int seg, int strips_x, int nitems,
                                                           int vec, Pol pol) {
    typedef typename Pol::T T;
    constexpr int R = Pol::R;
    constexpr int RP = (R + 3) / 4 * 4;      // halo padded to whole quads
    constexpr int LDSW = kStripW + 2 * RP;   // floats per staged row
    constexpr int NQ = (2 * RP) / 4 + 1;     // 16-byte window reads per lane
    constexpr int PER = 2 * R + 1;           // ring period
    constexpr int D = PrefetchDepth<R>::value;
    constexpr int U = PER * D;               // unroll: multiple of PER, D and 2
    __shared__ __attribute__((aligned(16))) T lds[2][LDSW];

    const int lane = threadIdx.x;
    const int item = (int)xcd_tile(blockIdx.x, gridDim.x);
    if (item >= nitems) return;
    const int sx = item % strips_x;
    const int sy = item / strips_x;
    const int xs = sx * kStripW;
    const int x0 = xs + 4 * lane;
    const int y0 = sy * seg;
    const int yend = min(y0 + seg, rows);
    const int total = (yend - y0) + 2 * R;  // input rows y0-R .. yend-1+R
    // halo quad owned by this lane: left quads for lanes [0, RP/4), right
    // quads for lanes [RP/4, RP/2); column of its first element.
    const bool has_halo = lane < RP / 2;
    const int hx = lane < RP / 4 ? xs - RP + 4 * lane : xs + kStripW + 4 * (lane - RP / 4);
    const int hslot = lane < RP / 4 ? 4 * lane : RP + kStripW + 4 * (lane - RP / 4);
    const bool vecb = vec != 0;

    Vec4<T> pc[D], ph[D];  // prefetched centre / halo quads
    Vec4<T> ring[PER];     // row sums
    Vec4<T> cring[Pol::kCenter ? PER : 1];

    auto fetch = [&](int t, Vec4<T> &c, Vec4<T> &h) {
        const int yy = y0 - R + t;
        if (t < total && yy >= 0 && yy < rows) {
            const T *rowp = in + (size_t)yy * cols;
            c = load_quad<T>(rowp, x0, cols, nv, vecb);
            if (has_halo) h = load_quad<T>(rowp, hx, cols, nv, vecb);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                c.v[k] = nv;
                h.v[k] = nv;
            }
        }
    };

#pragma unroll
    for (int d = 0; d < D; d++) fetch(d, pc[d], ph[d]);

    for (int t0 = 0; t0 < total; t0 += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u;
            if (t < total) {
                T *buf = lds[u & 1];
                // stage row t
                *reinterpret_cast<typename NativeVec<T>::type *>(buf + RP + 4 * lane) =
                    typename NativeVec<T>::type{pc[u % D].v[0], pc[u % D].v[1], pc[u % D].v[2], pc[u % D].v[3]};
                if (has_halo)
                    *reinterpret_cast<typename NativeVec<T>::type *>(buf + hslot) =
                        typename NativeVec<T>::type{ph[u % D].v[0], ph[u % D].v[1], ph[u % D].v[2], ph[u % D].v[3]};
                // prefetch row t + D into the slot just consumed
                fetch(t + D, pc[u % D], ph[u % D]);
                __syncthreads();
                T w[4 * NQ];
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    typename NativeVec<T>::type v =
                        *reinterpret_cast<const typename NativeVec<T>::type *>(buf + 4 * lane + 4 * q);
                    w[4 * q + 0] = v.x;
                    w[4 * q + 1] = v.y;
                    w[4 * q + 2] = v.z;
                    w[4 * q + 3] = v.w;
                }
                Vec4<T> rs;
#pragma unroll
                for (int c = 0; c < 4; c++) rs.v[c] = pol.template rowsum<RP>(w, c);
                ring[u % PER] = rs;
                if constexpr (Pol::kCenter) {
#pragma unroll
                    for (int c = 0; c < 4; c++) cring[u % PER].v[c] = w[RP + c];
                }
                if (t >= 2 * R) {
                    const int y = y0 - R + t - R;
                    Vec4<T> o;
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        T s = pol.colsum_init();
#pragma unroll
                        for (int k = 0; k < PER; k++) s = pol.acc(s, ring[(u + 1 + k) % PER].v[c]);
                        T centre = T(0);
                        if constexpr (Pol::kCenter) centre = cring[(u + 1 + R) % PER].v[c];
                        o.v[c] = pol.finish(s, centre);
                    }
                    T *orow = out + (size_t)y * cols;
                    if (vecb && x0 + 3 < cols) {
                        __builtin_nontemporal_store(typename NativeVec<T>::type{o.v[0], o.v[1], o.v[2], o.v[3]},
                                                    reinterpret_cast<typename NativeVec<T>::type *>(orow + x0));
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; c++)
                            if (x0 + c < cols) orow[x0 + c] = o.v[c];
                    }
                }
            }
        }
    }
}

struct StripShape {
    int seg, strips_x, segs, nitems;
};

inline StripShape strip_shape(int rows, int cols, int seg) {
    StripShape s;
    s.strips_x = (cols + kStripW - 1) / kStripW;
    if (seg <= 0) {
        // Aim for >= 16 waves per CU without very short segments.
        long target = (long)ctx().num_cus * 16;
        long segs = (target + s.strips_x - 1) / s.strips_x;
        seg = (int)((rows + segs - 1) / segs);
        if (seg < 32) seg = 32;
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
    int vec = ((cols & 3) == 0) && ((((uintptr_t)in) & 15u) == 0) && ((((uintptr_t)out) & 15u) == 0);
    strip_stencil_kernel<Pol><<<s.nitems, 64, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg, s.strips_x,
                                                                   s.nitems, vec, pol);
    return check_launch(name);
}

template <int R>
int launch_mean_r(float *out, const float *in, int rows, int cols, float nv, int seg) {
    MeanPolicy<R> pol;
    pol.denom = (float)((2 * R + 1) * (2 * R + 1));
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

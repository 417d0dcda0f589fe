// 2-D stencils on a rows x cols grid: mean blur of radius R and Game of Life.
//
// Reference: DM::mapStencil<f> launches one thread per output that makes
// (2r+1)^2 bounds-checked global reads through NPLMatrix::operator()
// (include/detail/map_kernels.cuh:159-168, include/nplmatrix.h:123-139).
//
// Here one wavefront owns a strip of 256 columns (64 lanes x 4 columns) and
// a segment of rows and marches down it with a register ring of row sums
// (strip_march): each input element is read once from HBM, each output is
// written once, and every output keeps the exact operation order of the
// scalar definition (bit-exact), with 2(2R+1) adds per output instead of
// (2R+1)^2.  The division by (2R+1)^2 is div_odd_square (multiply plus one
// FMA correction, checked bit-identical to IEEE division on all inputs).
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "strip.hpp"

namespace mslhip {

constexpr int kStripW = StripConsts<1, 4>::SW;  // columns per wave strip

// One wave per (strip, segment) item, four independent waves per 256-thread
// workgroup; each wave stages rows in its own LDS region, so only
// wave-level ordering is needed (no workgroup barrier).  C columns per lane:
// 4 (256-column strips) in general, 1 (64-column strips) for small grids,
// where four times as many strips keep more waves busy.
// Radii 9 and 10 with 256-column strips: at most two waves per SIMD (the
// register budget of the wave-uniform row addressing would allow three, but
// three waves per SIMD measured slower: spills, and a partial last round of
// the waves strip_shape launches).
template <int R, int C>
constexpr int strip_max_waves_per_eu() { return (C == 4 && R >= 9 && R <= 10) ? 2 : 8; }

// How the single-grid march ends its last ring period (strip_march_dir's
// kTail): r = 10 runs 128-row segments on streamed grids (see strip_shape),
// 148 input rows, one more than 7 ring periods of 21, so it ends with one
// checked period (8192^2: 4.39 -> 4.69 TB/s against whole periods of 64-row
// segments); every other radius keeps whole periods (r = 9 at 128 rows
// marches 146 of 147 rows anyway and lost 2 % to the checked copy).
template <int R, int C>
constexpr int strip_tail_mode() { return (C == 4 && R == 10) ? 2 : 0; }

template <typename Pol, int MODE, bool NT = true, int C = 4>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, strip_max_waves_per_eu<Pol::R, C>())))
void strip_stencil_kernel(typename Pol::T *__restrict__ out,
                                                            const typename Pol::T *__restrict__ in, int rows, int cols,
                                                            typename Pol::T nv, int seg, int strips_x, int nitems,
                                                            Pol pol) {
    typedef typename Pol::T T;
    constexpr int R = Pol::R;
    __shared__ __attribute__((aligned(16))) T lds[4][2][StripConsts<R, C>::LDSW];

    const int wave = threadIdx.x >> 6;
    // the item is wave-uniform: with readfirstlane the row indices, clamps
    // and row bases of the march are scalar (r = 10: 205 -> 190 VALU
    // instructions per marched row, 170 -> 154 VGPRs)
    const int item = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave);
    if (item >= nitems) return;
    const int sx = item % strips_x;
    const int sy = item / strips_x;
    const int y0 = sy * seg;
    // Grids that stay in the Infinity Cache run fewer waves per CU (see
    // strip_shape), so radii up to 10 that otherwise prefetch one row ahead
    // prefetch three there (measured: r = 5, 6 and 8 gain 4-25 % on 4096^2).
    constexpr int DQ = (NT || StripConsts<R, C>::D > 1 || R > 10) ? 0 : 3;
    strip_march<Pol, C, MODE, false, DQ, NT, strip_tail_mode<R, C>()>(out, in, rows, cols, nv, sx * StripConsts<R, C>::SW, y0,
                                             min(y0 + seg, rows), pol, lds[wave], nullptr, sy & 1);
}

// Two generations of the radius-R mean blur per wave item (strip2_march_dir):
// output strips of 256 - 2RP columns, segments marched in alternating
// directions as in strip_stencil_kernel.  Aligned or unaligned rows.
template <int R, int MODE, bool NT>
__global__ __launch_bounds__(256) void strip2_mean_kernel(float *__restrict__ out, const float *__restrict__ in,
                                                          int rows, int cols, float nv, int seg, int strips_x,
                                                          int nitems, MeanPolicy<R> pol) {
    typedef StripConsts<R, 4> K;
    __shared__ __attribute__((aligned(16))) float lds[4][4][K::LDSW];
    const int wave = threadIdx.x >> 6;
    // wave-uniform item: row indices and row bases of the march are scalar
    const int item = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave);
    if (item >= nitems) return;
    const int sx = item % strips_x;
    const int sy = item / strips_x;
    const int y0 = sy * seg;
    const int xs = sx * (K::SW - 2 * K::RP);
    if (sy & 1)
        strip2_march_dir<R, MODE, true, NT>(out, in, rows, cols, nv, xs, y0, min(y0 + seg, rows), pol, lds[wave]);
    else
        strip2_march_dir<R, MODE, false, NT>(out, in, rows, cols, nv, xs, y0, min(y0 + seg, rows), pol, lds[wave]);
}

// Row access mode of a 4-column-per-lane march over these buffers.
template <typename T>
inline int row_mode(const T *in, const T *out, int cols) {
    const uintptr_t a = (uintptr_t)in | (uintptr_t)out;
    if ((cols & 3) == 0 && (a & 15u) == 0) return kAligned;
    if (cols >= 4 && (a & 3u) == 0) return kUnaligned;
    return kScalar;
}

struct StripShape {
    int seg, strips_x, segs, nitems, c;
};

// Columns per lane and segment length (measured on 512^2 .. 8192^2 grids,
// r = 1, 5, 10).  A grid whose 256-column strips cannot give 8 waves per CU
// at 8-row segments marches 64-column strips (one column per lane) instead.
// Segments aim at 8 waves per CU with 256-column strips on grids that stay
// in the Infinity Cache, 16 on grids streamed from HBM (longer load latency
// to cover), and 32 with 64-column strips, which carry a quarter of the
// loads per row each.  They stay at least 4, 8 or 16 rows long for R <= 2,
// <= 6 and larger, so that the 2R halo rows read by each segment stay a
// modest fraction of its rows.
inline StripShape strip_shape(int rows, int cols, int r, int seg, bool streamed) {
    StripShape s;
    const long cus = ctx().num_cus;
    const long strips4 = (cols + kStripW - 1) / kStripW;
    s.c = tune_int("MSL_STRIP_C", 0);
    if (s.c != 1 && s.c != 4) s.c = strips4 * ((rows + 7) / 8) < 8 * cus ? 1 : 4;
    const int sw = 64 * s.c;
    s.strips_x = (cols + sw - 1) / sw;
    if (seg <= 0) seg = tune_int("MSL_STRIP_SEG", 0);
    if (seg <= 0) {
        // radii 9 and 10 run at most two waves per SIMD with 256-column
        // strips: one round of them (8 per CU) halves the 2R halo rows of
        // each output row against two rounds (8192^2: r = 9 4.45 -> 4.68
        // TB/s, r = 10 with the checked last period 4.39 -> 4.69)
        const bool two_waves = s.c == 4 && r >= 9 && r <= 10;  // strip_max_waves_per_eu
        const long target = cus * (s.c == 1 ? 32 : (streamed && !two_waves ? 16 : 8));
        // at most target waves (whole rounds, no trailing partial one)
        const long segs = std::max(1L, target / s.strips_x);
        seg = (int)((rows + segs - 1) / segs);
        const int seg_min = r <= 2 ? 4 : (r <= 6 ? 8 : 16);
        if (seg < seg_min) seg = seg_min;
    }
    s.seg = seg;
    s.segs = (rows + seg - 1) / seg;
    s.nitems = s.strips_x * s.segs;
    return s;
}

template <typename Pol>
int launch_strip(typename Pol::T *out, const typename Pol::T *in, int rows, int cols, typename Pol::T nv, Pol pol,
                 int seg, const char *name) {
    const bool nt = stream_stores((size_t)rows * cols * sizeof(*out));
    StripShape s = strip_shape(rows, cols, Pol::R, seg, nt);
    const int blocks = (s.nitems + 3) / 4;
    const int mode = row_mode(in, out, cols);
    if (s.c == 1 && nt)
        strip_stencil_kernel<Pol, kAligned, true, 1><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg,
                                                                                         s.strips_x, s.nitems, pol);
    else if (s.c == 1)
        strip_stencil_kernel<Pol, kAligned, false, 1><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg,
                                                                                          s.strips_x, s.nitems, pol);
    else if (mode == kAligned && nt)
        strip_stencil_kernel<Pol, kAligned, true><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg,
                                                                                      s.strips_x, s.nitems, pol);
    else if (mode == kAligned)
        strip_stencil_kernel<Pol, kAligned, false><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg,
                                                                                       s.strips_x, s.nitems, pol);
    else if (mode == kUnaligned)
        strip_stencil_kernel<Pol, kUnaligned><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg,
                                                                                  s.strips_x, s.nitems, pol);
    else
        strip_stencil_kernel<Pol, kScalar><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, s.seg,
                                                                               s.strips_x, s.nitems, pol);
    return check_launch(name);
}

template <int R>
int launch_mean_r(float *out, const float *in, int rows, int cols, float nv, int seg) {
    return launch_strip(out, in, rows, cols, nv, MeanPolicy<R>::make(), seg, "msl_stencil2d_mean_f32");
}

typedef int (*MeanLauncher)(float *, const float *, int, int, float, int);
static const MeanLauncher kMeanLaunchers[] = {
    nullptr,           launch_mean_r<1>,  launch_mean_r<2>,  launch_mean_r<3>,  launch_mean_r<4>,
    launch_mean_r<5>,  launch_mean_r<6>,  launch_mean_r<7>,  launch_mean_r<8>,  launch_mean_r<9>,
    launch_mean_r<10>, launch_mean_r<11>, launch_mean_r<12>, launch_mean_r<13>, launch_mean_r<14>,
    launch_mean_r<15>, launch_mean_r<16>};
constexpr int kMaxMeanR = 16;

// Two generations per launch (strip2_mean_kernel) for radii 2 .. kMaxMean2R.
template <int R>
int launch_mean2_r(float *out, const float *in, int rows, int cols, float nv, int seg) {
    typedef StripConsts<R, 4> K;
    constexpr int sw = K::SW - 2 * K::RP;  // output columns per strip
    const bool nt = stream_stores((size_t)rows * cols * sizeof(float));
    const long cus = ctx().num_cus;
    const int strips_x = (cols + sw - 1) / sw;
    if (seg <= 0) seg = tune_int("MSL_STRIP2_SEG", 0);
    if (seg <= 0) {
        // waves per CU on streamed grids: 24 for r = 2, 8 above: one round
        // of long segments halves the 4R warm-up rows of each segment
        // (8192^2, 20 iterations, TB/s with 8 / 16 / 24 waves per CU:
        // r = 2 8.70 / 8.68 / 8.79, r = 3 8.05 / 7.44 / 7.83, r = 5
        // 7.05 / 6.06 / 5.86, r = 7 6.01 / 5.59 / 5.12);
        // MSL_STRIP2_WAVES overrides (tune_int caches one value per name, so
        // the radius-dependent default is applied here, not passed to it)
        int waves = tune_int("MSL_STRIP2_WAVES", 0);
        if (waves <= 0) waves = nt && R <= 2 ? 24 : 8;
        // whole rounds: at most target waves, so no partial round trails
        const long target = cus * waves;
        const long segs = std::max(1L, target / strips_x);
        seg = (int)((rows + segs - 1) / segs);
        if (seg < 8 * R) seg = 8 * R;
    }
    const long nitems = (long)strips_x * ((rows + seg - 1) / seg);
    MSL_REQUIRE(nitems <= INT32_MAX - 3, "msl_stencil2d_mean_iterate_f32: grid too large");
    const int blocks = (int)((nitems + 3) / 4);
    const MeanPolicy<R> pol = MeanPolicy<R>::make();
    const int mode = row_mode(in, out, cols);
    MSL_REQUIRE(mode != kScalar, "msl_stencil2d_mean_iterate_f32: two generations need rows of 4 or more cells");
    if (mode == kAligned && nt)
        strip2_mean_kernel<R, kAligned, true><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, seg,
                                                                               strips_x, (int)nitems, pol);
    else if (mode == kAligned)
        strip2_mean_kernel<R, kAligned, false><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, seg,
                                                                                strips_x, (int)nitems, pol);
    else
        strip2_mean_kernel<R, kUnaligned, false><<<blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, nv, seg,
                                                                                  strips_x, (int)nitems, pol);
    return check_launch("msl_stencil2d_mean_iterate_f32");
}

static const MeanLauncher kMean2Launchers[] = {nullptr, nullptr, launch_mean2_r<2>, launch_mean2_r<3>,
                                               launch_mean2_r<4>, launch_mean2_r<5>, launch_mean2_r<6>,
                                               launch_mean2_r<7>};
constexpr int kMaxMean2R = 7;

int mean_step(float *out, const float *in, int rows, int cols, int r, float nv, int seg) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(rows >= 0 && cols >= 0, "msl_stencil2d_mean_f32: negative shape");
    MSL_REQUIRE(r >= 1 && r <= kMaxMeanR, "msl_stencil2d_mean_f32: radius must be in [1, 16]");
    MSL_REQUIRE((int64_t)rows * cols <= INT32_MAX && cols < (1 << 30), "msl_stencil2d_mean_f32: grid too large");
    if (rows == 0 || cols == 0) return MSL_OK;
    MSL_REQUIRE(out && in && out != in, "msl_stencil2d_mean_f32: null or aliased buffers");
    return kMeanLaunchers[r](out, in, rows, cols, nv, seg);
}

int gol_step(int32_t *out, const int32_t *in, int rows, int cols) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(rows >= 0 && cols >= 0, "msl_gol_i32: negative shape");
    MSL_REQUIRE((int64_t)rows * cols <= INT32_MAX && cols < (1 << 30), "msl_gol_i32: grid too large");
    if (rows == 0 || cols == 0) return MSL_OK;
    MSL_REQUIRE(out && in && out != in, "msl_gol_i32: null or aliased buffers");
    LifePolicy pol;
    pol.denom = 1.0f;
    return launch_strip(out, in, rows, cols, (int32_t)0, pol, 0, "msl_gol_i32");
}

// ---- Game of Life, several generations per launch ---------------------------
// A block owns a 64-column x (4 * RPW)-row window of the grid (cells outside
// the grid read 0): lane l holds column l, wave w rows [w*RPW, (w+1)*RPW)
// in registers.  Each generation exchanges the waves' first and last rows
// through LDS, forms vertical 3-sums in registers and the horizontal ones by
// DPP lane shifts, and applies B3/S23; cells outside the grid stay 0, as in
// the one-step kernel.  The window edges lack neighbours, so after s
// generations only cells at least s from every window edge are exact; the
// block stores the central (64 - 2K) x (4*RPW - 2K) tile, exact after up to
// K generations.  Integer rule: bit-identical to s single steps.
// Whole-wave lane shifts with the vacated lane reading 0 (bound_ctrl): no
// 'old' operand to materialise, so the compiler folds the shift into the
// add that consumes it (v_add_f32_dpp / v_add_u32_dpp).
__device__ __forceinline__ int32_t lane_from_left(int32_t v) {  // lane l gets lane l-1 (lane 0: 0)
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int32_t lane_from_right(int32_t v) {  // lane l gets lane l+1 (lane 63: 0)
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true);
}

template <int K, int RPW>
__global__ __launch_bounds__(256) void gol_multi_kernel(int32_t *__restrict__ out, const int32_t *__restrict__ in,
                                                        int rows, int cols, int steps, int tiles_x) {
    constexpr int H = 4 * RPW, TW = 64 - 2 * K, TH = H - 2 * K;
    __shared__ int32_t edge[4][2][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int gx = tx * TW - K + lane;
    const int gy0 = ty * TH - K + wave * RPW;
    const bool colok = (unsigned)gx < (unsigned)cols;
    int32_t c[RPW + 2];  // rows gy0-1 .. gy0+RPW
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int gy = gy0 + i;
        c[i + 1] = (colok && (unsigned)gy < (unsigned)rows) ? in[(size_t)gy * cols + gx] : 0;
    }
    for (int s = 0; s < steps; s++) {
        edge[wave][0][lane] = c[1];
        edge[wave][1][lane] = c[RPW];
        __syncthreads();
        c[0] = wave > 0 ? edge[wave - 1][1][lane] : 0;
        c[RPW + 1] = wave < 3 ? edge[wave + 1][0][lane] : 0;
        __syncthreads();
        int32_t nw[RPW];
#pragma unroll
        for (int i = 0; i < RPW; i++) {
            const int32_t vs = c[i] + c[i + 1] + c[i + 2];
            const int32_t live = c[i + 1];
            const int32_t sum = lane_from_left(vs) + vs + lane_from_right(vs) - live;
            // B3/S23 as in examples/gameoflife.cpp:18-43
            int32_t next = 0;
            if (live == 1 && (sum == 2 || sum == 3)) next = 1;
            if (live == 0 && sum == 3) next = 1;
            const int gy = gy0 + i;
            nw[i] = (colok && (unsigned)gy < (unsigned)rows) ? next : 0;
        }
#pragma unroll
        for (int i = 0; i < RPW; i++) c[i + 1] = nw[i];
    }
    if (lane >= K && lane < 64 - K && colok) {
#pragma unroll
        for (int i = 0; i < RPW; i++) {
            const int wy = wave * RPW + i, gy = gy0 + i;
            if (wy >= K && wy < H - K && gy < rows) out[(size_t)gy * cols + gx] = c[i + 1];
        }
    }
}

constexpr int kGolSteps = 4;  // generations per launch
constexpr int kGolRows = 8;   // rows per wave: 64 x 32 windows, 56 x 24 tiles
// Grids of at most kGolSmallCells cells are launch-bound: there, 8
// generations per launch on 64 x 64 windows (48 x 48 tiles) run faster
// despite the wider overlap (1024^2: 2.16 -> 2.01 us per generation); from
// 2048^2 on the 4-generation windows are 7-9 % faster.
constexpr int kGolSmallSteps = 8;
constexpr int kGolSmallRows = 16;
constexpr long kGolSmallCells = 1L << 20;

// Mean blur of radius 1, several generations per launch, laid out like the
// Game of Life kernel above: lane = column, wave = RPW rows, row sums
// ((0 + left) + centre) + right with the side cells by DPP, column sums
// ((0 + above) + row) + below with the wave edge rows through LDS, then the
// exact division by 9 -- the operation order of the one-step kernel (the
// leading "0 +" of both sums folded into div_plus0), so every generation is
// bit-identical to it.  Cells outside the grid hold nv
// at every generation.
template <int K, int RPW>
__global__ __launch_bounds__(256) void mean3_multi_kernel(float *__restrict__ out, const float *__restrict__ in,
                                                          int rows, int cols, float nv, int steps, int tiles_x,
                                                          float denom, float rcp) {
    constexpr int H = 4 * RPW, TW = 64 - 2 * K, TH = H - 2 * K;
    __shared__ float edge[4][2][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int gx = tx * TW - K + lane;
    const int gy0 = ty * TH - K + wave * RPW;
    const bool colok = (unsigned)gx < (unsigned)cols;
    float c[RPW];
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int gy = gy0 + i;
        c[i] = (colok && (unsigned)gy < (unsigned)rows) ? in[(size_t)gy * cols + gx] : nv;
    }
    // row sum of row i of the window (no leading +0: the column sum below
    // is finished as if it started at +0, which makes a -0 row sum and a
    // +0 one give the same result)
    auto rowsum = [&](int i) {
        const float l = __int_as_float(lane_from_left(__float_as_int(c[i])));
        const float r = __int_as_float(lane_from_right(__float_as_int(c[i])));
        return __fadd_rn(__fadd_rn(l, c[i]), r);
    };
    for (int s = 0; s < steps; s++) {
        float rs[RPW + 2];  // row sums of rows gy0-1 .. gy0+RPW
        // the wave's edge rows first: their sums go to the neighbouring
        // waves; the inner rows' sums follow the exchange, in the block of
        // the column sums, where each lane shift folds into its add
        rs[1] = rowsum(0);
        rs[RPW] = rowsum(RPW - 1);
        edge[wave][0][lane] = rs[1];
        edge[wave][1][lane] = rs[RPW];
        __syncthreads();
        rs[0] = wave > 0 ? edge[wave - 1][1][lane] : 0.0f;
        rs[RPW + 1] = wave < 3 ? edge[wave + 1][0][lane] : 0.0f;
        __syncthreads();
#pragma unroll
        for (int i = 1; i < RPW - 1; i++) rs[i + 1] = rowsum(i);
#pragma unroll
        for (int i = 0; i < RPW; i++) {
            // the column sum without its leading +0, which div_plus0 restores
            const float t = __fadd_rn(__fadd_rn(rs[i], rs[i + 1]), rs[i + 2]);
            const int gy = gy0 + i;
            c[i] = (colok && (unsigned)gy < (unsigned)rows) ? div_plus0(t, denom, rcp) : nv;
        }
    }
    if (lane >= K && lane < 64 - K && colok) {
#pragma unroll
        for (int i = 0; i < RPW; i++) {
            const int wy = wave * RPW + i, gy = gy0 + i;
            if (wy >= K && wy < H - K && gy < rows) out[(size_t)gy * cols + gx] = c[i];
        }
    }
}

constexpr int kMean3Steps = 6;  // generations per launch of the radius-1 mean blur
constexpr int kMean3Rows = 16;  // rows per wave: 64 x 64 windows, 52 x 52 tiles

// Launch plan for `iters` generations of at most k per launch, alternating
// buffers: the result must end in b exactly when iters is odd (as with one
// launch per generation), so the number of launches must have the parity of
// iters; when it does not, one launch of several generations is split in
// two (there is one, since launches of one generation each number iters).
static std::vector<int> generation_plan(int iters, int k) {
    std::vector<int> plan;
    for (int done = 0; done < iters; done += plan.back()) plan.push_back(std::min(k, iters - done));
    if ((int)(plan.size() & 1) != (iters & 1)) {
        for (size_t j = plan.size(); j-- > 0;)
            if (plan[j] >= 2) {
                plan[j] -= 1;
                plan.insert(plan.begin() + j + 1, 1);
                break;
            }
    }
    return plan;
}

// Whether iterations of radius r run two generations per launch by default.
// Measured (GB/s, one vs two generations per launch): radii 2-4 gain from
// 4096^2 on (r = 2: 5.7 -> 9.0 TB/s), radii 5-7 only on grids streamed from
// HBM (8192^2: r = 5 4.7 -> 7.2, r = 7 4.5 -> 6.1); on grids that stay in
// the Infinity Cache their longer halo recomputation costs more than the
// saved traffic (5792^2: r = 5 5.2 -> 4.6).
static bool mean2_default(int r, int rows, int cols) {
    const size_t bytes = (size_t)rows * cols * sizeof(float);
    if (r < 2 || r > kMaxMean2R) return false;
    return r <= 4 ? bytes >= ((size_t)32 << 20) : stream_stores(bytes);
}

static int mean3_multi_step(float *out, const float *in, int rows, int cols, float nv, int steps) {
    constexpr int TW = 64 - 2 * kMean3Steps, TH = 4 * kMean3Rows - 2 * kMean3Steps;
    const int tiles_x = (cols + TW - 1) / TW;
    const long blocks = (long)tiles_x * ((rows + TH - 1) / TH);
    MSL_REQUIRE(blocks <= INT32_MAX, "msl_stencil2d_mean_iterate_f32: grid too large");
    const MeanPolicy<1> pol = MeanPolicy<1>::make();
    mean3_multi_kernel<kMean3Steps, kMean3Rows><<<(int)blocks, 256, 0, ctx().stream>>>(
        out, in, rows, cols, nv, steps, tiles_x, pol.denom, pol.rcp);
    return check_launch("msl_stencil2d_mean_iterate_f32");
}

template <int K, int RPW>
static int gol_multi_step(int32_t *out, const int32_t *in, int rows, int cols, int steps) {
    constexpr int TW = 64 - 2 * K, TH = 4 * RPW - 2 * K;
    const int tiles_x = (cols + TW - 1) / TW;
    const long blocks = (long)tiles_x * ((rows + TH - 1) / TH);
    MSL_REQUIRE(blocks <= INT32_MAX, "msl_gol_iterate_i32: grid too large");
    MSL_REQUIRE(steps >= 1 && steps <= K, "msl_gol_iterate_i32: generations per launch out of range");
    gol_multi_kernel<K, RPW><<<(int)blocks, 256, 0, ctx().stream>>>(out, in, rows, cols, steps, tiles_x);
    return check_launch("msl_gol_iterate_i32");
}

// Generations per launch of msl_stencil2d_mean_iterate_f32, in launch order
// (1 = the one-step kernel; otherwise the six-generation register-window
// kernel at r = 1 or the two-generation march at r = 2 .. kMaxMean2R on rows
// of 4 or more cells).  MSL_MEAN_GENS = 1 or 2 forces one or two generations
// per launch for r >= 2 (read on every call, so a process can compare both).
static std::vector<int> mean_iterate_plan(const float *d_a, const float *d_b, int rows, int cols, int r, int iters) {
    if (r == 1) return generation_plan(iters, kMean3Steps);
    const char *ge = getenv("MSL_MEAN_GENS");
    const int gens_env = ge ? atoi(ge) : 0;
    const bool pairs = r <= kMaxMean2R && row_mode(d_a, d_b, cols) != kScalar &&
                       (gens_env == 2 || (gens_env != 1 && mean2_default(r, rows, cols)));
    return pairs ? generation_plan(iters, 2) : std::vector<int>(iters, 1);
}

static std::string plan_signature(const std::vector<int> &plan) {
    std::string sig = "|";
    for (int g : plan) sig += (char)('0' + g);
    return sig;
}

// ---- exhaustive check of div_odd_square against IEEE division -------------
__global__ void zero_u64_kernel(unsigned long long *p) { *p = 0ull; }

// Every 32-bit pattern x: compare div_odd_square(x) (or div6, the heat
// division, or div_plus0 against (0 + x) / d) with x / d bit for bit (NaN against NaN counts as equal);
// one atomic per block.
// kKind 0: div_odd_square, 1: div6, 2: div_plus0 and 3: div6_plus0, the last
// two against (0 + x) / d
template <int kKind>
__global__ __launch_bounds__(256) void div_check_kernel(float d, float rcp, unsigned long long *count) {
    __shared__ unsigned int part[256];
    unsigned int bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < (1ull << 32); i += stride) {
        const float x = __uint_as_float((unsigned int)i);
        const float ref = __fdiv_rn(kKind >= 2 ? __fadd_rn(0.0f, x) : x, d);
        float got;
        if constexpr (kKind == 1) {
            got = div6(x);
        } else if constexpr (kKind == 2) {
            got = div_plus0(x, d, rcp);
        } else if constexpr (kKind == 3) {
            got = div6_plus0(x);
        } else {
            got = div_odd_square(x, d, rcp);
        }
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
            // a cached graph may still be running: drain the stream first
            MSL_HIP_TRY(hipStreamSynchronize(ctx().stream));
            for (auto &kv : g_cache.graphs) (void)hipGraphExecDestroy(kv.second);
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
    MSL_REQUIRE((int64_t)rows * cols <= INT32_MAX && cols < (1 << 30), "msl_stencil2d_mean_iterate_f32: grid too large");
    MSL_REQUIRE(d_a && d_b && d_a != d_b, "msl_stencil2d_mean_iterate_f32: null or aliased buffers");
    const std::vector<int> plan = mean_iterate_plan(d_a, d_b, rows, cols, r, iters);
    std::string key = graph_key("mean2d", d_a, d_b, rows, cols, r, nvb.i, iters);
    key += plan_signature(plan);
    return run_cached_graph(key, [&]() {
        for (size_t k = 0; k < plan.size(); k++) {
            float *src = (k & 1) ? d_b : d_a;
            float *dst = (k & 1) ? d_a : d_b;
            int brc = plan[k] == 1   ? mean_step(dst, src, rows, cols, r, nv, 0)
                      : r == 1       ? mean3_multi_step(dst, src, rows, cols, nv, plan[k])
                                     : kMean2Launchers[r](dst, src, rows, cols, nv, 0);
            if (brc) return brc;
        }
        return MSL_OK;
    });
}

int msl_stencil2d_mean_iterate_plan(const float *d_a, const float *d_b, int rows, int cols, int r, int iters,
                                    int *h_gens, int capacity, int *h_launches) {
    MSL_REQUIRE(iters >= 0 && rows >= 0 && cols >= 0, "msl_stencil2d_mean_iterate_plan: negative argument");
    MSL_REQUIRE(r >= 1 && r <= kMaxMeanR, "msl_stencil2d_mean_iterate_plan: radius must be in [1, 16]");
    MSL_REQUIRE(h_launches != nullptr && capacity >= 0, "msl_stencil2d_mean_iterate_plan: bad output arguments");
    std::vector<int> plan;
    if (iters > 0 && rows > 0 && cols > 0) plan = mean_iterate_plan(d_a, d_b, rows, cols, r, iters);
    *h_launches = (int)plan.size();
    MSL_REQUIRE(h_gens != nullptr || plan.empty() || capacity == 0, "msl_stencil2d_mean_iterate_plan: null output");
    for (size_t k = 0; k < plan.size() && (int)k < capacity; k++) h_gens[k] = plan[k];
    return MSL_OK;
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
    div_check_kernel<0><<<ctx().num_cus * 8, 256, 0, ctx().stream>>>(d, 1.0f / d, d_count);
    div_check_kernel<2><<<ctx().num_cus * 8, 256, 0, ctx().stream>>>(d, 1.0f / d, d_count);
    rc = check_launch("msl_selftest_div_odd_square");
    unsigned long long h = 0;
    hipError_t e = hipMemcpyAsync(&h, d_count, sizeof h, hipMemcpyDeviceToHost, ctx().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx().stream);
    (void)hipFree(d_count);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "msl_selftest_div_odd_square");
    *h_mismatches = h;
    return MSL_OK;
}

int msl_selftest_div_six(uint64_t *h_mismatches) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(h_mismatches != nullptr, "msl_selftest_div_six: null out-pointer");
    unsigned long long *d_count = nullptr;
    MSL_HIP_TRY(hipMalloc(&d_count, sizeof(unsigned long long)));
    zero_u64_kernel<<<1, 1, 0, ctx().stream>>>(d_count);
    div_check_kernel<1><<<ctx().num_cus * 8, 256, 0, ctx().stream>>>(6.0f, 1.0f / 6.0f, d_count);
    div_check_kernel<3><<<ctx().num_cus * 8, 256, 0, ctx().stream>>>(6.0f, 1.0f / 6.0f, d_count);
    rc = check_launch("msl_selftest_div_six");
    unsigned long long h = 0;
    hipError_t e = hipMemcpyAsync(&h, d_count, sizeof h, hipMemcpyDeviceToHost, ctx().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx().stream);
    (void)hipFree(d_count);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "msl_selftest_div_six");
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
    MSL_REQUIRE((int64_t)rows * cols <= INT32_MAX, "msl_gol_iterate_i32: grid too large");
    MSL_REQUIRE(d_a && d_b && d_a != d_b, "msl_gol_iterate_i32: null or aliased buffers");
    const bool small = (long)rows * cols <= kGolSmallCells;
    const std::vector<int> plan = generation_plan(iters, small ? kGolSmallSteps : kGolSteps);
    std::string key = graph_key("gol", d_a, d_b, rows, cols, iters, 0, 0);
    return run_cached_graph(key, [&]() {
        for (size_t k = 0; k < plan.size(); k++) {
            int32_t *src = (k & 1) ? d_b : d_a;
            int32_t *dst = (k & 1) ? d_a : d_b;
            int brc = small ? gol_multi_step<kGolSmallSteps, kGolSmallRows>(dst, src, rows, cols, plan[k])
                            : gol_multi_step<kGolSteps, kGolRows>(dst, src, rows, cols, plan[k]);
            if (brc) return brc;
        }
        return MSL_OK;
    });
}

}  // extern "C"

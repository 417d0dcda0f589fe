// Batched mean blur: thousands of independent grids of different shapes and
// radii stencilled together, each step one persistent launch over a device
// work queue, the iteration loop and the per-grid fold captured in one HIP
// graph.
//
// Reference counterpart: none as one operation.  A Muesli program runs
// DM::mapStencil (src/dm.cpp:1535-1591) and DS::fold (src/ds.cpp:739-912)
// grid by grid, each call a launch plus synchronisations; the plain loop
// kept here for comparison (msl_batch_run_loop) does one launch per grid per
// step and one fold per grid.
//
// Work items.  A grid is cut into wave strips of 64*C columns (C = 4 or 1
// columns per lane) and segments of `seg` rows; one item = one (grid, strip,
// segment), processed by one wavefront with the same strip march as the
// single-grid kernel (so every output is bit-identical to it).  Items are
// ordered by (radius, row mode, grid) so that waves that run at the same
// time execute the same specialisation (the instruction cache is shared by
// two CUs).  The persistent launch holds `blocks_per_cu` 4-wave workgroups
// per CU; wave w takes items w, w+W, ... (static) or, with the dynamic
// queue, the items are split into one cost-balanced contiguous range per
// XCD and the XCD's waves take them from that range's atomic counter.
// Either way consecutive items run on one XCD, so the halo rows shared by
// neighbouring segments are served by one L2.
//
// Two generations per item.  In msl_batch_run, grids of radius up to
// kBatchPairMaxR with 4-column strips run two generations per item
// (strip2_march_dir: the generation-1 rows of a strip and its halo are kept
// in the wave's LDS and the second generation is formed from them, so each
// pair of generations reads and writes the grid once), in the launch plan
// of the single-grid iterations (pairs, with one pair split so that the
// number of launches keeps the parity of the step count); the other grids
// run one generation per launch.  Each launch round is a pair kernel over
// the pair items of that round and a step kernel over its single items.
// Every cell keeps the operation order of one generation per launch, so the
// states are bit-identical to the plain loop.
//
// Fold.  On the last step every lane keeps a TwoSum-compensated sum of the
// values it stores; a wave combines its lanes in a fixed butterfly and
// writes one (s, c) pair per item; a final kernel adds a grid's pairs in
// FP64 in item order.  Deterministic for a given tiling.
//
// Autotuner.  The launch shape (workgroups per CU, static or atomic queue)
// and the tile shape (C, seg) per width class and then per radius are chosen
// by coordinate descent, each candidate timed on whole-batch steps.  The
// table can be saved to and loaded from a text file keyed by the device.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "blur_item.hpp"

namespace mslhip {

int mean_step(float *out, const float *in, int rows, int cols, int r, float nv, int seg);

struct BGrid {
    float *a;
    float *b;
    int rows, cols, r, kind;
    float nv;
    int item0, nitems, pad;
};

struct BItem {
    int grid2;  // 2 * grid + march direction (1 = upwards)
    int xs, y0, y1;
};

constexpr int kXcd = 8;  // workgroups are dealt round-robin to the 8 XCDs

// Static mode (ctr == nullptr): logical wave w (XCD-major, xcd_tile) takes
// items w, w+W, w+2W, ...  Dynamic mode: the items are split into kXcd
// contiguous, cost-balanced ranges [bounds[x], bounds[x+1]); the waves of
// XCD x (blockIdx % kXcd == x) take their first item statically and every
// further one from that XCD's counter ctr[x], so neighbouring items stay in
// one L2 while the load is balanced within the XCD (requires gridDim.x to
// be a multiple of kXcd).  With ctr but no bounds: one global counter.
template <bool kFold>
__global__ __launch_bounds__(256) void batch_step_kernel(const BGrid *__restrict__ grids,
                                                         const BItem *__restrict__ items, int nitems, int parity,
                                                         float2 *__restrict__ partials, unsigned *__restrict__ ctr,
                                                         const int *__restrict__ bounds) {
    __shared__ __attribute__((aligned(16))) float lds[4][2 * kBatchLdsw];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    int i, end, stride;
    if (ctr && !bounds) {  // one global queue
        end = nitems;
        stride = gridDim.x * 4;
        i = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave);
    } else if (ctr) {
        const int x = blockIdx.x % kXcd;
        const int waves_x = (gridDim.x / kXcd) * 4;
        const int s0 = bounds[x];
        end = bounds[x + 1];
        stride = s0 + waves_x;
        i = __builtin_amdgcn_readfirstlane(s0 + (int)(blockIdx.x / kXcd) * 4 + wave);
        ctr += x;
    } else {
        end = nitems;
        stride = gridDim.x * 4;
        i = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave);
    }
    while (i < end) {
        const BItem it = items[i];
        const BGrid g = grids[it.grid2 >> 1];
        LaneSum ls;
        const float *src = parity ? g.b : g.a;
        float *dst = parity ? g.a : g.b;
        blur_item_r<kFold>(g.r, src, dst, g.rows, g.cols, g.nv, g.kind, it.xs, it.y0, it.y1, lds[wave], &ls,
                           it.grid2 & 1);
        if constexpr (kFold) {
            wave_pair_reduce(ls);
            if (lane == 0) partials[i] = make_float2(ls.s, ls.c);
        }
        if (ctr) {
            unsigned nx = 0;
            if (lane == 0) nx = atomicAdd(ctr, 1u);
            i = stride + __builtin_amdgcn_readfirstlane((int)nx);
        } else {
            i += stride;
        }
    }
}

// ---- launch rounds of msl_batch_run ----------------------------------------
// Radii up to 8 run two generations per item (the kernel needs 219 VGPRs,
// two waves per SIMD, as for r <= 7; r <= 10 needs 256 and measured no
// faster on the 4096-grid batch).  Items of 128 rows: the 4R extra input
// rows of a two-generation march cost 44% at 64 rows and r = 7 (10 mean
// blur steps of the 4096-grid batch: 5.69 ms at 64 rows, 5.50 at 96, 5.56
// at 128 with r <= 7; 5.43 at 128 with r <= 8; 5.95 one generation per item).
#ifndef MSL_BATCH_PAIR_MAX_R
#define MSL_BATCH_PAIR_MAX_R 8
#endif
constexpr int kBatchPairMaxR = MSL_BATCH_PAIR_MAX_R;
constexpr int kPairSeg = 128;  // rows per two-generation item
constexpr int kPairLdsw = StripConsts<kBatchPairMaxR, 4>::LDSW;

// One item of a launch round: grid g, strip at column xs, rows [y0, y1).
// flags: kRUp march upwards, kRSrcB read state b (else a), kRFold fold the
// stored values into partials[part].
struct RItem {
    int g, xs, y0, y1, flags, part;
};
constexpr int kRUp = 1, kRSrcB = 2, kRFold = 4;

template <int R, bool kFold>
__device__ __forceinline__ void pair_item(const float *src, float *dst, const BGrid &g, int xs, int y0, int y1,
                                          float *ldsw, LaneSum *ls, int up) {
    const MeanPolicy<R> pol = MeanPolicy<R>::make();
    typedef float Row[StripConsts<R, 4>::LDSW];
    Row *lds = reinterpret_cast<Row *>(ldsw);
    if (g.kind == kKindAligned4) {
        if (up)
            strip2_march_dir<R, kAligned, true, true, kFold>(dst, src, g.rows, g.cols, g.nv, xs, y0, y1, pol, lds, ls);
        else
            strip2_march_dir<R, kAligned, false, true, kFold>(dst, src, g.rows, g.cols, g.nv, xs, y0, y1, pol, lds, ls);
    } else {
        if (up)
            strip2_march_dir<R, kUnaligned, true, true, kFold>(dst, src, g.rows, g.cols, g.nv, xs, y0, y1, pol, lds, ls);
        else
            strip2_march_dir<R, kUnaligned, false, true, kFold>(dst, src, g.rows, g.cols, g.nv, xs, y0, y1, pol, lds,
                                                                ls);
    }
}

template <bool kFold, int R = 1>
__device__ __forceinline__ void pair_item_r(int r, const float *src, float *dst, const BGrid &g, int xs, int y0, int y1,
                                            float *ldsw, LaneSum *ls, int up) {
    if constexpr (R < kBatchPairMaxR) {
        if (r != R) return pair_item_r<kFold, R + 1>(r, src, dst, g, xs, y0, y1, ldsw, ls, up);
    }
    pair_item<R, kFold>(src, dst, g, xs, y0, y1, ldsw, ls, up);
}

// One launch of a round: every wave takes its first item statically and the
// next ones from the round's atomic counter; kPair: two generations per item.
template <bool kPair>
__global__ __launch_bounds__(256) void batch_round_kernel(const BGrid *__restrict__ grids,
                                                          const RItem *__restrict__ items, int nitems,
                                                          float2 *__restrict__ partials, unsigned *__restrict__ ctr) {
    __shared__ __attribute__((aligned(16))) float lds[4][kPair ? 4 * kPairLdsw : 2 * kBatchLdsw];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int stride = gridDim.x * 4;
    for (int i = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave); i < nitems;) {
        const RItem it = items[i];
        const BGrid g = grids[it.g];
        const float *src = (it.flags & kRSrcB) ? g.b : g.a;
        float *dst = (it.flags & kRSrcB) ? g.a : g.b;
        const int up = it.flags & kRUp;
        LaneSum ls;
        if (it.flags & kRFold) {
            if constexpr (kPair)
                pair_item_r<true>(g.r, src, dst, g, it.xs, it.y0, it.y1, lds[wave], &ls, up);
            else
                blur_item_r<true>(g.r, src, dst, g.rows, g.cols, g.nv, g.kind, it.xs, it.y0, it.y1, lds[wave], &ls, up);
            wave_pair_reduce(ls);
            if (lane == 0) partials[it.part] = make_float2(ls.s, ls.c);
        } else {
            if constexpr (kPair)
                pair_item_r<false>(g.r, src, dst, g, it.xs, it.y0, it.y1, lds[wave], &ls, up);
            else
                blur_item_r<false>(g.r, src, dst, g.rows, g.cols, g.nv, g.kind, it.xs, it.y0, it.y1, lds[wave], &ls,
                                   up);
        }
        unsigned nx = 0;
        if (lane == 0) nx = atomicAdd(ctr, 1u);
        i = stride + __builtin_amdgcn_readfirstlane((int)nx);
    }
}

// One wave per grid: FP64 sum of the (s, c) pairs of its last round's items
// (fold[g] = {first partial, count}), in item order.
__global__ __launch_bounds__(256) void batch_round_final_kernel(const int2 *__restrict__ fold, int ngrids,
                                                                const float2 *__restrict__ partials,
                                                                double *__restrict__ sums) {
    const int lane = threadIdx.x & 63;
    const int gi = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gi >= ngrids) return;
    const int2 f = fold[gi];
    double acc = 0.0;
    for (int k = lane; k < f.y; k += 64) {
        const float2 p = partials[f.x + k];
        acc += (double)p.x;
        acc += (double)p.y;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
    if (lane == 0) sums[gi] = acc;
}

// One wave per grid: sum = sum over the grid's items, in item order, of
// (double)s + (double)c; lanes take items l, l+64, ... and are combined in a
// fixed butterfly.
__global__ __launch_bounds__(256) void batch_fold_final_kernel(const BGrid *__restrict__ grids, int ngrids,
                                                               const float2 *__restrict__ partials,
                                                               double *__restrict__ sums) {
    const int lane = threadIdx.x & 63;
    const int gi = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gi >= ngrids) return;
    const BGrid g = grids[gi];
    double acc = 0.0;
    for (int k = lane; k < g.nitems; k += 64) {
        const float2 p = partials[g.item0 + k];
        acc += (double)p.x;
        acc += (double)p.y;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
    if (lane == 0) sums[gi] = acc;
}

__global__ void batch_zero_kernel(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void batch_zero_f64_kernel(double *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---- host side --------------------------------------------------------------
struct Tile {
    int c = 4;     // columns per lane (4 or 1)
    int seg = 64;  // rows per item
};

// A captured msl_batch_run graph and the device tables it reads.
struct RunGraph {
    hipGraphExec_t exec = nullptr;
    RItem *d_items = nullptr;
    int2 *d_fold = nullptr;
    float2 *d_partials = nullptr;
    unsigned *d_ctr = nullptr;
};

struct Batch {
    std::vector<msl_batch_grid> grids;
    Tile table[kBatchMaxR + 1][kWidthClasses];
    int blocks_per_cu = 2;
    int dynamic = 0;
    int max_blocks_per_cu = 4;
    // device state
    BGrid *d_grids = nullptr;
    BItem *d_items = nullptr;
    float2 *d_partials = nullptr;
    unsigned *d_ctr = nullptr;  // kXcd counters per step
    int ctr_cap = 0;
    int *d_bounds = nullptr;    // kXcd + 1 item-range bounds
    int nitems = 0, items_cap = 0;
    bool pairs = true;           // two generations per item in msl_batch_run
    int max_pair_blocks_per_cu = 2;
    int pair_seg = kPairSeg;
    std::map<std::pair<int, const void *>, struct RunGraph> graphs;
};

static void drop_graphs(Batch &b) {
    if (b.graphs.empty()) return;
    (void)hipStreamSynchronize(ctx().stream);
    for (auto &kv : b.graphs) {
        RunGraph &r = kv.second;
        if (r.exec) (void)hipGraphExecDestroy(r.exec);
        if (r.d_items) (void)hipFree(r.d_items);
        if (r.d_fold) (void)hipFree(r.d_fold);
        if (r.d_partials) (void)hipFree(r.d_partials);
        if (r.d_ctr) (void)hipFree(r.d_ctr);
    }
    b.graphs.clear();
}

static int grid_kind(const msl_batch_grid &g, int c) {
    if (c == 1 || g.cols < 4 || g.r > kBatchMaxR4) return kKindNarrow1;
    const uintptr_t a = (uintptr_t)g.a | (uintptr_t)g.b;
    if ((g.cols & 3) == 0 && (a & 15u) == 0) return kKindAligned4;
    return kKindUnaligned4;
}

// Items and device descriptors of the grids selected by `pick` (all when
// empty), with the given tiles.  Descriptors of unselected grids get no
// items.
static void plan_items(const Batch &b, const std::vector<int> &pick, const Tile (*table)[kWidthClasses],
                       std::vector<BGrid> &hg, std::vector<BItem> &hi) {
    const int n = (int)b.grids.size();
    hg.assign(n, BGrid{});
    hi.clear();
    std::vector<int> order;
    if (pick.empty()) {
        order.resize(n);
        for (int i = 0; i < n; i++) order[i] = i;
    } else {
        order = pick;
    }
    std::vector<int> kind(n);
    for (int g : order) {
        const msl_batch_grid &G = b.grids[g];
        kind[g] = grid_kind(G, table[G.r][width_class(G.cols)].c);
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (b.grids[x].r != b.grids[y].r) return b.grids[x].r < b.grids[y].r;
        return kind[x] < kind[y];
    });
    for (int g = 0; g < n; g++) {
        const msl_batch_grid &G = b.grids[g];
        hg[g] = BGrid{G.a, G.b, G.rows, G.cols, G.r, 0, G.nv, 0, 0, 0};
    }
    for (int g : order) {
        const msl_batch_grid &G = b.grids[g];
        const Tile t = table[G.r][width_class(G.cols)];
        const int k = kind[g];
        const int sw = 64 * (k == kKindNarrow1 ? 1 : 4);
        const int seg = std::max(1, t.seg);
        BGrid &d = hg[g];
        d.kind = k;
        d.item0 = (int)hi.size();
        if (G.rows > 0 && G.cols > 0) {
            for (int xs = 0; xs < G.cols; xs += sw)
                for (int y0 = 0; y0 < G.rows; y0 += seg)
                    hi.push_back(BItem{2 * g + ((y0 / seg) & 1), xs, y0, std::min(y0 + seg, G.rows)});
        }
        d.nitems = (int)hi.size() - d.item0;
    }
}

static int upload(Batch &b, const std::vector<BGrid> &hg, const std::vector<BItem> &hi) {
    drop_graphs(b);
    const int ni = (int)hi.size();
    if (ni > b.items_cap) {
        if (b.d_items) MSL_HIP_TRY(hipFree(b.d_items));
        if (b.d_partials) MSL_HIP_TRY(hipFree(b.d_partials));
        b.d_items = nullptr;
        b.d_partials = nullptr;
        const int cap = std::max(ni, 1);
        MSL_HIP_TRY(hipMalloc(&b.d_items, sizeof(BItem) * cap));
        MSL_HIP_TRY(hipMalloc(&b.d_partials, sizeof(float2) * cap));
        b.items_cap = cap;
    }
    if (!hg.empty())
        MSL_HIP_TRY(hipMemcpyAsync(b.d_grids, hg.data(), sizeof(BGrid) * hg.size(), hipMemcpyHostToDevice,
                                   ctx().stream));
    if (ni) MSL_HIP_TRY(hipMemcpyAsync(b.d_items, hi.data(), sizeof(BItem) * ni, hipMemcpyHostToDevice, ctx().stream));
    // cost-balanced contiguous item ranges, one per XCD: cost = cells read
    // by the item (its rows plus the 2r halo rows, times its columns)
    std::vector<double> cum(ni + 1, 0.0);
    for (int i = 0; i < ni; i++) {
        const BItem &it = hi[i];
        const BGrid &g = hg[it.grid2 >> 1];
        const int sw = g.kind == kKindNarrow1 ? 64 : 256;
        cum[i + 1] = cum[i] + (double)(it.y1 - it.y0 + 2 * g.r) * (double)std::min(sw, g.cols - it.xs);
    }
    int bounds[kXcd + 1];
    bounds[0] = 0;
    for (int x = 1; x < kXcd; x++) {
        const double target = cum[ni] * x / kXcd;
        bounds[x] = (int)(std::lower_bound(cum.begin(), cum.end(), target) - cum.begin());
        bounds[x] = std::max(bounds[x - 1], std::min(bounds[x], ni));
    }
    bounds[kXcd] = ni;
    if (!b.d_bounds) MSL_HIP_TRY(hipMalloc(&b.d_bounds, sizeof(int) * (kXcd + 1)));
    MSL_HIP_TRY(hipMemcpyAsync(b.d_bounds, bounds, sizeof(bounds), hipMemcpyHostToDevice, ctx().stream));
    MSL_HIP_TRY(hipStreamSynchronize(ctx().stream));
    b.nitems = ni;
    return MSL_OK;
}

static int rebuild(Batch &b) {
    std::vector<BGrid> hg;
    std::vector<BItem> hi;
    plan_items(b, {}, b.table, hg, hi);
    return upload(b, hg, hi);
}

// Queue modes: 0 static, 1 one atomic queue per XCD, 2 one global queue.
static const int *queue_bounds(const Batch &b) { return b.dynamic == 1 ? b.d_bounds : nullptr; }

static int launch_blocks(const Batch &b) {
    const int want = (b.nitems + 3) / 4;
    const int cap = ctx().num_cus * std::max(1, std::min(b.blocks_per_cu, b.max_blocks_per_cu));
    int blocks = std::max(1, std::min(want, cap));
    if (b.dynamic == 1) blocks = (blocks + kXcd - 1) / kXcd * kXcd;  // per-XCD queues
    return blocks;
}

static int ensure_ctr(Batch &b, int n) {
    if (n <= b.ctr_cap) return MSL_OK;
    drop_graphs(b);
    if (b.d_ctr) MSL_HIP_TRY(hipFree(b.d_ctr));
    b.d_ctr = nullptr;
    MSL_HIP_TRY(hipMalloc(&b.d_ctr, sizeof(unsigned) * n));
    b.ctr_cap = n;
    return MSL_OK;
}

// Mean time of `reps` back-to-back steps (parity 0, no fold), in ms.
static int time_steps(Batch &b, int reps, float *ms) {
    int rc = ensure_ctr(b, reps * kXcd);
    if (rc) return rc;
    const int blocks = launch_blocks(b);
    hipStream_t s = ctx().stream;
    for (int w = 0; w < 2; w++) {  // warm-up: code, TLB, caches
        if (b.dynamic) batch_zero_kernel<<<1, 256, 0, s>>>(b.d_ctr, kXcd);
        batch_step_kernel<false><<<blocks, 256, 0, s>>>(b.d_grids, b.d_items, b.nitems, 0, b.d_partials,
                                                        b.dynamic ? b.d_ctr : nullptr, queue_bounds(b));
    }
    if (b.dynamic) batch_zero_kernel<<<(reps * kXcd + 255) / 256, 256, 0, s>>>(b.d_ctr, reps * kXcd);
    MSL_HIP_TRY(hipEventRecord(ctx().t0, s));
    for (int k = 0; k < reps; k++)
        batch_step_kernel<false><<<blocks, 256, 0, s>>>(b.d_grids, b.d_items, b.nitems, 0, b.d_partials,
                                                        b.dynamic ? b.d_ctr + k * kXcd : nullptr, queue_bounds(b));
    MSL_HIP_TRY(hipEventRecord(ctx().t1, s));
    rc = check_launch("msl_batch_autotune");
    if (rc) return rc;
    MSL_HIP_TRY(hipEventSynchronize(ctx().t1));
    float t = 0.0f;
    MSL_HIP_TRY(hipEventElapsedTime(&t, ctx().t0, ctx().t1));
    *ms = t / reps;
    return MSL_OK;
}

static std::string device_tag() {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, ctx().device) != hipSuccess) return "unknown";
    std::string n = p.gcnArchName;
    for (char &ch : n)
        if (ch == ' ') ch = '_';
    return n + "/" + std::to_string(p.multiProcessorCount);
}

static bool load_table(Batch &b, const char *path, int *entries) {
    std::ifstream f(path);
    if (!f) return false;
    std::string magic, tag;
    f >> magic >> tag;
    if (magic != "msl-batch-tiles-v1" || tag != device_tag()) return false;
    Tile t[kBatchMaxR + 1][kWidthClasses];
    int n = 0, bpc = b.blocks_per_cu, dyn = b.dynamic;
    std::string key;
    while (f >> key) {
        if (key == "tile") {
            int r, wc, c, seg;
            if (!(f >> r >> wc >> c >> seg)) return false;
            if (r < 1 || r > kBatchMaxR || wc < 0 || wc >= kWidthClasses || (c != 1 && c != 4) || seg < 1) return false;
            t[r][wc] = Tile{c, seg};
            n++;
        } else if (key == "launch") {
            if (!(f >> bpc >> dyn)) return false;
        } else {
            return false;
        }
    }
    std::memcpy(b.table, t, sizeof(t));
    b.blocks_per_cu = bpc;
    b.dynamic = (dyn >= 0 && dyn <= 2) ? dyn : 0;
    *entries = n;
    return true;
}

static void save_table(const Batch &b, const char *path) {
    std::ofstream f(path);
    if (!f) return;
    f << "msl-batch-tiles-v1 " << device_tag() << "\n";
    for (int r = 1; r <= kBatchMaxR; r++)
        for (int wc = 0; wc < kWidthClasses; wc++)
            f << "tile " << r << " " << wc << " " << b.table[r][wc].c << " " << b.table[r][wc].seg << "\n";
    f << "launch " << b.blocks_per_cu << " " << b.dynamic << "\n";
}

// Whole-batch objective: the smallest of three averages of `reps`
// back-to-back steps (the minimum is the least noisy estimate of what a
// configuration can do).
static int objective(Batch &b, float *ms) {
    int rc = rebuild(b);
    if (rc) return rc;
    float best = 1e30f;
    for (int k = 0; k < 3; k++) {
        float t = 0.0f;
        rc = time_steps(b, 3, &t);
        if (rc) return rc;
        best = std::min(best, t);
    }
    *ms = best;
    return MSL_OK;
}

// Coordinate descent on the whole batch, so every candidate is measured with
// the device full of the other classes' work (a class timed alone under-fills
// the device and favours short segments):
//   1. launch shape: workgroups per CU (the two largest that fit) x queue;
//   2. per width class, for all radii together: C in {1, 4} x seg in
//      {32, 64, 128};
//   3. per radius: seg in {32, 64, 128} with the chosen C.
// A candidate replaces the current setting only when it is at least 1 %
// faster.
static int autotune(Batch &b, int *entries) {
    bool present[kBatchMaxR + 1][kWidthClasses] = {};
    bool any_r[kBatchMaxR + 1] = {};
    for (const msl_batch_grid &G : b.grids)
        if (G.rows > 0 && G.cols > 0) {
            present[G.r][width_class(G.cols)] = true;
            any_r[G.r] = true;
        }
    float cur = 0.0f;
    int rc = objective(b, &cur);
    if (rc) return rc;
    auto try_change = [&](auto apply, auto undo) -> int {
        apply();
        float t = 0.0f;
        int trc = objective(b, &t);
        if (trc) return trc;
        if (t < 0.99f * cur) cur = t;
        else undo();
        return MSL_OK;
    };
    // 1. launch shape
    for (int bpc = std::max(1, b.max_blocks_per_cu - 1); bpc <= b.max_blocks_per_cu; bpc++) {
        for (int dyn = 0; dyn <= 2; dyn++) {
            const int obpc = b.blocks_per_cu, odyn = b.dynamic;
            if (bpc == obpc && dyn == odyn) continue;
            rc = try_change([&] { b.blocks_per_cu = bpc; b.dynamic = dyn; },
                            [&] { b.blocks_per_cu = obpc; b.dynamic = odyn; });
            if (rc) return rc;
        }
    }
    // 2. per width class
    static const Tile kClassCands[] = {{4, 32}, {4, 64}, {4, 128}, {1, 32}, {1, 64}, {1, 128}};
    for (int wc = 0; wc < kWidthClasses; wc++) {
        bool used = false;
        for (int r = 1; r <= kBatchMaxR; r++) used |= present[r][wc];
        if (!used) continue;
        for (const Tile &cand : kClassCands) {
            Tile saved[kBatchMaxR + 1];
            for (int r = 0; r <= kBatchMaxR; r++) saved[r] = b.table[r][wc];
            rc = try_change([&] { for (int r = 1; r <= kBatchMaxR; r++) b.table[r][wc] = cand; },
                            [&] { for (int r = 0; r <= kBatchMaxR; r++) b.table[r][wc] = saved[r]; });
            if (rc) return rc;
        }
    }
    // 3. per radius
    static const int kSegs[] = {32, 64, 128};
    for (int r = 1; r <= kBatchMaxR; r++) {
        if (!any_r[r]) continue;
        for (int seg : kSegs) {
            Tile saved[kWidthClasses];
            for (int wc = 0; wc < kWidthClasses; wc++) saved[wc] = b.table[r][wc];
            rc = try_change([&] { for (int wc = 0; wc < kWidthClasses; wc++) b.table[r][wc].seg = seg; },
                            [&] { for (int wc = 0; wc < kWidthClasses; wc++) b.table[r][wc] = saved[wc]; });
            if (rc) return rc;
        }
    }
    int tuned = 0;
    for (int r = 1; r <= kBatchMaxR; r++)
        for (int wc = 0; wc < kWidthClasses; wc++) tuned += present[r][wc] ? 1 : 0;
    *entries = tuned;
    return rebuild(b);
}

// Generations per launch of a grid that runs two per item: pairs, one split
// in two when the number of launches would not have the parity of iters (the
// result then ends in b exactly when iters is odd, as for the other grids).
static std::vector<int> pair_plan(int iters) {
    std::vector<int> plan;
    for (int done = 0; done < iters; done += plan.back()) plan.push_back(std::min(2, iters - done));
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

struct RoundRange {
    int pair_off, pair_n, single_off, single_n;
};

// The launch rounds of msl_batch_run: per round the pair items and the
// single items (tiles of the single-step table), grids by descending radius,
// each grid's items in its round j reading state b when j is odd; the items
// of a grid's last round fold into consecutive partials (fold[g] = {first,
// count}).
struct RoundPlan {
    std::vector<RItem> items;
    std::vector<int2> fold;  // per grid {first partial, count}
    std::vector<RoundRange> rounds;
    int nparts = 0;
    int pair_grids = 0;
};

static void plan_rounds(const Batch &b, int iters, bool fold, RoundPlan &rp) {
    const int ng = (int)b.grids.size();
    std::vector<std::vector<int>> plan(ng);
    std::vector<int> kind(ng, kKindNarrow1), order;
    for (int g = 0; g < ng; g++) {
        const msl_batch_grid &G = b.grids[g];
        if (G.rows == 0 || G.cols == 0) continue;
        kind[g] = grid_kind(G, b.table[G.r][width_class(G.cols)].c);
        const bool pair = b.pairs && G.r <= kBatchPairMaxR && kind[g] != kKindNarrow1;
        plan[g] = pair ? pair_plan(iters) : std::vector<int>(iters, 1);
        rp.pair_grids += pair ? 1 : 0;
        order.push_back(g);
    }
    // largest radius first: the costliest items start first and the cheap
    // ones fill the end of each round (10 steps of the 4096-grid batch:
    // 5.38 -> 5.04 ms against ascending radii); one radius and row mode at a
    // time, so concurrent waves run the same specialisation
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (b.grids[x].r != b.grids[y].r) return b.grids[x].r > b.grids[y].r;
        return kind[x] < kind[y];
    });
    size_t nrounds = 0;
    for (int g : order) nrounds = std::max(nrounds, plan[g].size());
    std::vector<RItem> &items = rp.items;
    std::vector<int2> &hfold = rp.fold;
    hfold.assign(std::max(ng, 1), make_int2(0, 0));
    int &nparts = rp.nparts;
    std::vector<RoundRange> &rounds = rp.rounds;
    for (size_t j = 0; j < nrounds; j++) {
        RoundRange rr{};
        for (int pass = 0; pass < 2; pass++) {  // pass 0: pair items, 1: single items
            const int off = (int)items.size();
            for (int g : order) {
                if (j >= plan[g].size() || (plan[g][j] == 2) != (pass == 0)) continue;
                const msl_batch_grid &G = b.grids[g];
                const bool last = fold && j + 1 == plan[g].size();
                int sw, seg;
                if (pass == 0) {
                    const int halo = (G.r + 3) / 4 * 4;
                    sw = 256 - 2 * halo;
                    seg = b.pair_seg;
                } else {
                    const Tile t = b.table[G.r][width_class(G.cols)];
                    sw = 64 * (kind[g] == kKindNarrow1 ? 1 : 4);
                    seg = std::max(1, t.seg);
                }
                if (last) hfold[g] = make_int2(nparts, 0);
                for (int xs = 0; xs < G.cols; xs += sw)
                    for (int y0 = 0; y0 < G.rows; y0 += seg) {
                        int flags = ((j & 1) ? kRSrcB : 0) | (((y0 / seg) & 1) ? kRUp : 0);
                        int part = 0;
                        if (last) {
                            flags |= kRFold;
                            part = nparts++;
                            hfold[g].y++;
                        }
                        items.push_back(RItem{g, xs, y0, std::min(y0 + seg, G.rows), flags, part});
                    }
            }
            if (pass == 0) {
                rr.pair_off = off;
                rr.pair_n = (int)items.size() - off;
            } else {
                rr.single_off = off;
                rr.single_n = (int)items.size() - off;
            }
        }
        rounds.push_back(rr);
    }
}

// Plans the rounds and uploads items and fold ranges into rg's buffers.
static int build_rounds(Batch &b, int iters, bool fold, RunGraph &rg, std::vector<RoundRange> &rounds) {
    RoundPlan rp;
    plan_rounds(b, iters, fold, rp);
    rounds = rp.rounds;
    const std::vector<RItem> &items = rp.items;
    const std::vector<int2> &hfold = rp.fold;
    const int nparts = rp.nparts;
    const size_t nrounds = rounds.size();
    MSL_HIP_TRY(hipMalloc(&rg.d_items, sizeof(RItem) * std::max<size_t>(items.size(), 1)));
    MSL_HIP_TRY(hipMalloc(&rg.d_fold, sizeof(int2) * hfold.size()));
    MSL_HIP_TRY(hipMalloc(&rg.d_partials, sizeof(float2) * std::max(nparts, 1)));
    MSL_HIP_TRY(hipMalloc(&rg.d_ctr, sizeof(unsigned) * std::max<size_t>(2 * nrounds, 1)));
    if (!items.empty())
        MSL_HIP_TRY(hipMemcpyAsync(rg.d_items, items.data(), sizeof(RItem) * items.size(), hipMemcpyHostToDevice,
                                   ctx().stream));
    MSL_HIP_TRY(hipMemcpyAsync(rg.d_fold, hfold.data(), sizeof(int2) * hfold.size(), hipMemcpyHostToDevice,
                               ctx().stream));
    MSL_HIP_TRY(hipStreamSynchronize(ctx().stream));
    return MSL_OK;
}

// Enqueues the rounds (counters cleared by a kernel, never a memset node)
// and the per-grid fold combine.
static int enqueue_rounds(Batch &b, const std::vector<RoundRange> &rounds, double *d_sums, const RunGraph &rg) {
    hipStream_t s = ctx().stream;
    const int nc = (int)(2 * rounds.size());
    if (nc > 0) {
        batch_zero_kernel<<<(nc + 255) / 256, 256, 0, s>>>(rg.d_ctr, nc);
        int rc = check_launch("msl_batch_run");
        if (rc) return rc;
    }
    const int cus = ctx().num_cus;
    for (size_t j = 0; j < rounds.size(); j++) {
        const RoundRange &rr = rounds[j];
        if (rr.pair_n) {
            const int blocks = std::max(1, std::min((rr.pair_n + 3) / 4, cus * b.max_pair_blocks_per_cu));
            batch_round_kernel<true><<<blocks, 256, 0, s>>>(b.d_grids, rg.d_items + rr.pair_off, rr.pair_n,
                                                            rg.d_partials, rg.d_ctr + 2 * j);
        }
        if (rr.single_n) {
            const int bpc = std::max(1, std::min(b.blocks_per_cu, b.max_blocks_per_cu));
            const int blocks = std::max(1, std::min((rr.single_n + 3) / 4, cus * bpc));
            batch_round_kernel<false><<<blocks, 256, 0, s>>>(b.d_grids, rg.d_items + rr.single_off, rr.single_n,
                                                             rg.d_partials, rg.d_ctr + 2 * j + 1);
        }
        int rc = check_launch("msl_batch_run");
        if (rc) return rc;
    }
    if (d_sums) {
        const int ng = (int)b.grids.size();
        batch_round_final_kernel<<<(ng + 3) / 4, 256, 0, s>>>(rg.d_fold, ng, rg.d_partials, d_sums);
        return check_launch("msl_batch_run");
    }
    return MSL_OK;
}

static int check_batch(void *h, const char *what) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (!h) return fail(MSL_E_INVALID, std::string(what) + ": null batch");
    if (ctx().capturing) return fail(MSL_E_STATE, std::string(what) + ": not allowed during graph capture");
    return MSL_OK;
}

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_batch_create(const msl_batch_grid *grids, int ngrids, void **h_batch) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(h_batch != nullptr, "msl_batch_create: null out-pointer");
    MSL_REQUIRE(ngrids >= 0 && (ngrids == 0 || grids != nullptr), "msl_batch_create: bad grid list");
    if (ctx().capturing) return fail(MSL_E_STATE, "msl_batch_create: not allowed during graph capture");
    for (int g = 0; g < ngrids; g++) {
        const msl_batch_grid &G = grids[g];
        MSL_REQUIRE(G.rows >= 0 && G.cols >= 0, "msl_batch_create: negative shape");
        MSL_REQUIRE((int64_t)G.rows * G.cols <= INT32_MAX && G.cols < (1 << 30), "msl_batch_create: grid too large");
        MSL_REQUIRE(G.r >= 1 && G.r <= kBatchMaxR, "msl_batch_create: radius must be in [1, 10]");
        if (G.rows > 0 && G.cols > 0) {
            MSL_REQUIRE(G.a && G.b && G.a != G.b, "msl_batch_create: null or aliased grid buffers");
            MSL_REQUIRE((((uintptr_t)G.a | (uintptr_t)G.b) & 3u) == 0, "msl_batch_create: buffers must be 4-byte aligned");
        }
    }
    Batch *b = new Batch();
    b->grids.assign(grids, grids + ngrids);
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_step_kernel<true>, 256, 0) == hipSuccess && occ > 0)
        b->max_blocks_per_cu = occ;
    b->blocks_per_cu = tune_int("MSL_BATCH_BLOCKS_PER_CU", 0);  // (cached per name: default applied below)
    if (b->blocks_per_cu <= 0) b->blocks_per_cu = b->max_blocks_per_cu;
    b->dynamic = tune_int("MSL_BATCH_DYNAMIC", 1);
    if (const char *e = getenv("MSL_BATCH_PAIRS")) b->pairs = atoi(e) != 0;
    if (const char *e = getenv("MSL_BATCH_PAIR_SEG")) b->pair_seg = std::max(8, atoi(e));
    int pocc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pocc, batch_round_kernel<true>, 256, 0) == hipSuccess && pocc > 0)
        b->max_pair_blocks_per_cu = pocc;
    for (int r = 0; r <= kBatchMaxR; r++)
        for (int wc = 0; wc < kWidthClasses; wc++) b->table[r][wc] = wc < 2 ? Tile{1, 64} : Tile{4, 64};
    if (ngrids > 0) {
        hipError_t e = hipMalloc(&b->d_grids, sizeof(BGrid) * ngrids);
        if (e != hipSuccess) {
            delete b;
            return hip_fail(e, "msl_batch_create");
        }
    }
    rc = rebuild(*b);
    if (rc) {
        msl_batch_destroy(b);
        return rc;
    }
    *h_batch = b;
    return MSL_OK;
}

int msl_batch_destroy(void *batch) {
    if (!batch) return MSL_OK;
    Batch *b = static_cast<Batch *>(batch);
    if (ctx().ready) {
        drop_graphs(*b);
        (void)hipStreamSynchronize(ctx().stream);
        if (b->d_grids) (void)hipFree(b->d_grids);
        if (b->d_items) (void)hipFree(b->d_items);
        if (b->d_partials) (void)hipFree(b->d_partials);
        if (b->d_ctr) (void)hipFree(b->d_ctr);
        if (b->d_bounds) (void)hipFree(b->d_bounds);
    }
    delete b;
    return MSL_OK;
}

int msl_batch_set_tile(void *batch, int r, int width_class_idx, int cols_per_lane, int seg) {
    int rc = check_batch(batch, "msl_batch_set_tile");
    if (rc) return rc;
    MSL_REQUIRE(r >= 1 && r <= kBatchMaxR, "msl_batch_set_tile: radius must be in [1, 10]");
    MSL_REQUIRE(width_class_idx >= 0 && width_class_idx < kWidthClasses, "msl_batch_set_tile: width class in [0, 3]");
    MSL_REQUIRE(cols_per_lane == 1 || cols_per_lane == 4, "msl_batch_set_tile: cols_per_lane must be 1 or 4");
    MSL_REQUIRE(seg >= 1, "msl_batch_set_tile: seg must be positive");
    Batch *b = static_cast<Batch *>(batch);
    b->table[r][width_class_idx] = Tile{cols_per_lane, seg};
    return rebuild(*b);
}

int msl_batch_set_launch(void *batch, int blocks_per_cu, int dynamic_queue) {
    int rc = check_batch(batch, "msl_batch_set_launch");
    if (rc) return rc;
    MSL_REQUIRE(blocks_per_cu >= 1, "msl_batch_set_launch: blocks_per_cu must be positive");
    Batch *b = static_cast<Batch *>(batch);
    drop_graphs(*b);
    b->blocks_per_cu = blocks_per_cu;
    MSL_REQUIRE(dynamic_queue >= 0 && dynamic_queue <= 2, "msl_batch_set_launch: queue mode must be 0, 1 or 2");
    b->dynamic = dynamic_queue;
    return MSL_OK;
}

int msl_batch_autotune(void *batch, const char *cache_path, int *h_entries) {
    int rc = check_batch(batch, "msl_batch_autotune");
    if (rc) return rc;
    Batch *b = static_cast<Batch *>(batch);
    int entries = 0;
    if (cache_path && load_table(*b, cache_path, &entries)) {
        if (h_entries) *h_entries = -entries;  // negative: loaded from the cache
        return rebuild(*b);
    }
    rc = autotune(*b, &entries);
    if (rc) {
        rebuild(*b);
        return rc;
    }
    if (cache_path) save_table(*b, cache_path);
    if (h_entries) *h_entries = entries;
    return MSL_OK;
}

int msl_batch_info(const void *batch, int *h_items, int *h_blocks, int *h_blocks_per_cu, int *h_dynamic) {
    MSL_REQUIRE(batch != nullptr, "msl_batch_info: null batch");
    const Batch *b = static_cast<const Batch *>(batch);
    if (h_items) *h_items = b->nitems;
    if (h_blocks) *h_blocks = launch_blocks(*b);
    if (h_blocks_per_cu) *h_blocks_per_cu = std::min(b->blocks_per_cu, b->max_blocks_per_cu);
    if (h_dynamic) *h_dynamic = b->dynamic;
    return MSL_OK;
}

int msl_batch_get_tile(const void *batch, int r, int width_class_idx, int *h_cols_per_lane, int *h_seg) {
    MSL_REQUIRE(batch != nullptr, "msl_batch_get_tile: null batch");
    MSL_REQUIRE(r >= 1 && r <= kBatchMaxR && width_class_idx >= 0 && width_class_idx < kWidthClasses,
                "msl_batch_get_tile: bad class");
    const Batch *b = static_cast<const Batch *>(batch);
    if (h_cols_per_lane) *h_cols_per_lane = b->table[r][width_class_idx].c;
    if (h_seg) *h_seg = b->table[r][width_class_idx].seg;
    return MSL_OK;
}

int msl_batch_run_plan(const void *batch, int iters, int with_fold, int *h_rounds, int *h_launches,
                       int *h_pair_grids, int *h_items) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(batch != nullptr, "msl_batch_run_plan: null batch");
    MSL_REQUIRE(iters >= 0, "msl_batch_run_plan: negative iteration count");
    const Batch *b = static_cast<const Batch *>(batch);
    RoundPlan rp;
    plan_rounds(*b, iters, with_fold != 0, rp);
    int launches = rp.rounds.empty() ? 0 : 1;  // counter clear
    for (const RoundRange &rr : rp.rounds) launches += (rr.pair_n > 0) + (rr.single_n > 0);
    if (with_fold && iters > 0) launches++;
    if (h_rounds) *h_rounds = (int)rp.rounds.size();
    if (h_launches) *h_launches = launches;
    if (h_pair_grids) *h_pair_grids = rp.pair_grids;
    if (h_items) *h_items = (int)rp.items.size();
    return MSL_OK;
}

int msl_batch_run(void *batch, int iters, double *d_sums, int *h_result_in_b) {
    int rc = check_batch(batch, "msl_batch_run");
    if (rc) return rc;
    MSL_REQUIRE(iters >= 0, "msl_batch_run: negative iteration count");
    MSL_REQUIRE(iters >= 1 || d_sums == nullptr, "msl_batch_run: the fold needs at least one step");
    Batch *b = static_cast<Batch *>(batch);
    if (h_result_in_b) *h_result_in_b = iters & 1;
    if (b->grids.empty()) return MSL_OK;
    const auto key = std::make_pair(iters, (const void *)d_sums);
    auto it = b->graphs.find(key);
    hipGraphExec_t exec = it == b->graphs.end() ? nullptr : it->second.exec;
    if (!exec) {
        RunGraph rg;
        std::vector<RoundRange> rounds;
        rc = build_rounds(*b, iters, d_sums != nullptr, rg, rounds);
        if (!rc) rc = msl_graph_begin();
        if (rc) {
            if (rg.d_items) (void)hipFree(rg.d_items);
            if (rg.d_fold) (void)hipFree(rg.d_fold);
            if (rg.d_partials) (void)hipFree(rg.d_partials);
            if (rg.d_ctr) (void)hipFree(rg.d_ctr);
            return rc;
        }
        int brc = enqueue_rounds(*b, rounds, d_sums, rg);
        void *g = nullptr;
        int erc = msl_graph_end(&g);
        if (brc || erc) {
            if (g) msl_graph_destroy(g);
            (void)hipFree(rg.d_items);
            (void)hipFree(rg.d_fold);
            (void)hipFree(rg.d_partials);
            (void)hipFree(rg.d_ctr);
            return brc ? brc : erc;
        }
        rg.exec = (hipGraphExec_t)g;
        exec = rg.exec;
        b->graphs[key] = rg;
    }
    MSL_HIP_TRY(hipGraphLaunch(exec, ctx().stream));
    return MSL_OK;
}

int msl_batch_run_loop(void *batch, int iters, double *d_sums, int *h_result_in_b) {
    int rc = check_batch(batch, "msl_batch_run_loop");
    if (rc) return rc;
    MSL_REQUIRE(iters >= 0, "msl_batch_run_loop: negative iteration count");
    MSL_REQUIRE(iters >= 1 || d_sums == nullptr, "msl_batch_run_loop: the fold needs at least one step");
    Batch *b = static_cast<Batch *>(batch);
    if (h_result_in_b) *h_result_in_b = iters & 1;
    for (int it = 0; it < iters; it++) {
        for (const msl_batch_grid &G : b->grids) {
            if (G.rows == 0 || G.cols == 0) continue;
            float *src = (it & 1) ? G.b : G.a;
            float *dst = (it & 1) ? G.a : G.b;
            rc = mean_step(dst, src, G.rows, G.cols, G.r, G.nv, 0);
            if (rc) return rc;
        }
    }
    if (d_sums) {
        const int ng = (int)b->grids.size();
        if (ng > 0) {
            batch_zero_f64_kernel<<<(ng + 255) / 256, 256, 0, ctx().stream>>>(d_sums, ng);
            rc = check_launch("msl_batch_run_loop");
            if (rc) return rc;
        }
        for (int g = 0; g < ng; g++) {
            const msl_batch_grid &G = b->grids[g];
            if (G.rows == 0 || G.cols == 0) continue;
            const float *res = (iters & 1) ? G.b : G.a;
            rc = msl_fold_sum_f32(res, (size_t)G.rows * G.cols, d_sums + g);
            if (rc) return rc;
        }
    }
    return MSL_OK;
}

}  // extern "C"

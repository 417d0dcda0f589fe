// Serving engine: skeleton jobs (2-D mean blur, 3-D heat, u32 map+fold)
// arrive over time and are served with continuous batching.
//
// Reference counterpart: none.  A Muesli program serves one request at a
// time: it builds the containers (src/ds.cpp:63-131), runs each skeleton
// with a synchronisation after every call (src/muesli.cpp:213-223) and
// folds (src/ds.cpp:739-912).  msl_serve_plain_loop keeps that model as the
// comparison baseline: one job at a time, one launch per step followed by a
// stream synchronisation, a fold copied to the host.
//
// The engine works in rounds (iteration-level batching):
//   * admission: pending jobs are admitted in arrival order while a slot is
//     free and their HBM bytes (two state buffers, or one for map+fold) fit
//     in the engine's arena, a single device allocation carved by a
//     first-fit allocator with coalescing (no per-job hipMalloc, no
//     fragmentation growth beyond the free list);
//   * every round advances each active stencil job by one step: one
//     persistent launch covers all active blur jobs' (strip, segment) items,
//     one covers all heat jobs' (tile, z-segment) items, one covers the
//     chunks of the map+fold jobs admitted this round; a job's state stays
//     resident in the arena across rounds.  When the active jobs' items
//     cannot fill the GPU (light load), a round advances each job by up to
//     kMaxSub steps instead, one launch per step, so that the per-round
//     host work and launches are spread over several steps;
//   * after a job's last step its final state is folded (compensated FP32
//     pairs per fixed 64K-element chunk, combined in FP64 in chunk order;
//     wrapping u32 for map+fold) and the round's results are copied to
//     pinned host memory;
//   * up to kRing rounds are in flight; a round's event retires its
//     finished jobs (results, latency stamps) and frees their memory, after
//     which waiting jobs can be admitted.
// Jobs that arrive while others are on the GPU join at the next round, so
// the work of differently sized jobs overlaps inside the same launches.
// Every stencil cell is computed in the operation order of the single-grid
// kernels whatever the tiling, and the fold partition depends only on the
// job's size, so every result is bit-identical whether the job runs alone
// or in a mix.  The tiling itself follows the load: at light load the blur
// segments are shortened until a job alone spreads over the device.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "blur_item.hpp"
#include "heat.hpp"

namespace mslhip {

int mean_step(float *out, const float *in, int rows, int cols, int r, float nv, int seg);
int heat_step(float *out, const float *in, int nx, int ny, int nz, float nv, float *maxdiff);

constexpr int kRing = 3;             // rounds in flight
constexpr int kMaxSub = 8;           // steps per round while the active items cannot fill the GPU
constexpr int kMapChunk = 1 << 16;   // elements per map+fold work item
constexpr int kHeatZseg = 16;        // z-planes per heat work item
constexpr size_t kAlign = 256;
const float kScale255 = 255.0f / 16777216.0f;
const float kScaleUnit = 1.0f / 16777216.0f;

// ---- per-round descriptors (device) ------------------------------------------
// A stencil job's step k of the round (k < nsub) reads state a when
// parity0 + k is even (b otherwise) and writes the other one.
struct SBlur {
    float *a;
    float *b;
    int rows, cols, r, kind;
    float nv;
    int strips_x, seg, item0, nitems, parity0, nsub;
};
struct SHeat {
    float *a;
    float *b;
    int nx, ny, nz;
    float nv;
    int xtiles, ytiles, zseg, item0, nitems, parity0, nsub;
};
// The final state of a finishing stencil job, folded in kMapChunk chunks.
struct SSum {
    const float *data;
    long long n;
    int item0, nitems;
};
struct SMap {
    uint32_t *data;
    long long n;
    uint32_t mul, add;
    int item0, nitems, fold, pad;
};
struct SFinal {
    int kind, item0, nitems, slot;
};

// item_desc[i] = largest j with starts[j] <= i (starts ascending, starts[0] = 0);
// also clears the round's nctr work-queue counters of this kind, one per step
// of the round (the work kernels run after this one on the same stream).
__global__ void serve_expand_kernel(const int *__restrict__ starts, int ndesc, int nitems, int *__restrict__ item_desc,
                                    unsigned *__restrict__ ctr, int nctr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nctr) ctr[i] = 0u;
    if (i >= nitems) return;
    int lo = 0, hi = ndesc - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    item_desc[i] = lo;
}

// Each wave takes its first item statically and every further one from the
// round's atomic counter (items differ in cost by radius and shape, so a
// static split leaves a long tail).
__device__ __forceinline__ int serve_next_item(unsigned *ctr, int stride) {
    unsigned nx = 0;
    if ((threadIdx.x & 63) == 0) nx = atomicAdd(ctr, 1u);
    return stride + __builtin_amdgcn_readfirstlane((int)nx);
}

// Step k of the round over every item of every active blur job (items of
// jobs with fewer steps in this round are skipped).
__global__ __launch_bounds__(256) void serve_blur_kernel(const SBlur *__restrict__ descs,
                                                         const int *__restrict__ item_desc, int nitems,
                                                         unsigned *__restrict__ ctr, int k_step) {
    __shared__ __attribute__((aligned(16))) float lds[4][2 * kBatchLdsw];
    const int wave = threadIdx.x >> 6;
    const int nw = gridDim.x * 4;
    for (int i = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave); i < nitems;
         i = serve_next_item(ctr, nw)) {
        const int j = __builtin_amdgcn_readfirstlane(item_desc[i]);
        const SBlur g = descs[j];
        if (k_step >= g.nsub) continue;
        const bool odd = (g.parity0 + k_step) & 1;
        const float *src = odd ? g.b : g.a;
        float *dst = odd ? g.a : g.b;
        const int k = i - g.item0;
        const int sw = g.kind == kKindNarrow1 ? 64 : 256;
        const int xs = (k % g.strips_x) * sw;
        const int sy = k / g.strips_x;
        const int y0 = sy * g.seg;
        const int y1 = min(y0 + g.seg, g.rows);
        LaneSum ls;
        blur_item_r<false>(g.r, src, dst, g.rows, g.cols, g.nv, g.kind, xs, y0, y1, lds[wave], &ls, sy & 1);
    }
}

__global__ __launch_bounds__(256) void serve_heat_kernel(const SHeat *__restrict__ descs,
                                                         const int *__restrict__ item_desc, int nitems,
                                                         unsigned *__restrict__ ctr, int k_step) {
    const int wave = threadIdx.x >> 6;
    const int nw = gridDim.x * 4;
    for (int i = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave); i < nitems;
         i = serve_next_item(ctr, nw)) {
        const int j = __builtin_amdgcn_readfirstlane(item_desc[i]);
        const SHeat g = descs[j];
        if (k_step >= g.nsub) continue;
        const bool odd = (g.parity0 + k_step) & 1;
        const float *src = odd ? g.b : g.a;
        float *dst = odd ? g.a : g.b;
        const int k = i - g.item0;
        const int xt = k % g.xtiles;
        const int rest = k / g.xtiles;
        const int yt = rest % g.ytiles;
        const int z0 = (rest / g.ytiles) * g.zseg;
        const int z1 = min(z0 + g.zseg, g.nz);
        float dmax = 0.0f;
        heat7_march<false, false>(dst, src, g.nx, g.ny, g.nz, g.nv, xt * kTX, yt * kTY, z0, z1, &dmax, nullptr);
    }
}

typedef uint32_t v4u_t __attribute__((ext_vector_type(4)));

// x <- x*mul + add over one chunk (wrapping u32), and the chunk's wrapping sum.
__global__ __launch_bounds__(256) void serve_map_kernel(const SMap *__restrict__ descs,
                                                        const int *__restrict__ item_desc, int nitems,
                                                        uint32_t *__restrict__ partials) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    for (int i = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave); i < nitems;
         i += nw) {
        const int j = __builtin_amdgcn_readfirstlane(item_desc[i]);
        const SMap g = descs[j];
        const long long e0 = (long long)(i - g.item0) * kMapChunk;
        const long long e1 = min(e0 + (long long)kMapChunk, g.n);
        const long long nq = (e1 - e0) / 4;
        uint32_t acc = 0;
        v4u_t *q = reinterpret_cast<v4u_t *>(g.data + e0);
        for (long long t = lane; t < nq; t += 64) {
            v4u_t v = q[t];
            v = v * g.mul + g.add;
            __builtin_nontemporal_store(v, q + t);
            acc += v.x + v.y + v.z + v.w;
        }
        for (long long e = e0 + nq * 4 + lane; e < e1; e += 64) {
            const uint32_t v = g.data[e] * g.mul + g.add;
            g.data[e] = v;
            acc += v;
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) acc += (uint32_t)__shfl_xor((int)acc, m);
        if (lane == 0) partials[i] = acc;
    }
}

// Compensated sum of one kMapChunk chunk of a finished stencil job's final
// state: each lane pushes its elements (16-byte loads, lane-strided) into a
// TwoSum pair, the wave combines the lanes in a fixed butterfly.
__global__ __launch_bounds__(256) void serve_sum_kernel(const SSum *__restrict__ descs,
                                                        const int *__restrict__ item_desc, int nitems,
                                                        float2 *__restrict__ partials) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    for (int i = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + wave); i < nitems;
         i += nw) {
        const int j = __builtin_amdgcn_readfirstlane(item_desc[i]);
        const SSum g = descs[j];
        const long long e0 = (long long)(i - g.item0) * kMapChunk;
        const long long e1 = min(e0 + (long long)kMapChunk, g.n);
        LaneSum ls;
        long long e = e0;
        if ((((uintptr_t)(g.data + e0)) & 15u) == 0) {
            typedef float v4f_t __attribute__((ext_vector_type(4)));
            const long long nq = (e1 - e0) / 4;
            const v4f_t *q = reinterpret_cast<const v4f_t *>(g.data + e0);
            for (long long t = lane; t < nq; t += 64) {
                const v4f_t v = q[t];
                ls.push(v.x);
                ls.push(v.y);
                ls.push(v.z);
                ls.push(v.w);
            }
            e = e0 + nq * 4;
        }
        for (e += lane; e < e1; e += 64) ls.push(g.data[e]);
        wave_pair_reduce(ls);
        if (lane == 0) partials[i] = make_float2(ls.s, ls.c);
    }
}

// One wave per finishing job: FP64 sum of its (s, c) chunk pairs in chunk
// order, or the wrapping sum of its u32 partials.
__global__ __launch_bounds__(256) void serve_final_kernel(const SFinal *__restrict__ finals, int nf,
                                                          const float2 *__restrict__ ps,
                                                          const uint32_t *__restrict__ pm, double *__restrict__ rf,
                                                          uint32_t *__restrict__ ru) {
    const int lane = threadIdx.x & 63;
    const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (f >= nf) return;
    const SFinal F = finals[f];
    if (F.kind == MSL_JOB_MAPFOLD_U32) {
        uint32_t acc = 0;
        for (int k = lane; k < F.nitems; k += 64) acc += pm[F.item0 + k];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) acc += (uint32_t)__shfl_xor((int)acc, m);
        if (lane == 0) ru[F.slot] = acc;
    } else {
        double acc = 0.0;
        for (int k = lane; k < F.nitems; k += 64) {
            const float2 v = ps[F.item0 + k];
            acc += (double)v.x;
            acc += (double)v.y;
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
        if (lane == 0) rf[F.slot] = acc;
    }
}

// ---- host side ----------------------------------------------------------------
static double now_s() {
    using clk = std::chrono::steady_clock;
    static const clk::time_point t0 = clk::now();
    return std::chrono::duration<double>(clk::now() - t0).count();
}

// First-fit allocator over [0, cap) with coalescing of free neighbours.
class Arena {
public:
    void reset(size_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
        cap_ = cap;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = (bytes + kAlign - 1) / kAlign * kAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t roff = it->first + bytes;
                free_.erase(it);
                if (rest) free_[roff] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + kAlign - 1) / kAlign * kAlign;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t capacity() const { return cap_; }

private:
    std::map<size_t, size_t> free_;
    size_t cap_ = 0, used_ = 0;
};

enum JobState { kQueued = 0, kActive = 1, kDone = 2 };

struct SJob {
    int64_t id = 0;
    msl_job spec{};
    JobState state = kQueued;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    size_t off = 0, bytes = 0;
    bool holds_memory = false;
    float *a = nullptr, *b = nullptr;
    uint32_t *u = nullptr;
    int steps_done = 0;
    int nsub = 0;  // steps in the round being planned
    // tiling
    int kind = 0, strips_x = 0, seg = 0, xtiles = 0, ytiles = 0, nitems = 0;
    double sum = 0.0;
    uint32_t sum_u32 = 0;
};

struct RoundSlot {
    bool inflight = false;
    hipEvent_t ev = nullptr;
    size_t stage_cap = 0;
    char *h_stage = nullptr;  // pinned
    char *d_stage = nullptr;
    int item_cap[4] = {0, 0, 0, 0};  // blur, heat, map, sum items
    int *d_item_desc[4] = {nullptr, nullptr, nullptr, nullptr};
    float2 *d_ps = nullptr;  // sum chunk partials
    uint32_t *d_pm = nullptr;
    unsigned *d_ctr = nullptr;  // work-queue counters, kMaxSub each: blur, heat, then map
    double *d_rf = nullptr, *h_rf = nullptr;
    uint32_t *d_ru = nullptr, *h_ru = nullptr;
    std::vector<std::pair<SJob *, int>> finishing;
};

struct Server {
    std::mutex mu;  // pending queue and job table
    std::deque<SJob *> pending;
    std::unordered_map<int64_t, SJob *> jobs;
    int64_t next_id = 1;
    int max_active = 0;
    char *base = nullptr;
    Arena arena;
    std::vector<SJob *> blur, heat;  // active stencil jobs
    RoundSlot ring[kRing];
    int64_t round_no = 0;
    int64_t done_rounds = 0;
    size_t peak_bytes = 0;
    double active_sum = 0.0;  // sum over rounds of active jobs (for the mean)
    int64_t done_jobs = 0;
};

static size_t job_bytes(const msl_job &j) {
    auto al = [](size_t x) { return (x + kAlign - 1) / kAlign * kAlign; };
    switch (j.kind) {
        case MSL_JOB_BLUR2D: return 2 * al((size_t)j.d0 * j.d1 * 4);
        case MSL_JOB_HEAT3D: return 2 * al((size_t)j.d0 * j.d1 * j.d2 * 4);
        default: return al((size_t)j.n * 4);
    }
}

static int validate_job(const msl_job &j, size_t budget) {
    switch (j.kind) {
        case MSL_JOB_BLUR2D:
            MSL_REQUIRE(j.d0 >= 1 && j.d1 >= 1 && (int64_t)j.d0 * j.d1 <= INT32_MAX && j.d1 < (1 << 30),
                        "msl_job: bad 2-D shape");
            MSL_REQUIRE(j.r >= 1 && j.r <= kBatchMaxR, "msl_job: radius must be in [1, 10]");
            MSL_REQUIRE(j.steps >= 1, "msl_job: steps must be >= 1");
            break;
        case MSL_JOB_HEAT3D:
            MSL_REQUIRE(j.d0 >= 1 && j.d1 >= 1 && j.d2 >= 1 && (int64_t)j.d0 * j.d1 * j.d2 <= INT32_MAX &&
                            (int64_t)j.d0 * j.d1 < (1 << 30),
                        "msl_job: bad 3-D shape");
            MSL_REQUIRE(j.steps >= 1, "msl_job: steps must be >= 1");
            break;
        case MSL_JOB_MAPFOLD_U32:
            MSL_REQUIRE(j.n >= 1 && j.n <= ((int64_t)1 << 36), "msl_job: bad length");
            break;
        default: return fail(MSL_E_INVALID, "msl_job: unknown kind");
    }
    if (job_bytes(j) > budget) return fail(MSL_E_NOMEM, "msl_job: job larger than the server's HBM budget");
    return MSL_OK;
}

// A job's strip layout and (for the fold) chunk count, fixed at admission.
// Blur segments of a round: 64 rows (32 for one column per lane) under
// load; at light load (fine) short enough that a job alone spreads over the
// device, about 4 waves per CU (a large grid served alone otherwise ran on an
// eighth of the GPU), but at least 4 / 8 / 16 rows for r <= 2, <= 6 and
// above so that the 2r halo rows stay a modest share.  The results do not
// depend on the segments.
static void tile_blur(SJob &J, bool fine) {
    const msl_job &s = J.spec;
    const int rows = s.d0;
    const int seg_max = J.kind == kKindNarrow1 ? 32 : 64;
    J.seg = seg_max;
    if (fine) {
        const long target = (long)ctx().num_cus * 4;
        const long segs = (target + J.strips_x - 1) / J.strips_x;
        const int seg_min = s.r <= 2 ? 4 : (s.r <= 6 ? 8 : 16);
        J.seg = std::min(seg_max, std::max(seg_min, (int)((rows + segs - 1) / segs)));
    }
    J.nitems = J.strips_x * ((rows + J.seg - 1) / J.seg);
}

static void plan_job(SJob &J) {
    const msl_job &s = J.spec;
    if (s.kind == MSL_JOB_BLUR2D) {
        const int cols = s.d1;
        const bool narrow = cols < 192 || s.r > kBatchMaxR4 || cols < 4;
        const uintptr_t al = (uintptr_t)J.a | (uintptr_t)J.b;
        J.kind = narrow ? kKindNarrow1 : (((cols & 3) == 0 && (al & 15u) == 0) ? kKindAligned4 : kKindUnaligned4);
        const int sw = narrow ? 64 : 256;
        J.strips_x = (cols + sw - 1) / sw;
        tile_blur(J, false);
    } else if (s.kind == MSL_JOB_HEAT3D) {
        J.xtiles = (s.d0 + kTX - 1) / kTX;
        J.ytiles = (s.d1 + kTY - 1) / kTY;
        J.seg = kHeatZseg;
        J.nitems = J.xtiles * J.ytiles * ((s.d2 + kHeatZseg - 1) / kHeatZseg);
    } else {
        J.nitems = (int)((s.n + kMapChunk - 1) / kMapChunk);
    }
}

static int slot_alloc(RoundSlot &R, int max_active) {
    MSL_HIP_TRY(hipEventCreateWithFlags(&R.ev, hipEventDisableTiming));
    MSL_HIP_TRY(hipHostMalloc((void **)&R.h_rf, sizeof(double) * max_active, hipHostMallocDefault));
    MSL_HIP_TRY(hipHostMalloc((void **)&R.h_ru, sizeof(uint32_t) * max_active, hipHostMallocDefault));
    MSL_HIP_TRY(hipMalloc(&R.d_rf, sizeof(double) * max_active));
    MSL_HIP_TRY(hipMalloc(&R.d_ru, sizeof(uint32_t) * max_active));
    return MSL_OK;
}

static void slot_free(RoundSlot &R) {
    if (R.ev) (void)hipEventDestroy(R.ev);
    if (R.h_stage) (void)hipHostFree(R.h_stage);
    if (R.d_stage) (void)hipFree(R.d_stage);
    for (int k = 0; k < 4; k++)
        if (R.d_item_desc[k]) (void)hipFree(R.d_item_desc[k]);
    if (R.d_ps) (void)hipFree(R.d_ps);
    if (R.d_pm) (void)hipFree(R.d_pm);
    if (R.d_ctr) (void)hipFree(R.d_ctr);
    if (R.d_rf) (void)hipFree(R.d_rf);
    if (R.d_ru) (void)hipFree(R.d_ru);
    if (R.h_rf) (void)hipHostFree(R.h_rf);
    if (R.h_ru) (void)hipHostFree(R.h_ru);
    R = RoundSlot();
}

// Grows a slot's buffers (only called while the slot is not in flight).
static int slot_reserve(RoundSlot &R, size_t stage_bytes, const int items[4]) {
    if (!R.d_ctr) MSL_HIP_TRY(hipMalloc(&R.d_ctr, sizeof(unsigned) * 4 * kMaxSub));
    if (stage_bytes > R.stage_cap) {
        if (R.h_stage) MSL_HIP_TRY(hipHostFree(R.h_stage));
        if (R.d_stage) MSL_HIP_TRY(hipFree(R.d_stage));
        R.h_stage = nullptr;
        R.d_stage = nullptr;
        const size_t cap = std::max(stage_bytes * 2, (size_t)1 << 16);
        MSL_HIP_TRY(hipHostMalloc((void **)&R.h_stage, cap, hipHostMallocDefault));
        MSL_HIP_TRY(hipMalloc(&R.d_stage, cap));
        R.stage_cap = cap;
    }
    for (int k = 0; k < 4; k++) {
        if (items[k] <= R.item_cap[k]) continue;
        const int cap = std::max(items[k] + items[k] / 2, 4096);
        if (R.d_item_desc[k]) MSL_HIP_TRY(hipFree(R.d_item_desc[k]));
        R.d_item_desc[k] = nullptr;
        MSL_HIP_TRY(hipMalloc(&R.d_item_desc[k], sizeof(int) * cap));
        if (k == 2) {
            if (R.d_pm) MSL_HIP_TRY(hipFree(R.d_pm));
            R.d_pm = nullptr;
            MSL_HIP_TRY(hipMalloc(&R.d_pm, sizeof(uint32_t) * cap));
        } else if (k == 3) {
            if (R.d_ps) MSL_HIP_TRY(hipFree(R.d_ps));
            R.d_ps = nullptr;
            MSL_HIP_TRY(hipMalloc(&R.d_ps, sizeof(float2) * cap));
        }
        R.item_cap[k] = cap;
    }
    return MSL_OK;
}

static void release_job_memory(Server &S, SJob &J) {
    if (J.holds_memory) {
        S.arena.release(J.off, J.bytes);
        J.holds_memory = false;
    }
}

// Results of a completed round: stamps, sums, memory.
static void retire(Server &S, RoundSlot &R) {
    const double t = now_s();
    std::lock_guard<std::mutex> lock(S.mu);
    for (auto &fj : R.finishing) {
        SJob &J = *fj.first;
        if (J.spec.kind == MSL_JOB_MAPFOLD_U32) J.sum_u32 = R.h_ru[fj.second];
        else J.sum = R.h_rf[fj.second];
        J.t_done = t;
        J.state = kDone;
        if (!J.spec.keep_final) release_job_memory(S, J);
        S.done_jobs++;
    }
    R.finishing.clear();
    R.inflight = false;
    S.done_rounds++;
}

static int wait_slot(Server &S, RoundSlot &R) {
    if (!R.inflight) return MSL_OK;
    MSL_HIP_TRY(hipEventSynchronize(R.ev));
    retire(S, R);
    return MSL_OK;
}

// Retires every in-flight round that has completed, oldest first.
static int poll_rounds(Server &S) {
    for (int64_t k = S.round_no - kRing; k < S.round_no; k++) {
        if (k < 0) continue;
        RoundSlot &R = S.ring[k % kRing];
        if (!R.inflight) continue;
        hipError_t e = hipEventQuery(R.ev);
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) return hip_fail(e, "msl_server: round event");
        retire(S, R);
    }
    return MSL_OK;
}

// Waits for the oldest in-flight round (if any) and retires it.
static int wait_oldest(Server &S) {
    for (int64_t k = S.round_no - kRing; k < S.round_no; k++) {
        if (k < 0) continue;
        RoundSlot &R = S.ring[k % kRing];
        if (R.inflight) return wait_slot(S, R);
    }
    return MSL_OK;
}

static int inflight_rounds(const Server &S) {
    int n = 0;
    for (const RoundSlot &R : S.ring) n += R.inflight ? 1 : 0;
    return n;
}

template <typename T>
static T *stage_put(char *base, size_t &off, const T *src, size_t n) {
    off = (off + 15) / 16 * 16;
    T *p = reinterpret_cast<T *>(base + off);
    if (n) std::memcpy(p, src, sizeof(T) * n);
    off += sizeof(T) * n;
    return p;
}

// One round.  *did_work = 0 when there was nothing to launch.
static int run_round(Server &S, int *did_work) {
    *did_work = 0;
    RoundSlot &R = S.ring[S.round_no % kRing];
    int rc = wait_slot(S, R);
    if (rc) return rc;
    hipStream_t st = ctx().stream;

    // ---- admission
    std::vector<SJob *> maps;
    {
        std::lock_guard<std::mutex> lock(S.mu);
        while (!S.pending.empty()) {
            SJob *J = S.pending.front();
            const int active = (int)(S.blur.size() + S.heat.size() + maps.size());
            if (active >= S.max_active) break;
            size_t off = 0;
            if (!S.arena.alloc(J->bytes, &off)) break;
            S.pending.pop_front();
            J->off = off;
            J->holds_memory = true;
            J->state = kActive;
            J->t_admit = now_s();
            char *p = S.base + off;
            const msl_job &s = J->spec;
            if (s.kind == MSL_JOB_MAPFOLD_U32) {
                J->u = reinterpret_cast<uint32_t *>(p);
            } else {
                J->a = reinterpret_cast<float *>(p);
                J->b = reinterpret_cast<float *>(p + J->bytes / 2);
            }
            plan_job(*J);
            int frc = MSL_OK;
            if (s.kind == MSL_JOB_BLUR2D) {
                frc = msl_fill_uniform_f32(J->a, (size_t)s.d0 * s.d1, s.seed, kScale255);
                S.blur.push_back(J);
            } else if (s.kind == MSL_JOB_HEAT3D) {
                frc = msl_fill_uniform_f32(J->a, (size_t)s.d0 * s.d1 * s.d2, s.seed, kScaleUnit);
                S.heat.push_back(J);
            } else {
                frc = msl_fill_u32(J->u, (size_t)s.n, s.seed);
                maps.push_back(J);
            }
            if (frc) return frc;
        }
        S.peak_bytes = std::max(S.peak_bytes, S.arena.used());
    }
    if (S.blur.empty() && S.heat.empty() && maps.empty()) return MSL_OK;

    // ---- descriptors of this round
    std::stable_sort(S.blur.begin(), S.blur.end(), [](const SJob *x, const SJob *y) {
        if (x->spec.r != y->spec.r) return x->spec.r < y->spec.r;
        return x->kind < y->kind;
    });
    std::vector<SBlur> db;
    std::vector<SHeat> dh;
    std::vector<SMap> dm;
    std::vector<SSum> ds;
    std::vector<SFinal> fin;
    std::vector<int> sb, sh, sm, ss;
    std::vector<std::pair<SJob *, int>> finishing;
    int items[4] = {0, 0, 0, 0};
    // Light load: the active stencil jobs' items at the default tiling
    // cannot fill one persistent launch (3 workgroups of 4 waves per CU).
    // Then the blur segments are shortened and each job advances up to
    // kMaxSub steps this round (per-round host work and launch latency would
    // otherwise dominate); under load one step with the default tiling.
    long stencil_items = 0;
    for (SJob *J : S.blur) {
        tile_blur(*J, false);
        stencil_items += J->nitems;
    }
    for (SJob *J : S.heat) stencil_items += J->nitems;
    const bool light = stencil_items < (long)ctx().num_cus * 12;
    const int ksteps = light ? kMaxSub : 1;
    if (light)
        for (SJob *J : S.blur) tile_blur(*J, true);
    // the fold of a finishing stencil job: its final state in kMapChunk chunks
    auto add_sum = [&](SJob *J, const float *data, long long n, int slot) {
        const int nchunks = (int)((n + kMapChunk - 1) / kMapChunk);
        ds.push_back(SSum{data, n, items[3], nchunks});
        ss.push_back(items[3]);
        fin.push_back(SFinal{J->spec.kind, items[3], nchunks, slot});
        finishing.emplace_back(J, slot);
        items[3] += nchunks;
    };
    int nsub_max[2] = {0, 0};
    for (SJob *J : S.blur) {
        const msl_job &s = J->spec;
        J->nsub = std::min(ksteps, s.steps - J->steps_done);
        nsub_max[0] = std::max(nsub_max[0], J->nsub);
        db.push_back(SBlur{J->a, J->b, s.d0, s.d1, s.r, J->kind, s.nv, J->strips_x, J->seg, items[0], J->nitems,
                           J->steps_done & 1, J->nsub});
        sb.push_back(items[0]);
        if (J->steps_done + J->nsub == s.steps)
            add_sum(J, (s.steps & 1) ? J->b : J->a, (long long)s.d0 * s.d1, (int)fin.size());
        items[0] += J->nitems;
    }
    for (SJob *J : S.heat) {
        const msl_job &s = J->spec;
        J->nsub = std::min(ksteps, s.steps - J->steps_done);
        nsub_max[1] = std::max(nsub_max[1], J->nsub);
        dh.push_back(SHeat{J->a, J->b, s.d0, s.d1, s.d2, s.nv, J->xtiles, J->ytiles, J->seg, items[1], J->nitems,
                           J->steps_done & 1, J->nsub});
        sh.push_back(items[1]);
        if (J->steps_done + J->nsub == s.steps)
            add_sum(J, (s.steps & 1) ? J->b : J->a, (long long)s.d0 * s.d1 * s.d2, (int)fin.size());
        items[1] += J->nitems;
    }
    for (SJob *J : maps) {
        const msl_job &s = J->spec;
        const int slot = (int)fin.size();
        dm.push_back(SMap{J->u, (long long)s.n, s.mul, s.add, items[2], J->nitems, slot, 0});
        sm.push_back(items[2]);
        fin.push_back(SFinal{MSL_JOB_MAPFOLD_U32, items[2], J->nitems, slot});
        finishing.emplace_back(J, slot);
        items[2] += J->nitems;
    }

    // ---- stage and upload
    const size_t stage_bytes = 16 * 10 + sizeof(SBlur) * db.size() + sizeof(SHeat) * dh.size() +
                               sizeof(SMap) * dm.size() + sizeof(SSum) * ds.size() + sizeof(SFinal) * fin.size() +
                               sizeof(int) * (sb.size() + sh.size() + sm.size() + ss.size());
    rc = slot_reserve(R, stage_bytes, items);
    if (rc) return rc;
    size_t off = 0;
    char *hb = R.h_stage;
    SBlur *pb = stage_put(hb, off, db.data(), db.size());
    SHeat *ph = stage_put(hb, off, dh.data(), dh.size());
    SMap *pm = stage_put(hb, off, dm.data(), dm.size());
    SSum *pss = stage_put(hb, off, ds.data(), ds.size());
    SFinal *pf = stage_put(hb, off, fin.data(), fin.size());
    int *psb = stage_put(hb, off, sb.data(), sb.size());
    int *psh = stage_put(hb, off, sh.data(), sh.size());
    int *psm = stage_put(hb, off, sm.data(), sm.size());
    int *pss_start = stage_put(hb, off, ss.data(), ss.size());
    MSL_HIP_TRY(hipMemcpyAsync(R.d_stage, R.h_stage, off, hipMemcpyHostToDevice, st));
    auto dev = [&](const void *hp) { return R.d_stage + ((const char *)hp - hb); };
    const SBlur *gb = (const SBlur *)dev(pb);
    const SHeat *gh = (const SHeat *)dev(ph);
    const SMap *gm = (const SMap *)dev(pm);
    const SSum *gss = (const SSum *)dev(pss);
    const SFinal *gf = (const SFinal *)dev(pf);
    const int *gsb = (const int *)dev(psb);
    const int *gsh = (const int *)dev(psh);
    const int *gsm = (const int *)dev(psm);
    const int *gss_start = (const int *)dev(pss_start);

    // ---- launches
    const int cus = ctx().num_cus;
    auto persistent = [&](int n, int per_cu) { return std::max(1, std::min((n + 3) / 4, cus * per_cu)); };
    // (map+fold jobs run before the stencil steps: they are single-pass and
    // must not wait behind several steps of the stencil jobs)
    if (items[2]) {
        serve_expand_kernel<<<(items[2] + 255) / 256, 256, 0, st>>>(gsm, (int)dm.size(), items[2], R.d_item_desc[2],
                                                                    R.d_ctr + 2 * kMaxSub, 1);
        serve_map_kernel<<<persistent(items[2], 4), 256, 0, st>>>(gm, R.d_item_desc[2], items[2], R.d_pm);
    }
    if (items[0])
        serve_expand_kernel<<<(items[0] + 255) / 256, 256, 0, st>>>(gsb, (int)db.size(), items[0], R.d_item_desc[0],
                                                                    R.d_ctr, nsub_max[0]);
    if (items[1])
        serve_expand_kernel<<<(items[1] + 255) / 256, 256, 0, st>>>(gsh, (int)dh.size(), items[1], R.d_item_desc[1],
                                                                    R.d_ctr + kMaxSub, nsub_max[1]);
    for (int k = 0; k < std::max(nsub_max[0], nsub_max[1]); k++) {
        if (k < nsub_max[0])
            serve_blur_kernel<<<persistent(items[0], 3), 256, 0, st>>>(gb, R.d_item_desc[0], items[0], R.d_ctr + k,
                                                                       k);
        if (k < nsub_max[1])
            serve_heat_kernel<<<persistent(items[1], 3), 256, 0, st>>>(gh, R.d_item_desc[1], items[1],
                                                                       R.d_ctr + kMaxSub + k, k);
    }
    if (items[3]) {
        serve_expand_kernel<<<(items[3] + 255) / 256, 256, 0, st>>>(gss_start, (int)ds.size(), items[3],
                                                                    R.d_item_desc[3], R.d_ctr + 3 * kMaxSub, 1);
        serve_sum_kernel<<<persistent(items[3], 4), 256, 0, st>>>(gss, R.d_item_desc[3], items[3], R.d_ps);
    }
    if (!fin.empty()) {
        serve_final_kernel<<<((int)fin.size() + 3) / 4, 256, 0, st>>>(gf, (int)fin.size(), R.d_ps, R.d_pm,
                                                                      R.d_rf, R.d_ru);
        MSL_HIP_TRY(hipMemcpyAsync(R.h_rf, R.d_rf, sizeof(double) * fin.size(), hipMemcpyDeviceToHost, st));
        MSL_HIP_TRY(hipMemcpyAsync(R.h_ru, R.d_ru, sizeof(uint32_t) * fin.size(), hipMemcpyDeviceToHost, st));
    }
    rc = check_launch("msl_server round");
    if (rc) return rc;
    MSL_HIP_TRY(hipEventRecord(R.ev, st));
    R.inflight = true;
    R.finishing = std::move(finishing);

    // ---- advance the jobs
    S.active_sum += (double)(S.blur.size() + S.heat.size() + maps.size());
    auto advance = [](std::vector<SJob *> &v) {
        size_t w = 0;
        for (SJob *J : v) {
            J->steps_done += J->nsub;
            if (J->steps_done < J->spec.steps) v[w++] = J;
        }
        v.resize(w);
    };
    advance(S.blur);
    advance(S.heat);
    S.round_no++;
    *did_work = 1;
    return MSL_OK;
}

static bool has_work(Server &S) {
    std::lock_guard<std::mutex> lock(S.mu);
    return !S.blur.empty() || !S.heat.empty() || !S.pending.empty();
}

static void fill_result(const SJob &J, msl_job_result *r) {
    r->id = J.id;
    r->kind = J.spec.kind;
    r->sum = J.sum;
    r->sum_u32 = J.sum_u32;
    r->t_arrival = J.t_arrival;
    r->t_admit = J.t_admit;
    r->t_done = J.t_done;
    r->bytes = J.bytes;
}

static int submit(Server &S, const msl_job &job, double t_arrival, int64_t *id) {
    int rc = validate_job(job, S.arena.capacity());
    if (rc) return rc;
    SJob *J = new SJob();
    J->spec = job;
    J->bytes = job_bytes(job);
    J->t_arrival = t_arrival;
    std::lock_guard<std::mutex> lock(S.mu);
    J->id = S.next_id++;
    S.jobs[J->id] = J;
    S.pending.push_back(J);
    if (id) *id = J->id;
    return MSL_OK;
}

static int check_server(void *h, const char *what) {
    int rc = ensure_ready();
    if (rc) return rc;
    if (!h) return fail(MSL_E_INVALID, std::string(what) + ": null server");
    if (ctx().capturing) return fail(MSL_E_STATE, std::string(what) + ": not allowed during graph capture");
    return MSL_OK;
}

static int drain(Server &S) {
    for (;;) {
        int did = 0;
        int rc = run_round(S, &did);
        if (rc) return rc;
        if (did) continue;
        if (inflight_rounds(S) == 0) {
            if (!has_work(S)) return MSL_OK;
            return fail(MSL_E_STATE, "msl_server: pending jobs cannot be admitted (memory held by unpolled results)");
        }
        for (RoundSlot &R : S.ring) {
            rc = wait_slot(S, R);
            if (rc) return rc;
        }
    }
}

}  // namespace mslhip

using namespace mslhip;

extern "C" {

int msl_server_create(size_t hbm_budget_bytes, int max_active, void **h_server) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(h_server != nullptr, "msl_server_create: null out-pointer");
    MSL_REQUIRE(hbm_budget_bytes >= kAlign && max_active >= 1, "msl_server_create: bad budget or slot count");
    if (ctx().capturing) return fail(MSL_E_STATE, "msl_server_create: not allowed during graph capture");
    Server *S = new Server();
    S->max_active = max_active;
    hipError_t e = hipMalloc(&S->base, hbm_budget_bytes);
    if (e != hipSuccess) {
        delete S;
        return fail(MSL_E_NOMEM, "msl_server_create: cannot reserve the HBM budget");
    }
    S->arena.reset(hbm_budget_bytes);
    for (RoundSlot &R : S->ring) {
        rc = slot_alloc(R, max_active);
        if (rc) {
            msl_server_destroy(S);
            return rc;
        }
    }
    *h_server = S;
    return MSL_OK;
}

int msl_server_destroy(void *server) {
    if (!server) return MSL_OK;
    Server *S = static_cast<Server *>(server);
    if (ctx().ready) (void)hipStreamSynchronize(ctx().stream);
    for (RoundSlot &R : S->ring) slot_free(R);
    for (auto &kv : S->jobs) delete kv.second;
    if (S->base) (void)hipFree(S->base);
    delete S;
    return MSL_OK;
}

int msl_server_submit(void *server, const msl_job *job, double t_arrival, int64_t *h_id) {
    MSL_REQUIRE(server && job, "msl_server_submit: null argument");
    Server *S = static_cast<Server *>(server);
    return submit(*S, *job, t_arrival < 0 ? now_s() : t_arrival, h_id);
}

int msl_server_step(void *server, int *h_active, int *h_pending) {
    int rc = check_server(server, "msl_server_step");
    if (rc) return rc;
    Server *S = static_cast<Server *>(server);
    rc = poll_rounds(*S);
    if (rc) return rc;
    int did = 0;
    rc = run_round(*S, &did);
    if (rc) return rc;
    if (!did) {  // nothing to launch: make progress by retiring the oldest round
        rc = wait_oldest(*S);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lock(S->mu);
    if (h_active) *h_active = (int)(S->blur.size() + S->heat.size());
    if (h_pending) *h_pending = (int)S->pending.size();
    return MSL_OK;
}

int msl_server_drain(void *server) {
    int rc = check_server(server, "msl_server_drain");
    if (rc) return rc;
    return drain(*static_cast<Server *>(server));
}

int msl_server_poll(void *server, int64_t id, msl_job_result *h_result, void *h_final, size_t final_bytes) {
    int rc = check_server(server, "msl_server_poll");
    if (rc) return rc;
    Server *S = static_cast<Server *>(server);
    rc = poll_rounds(*S);
    if (rc) return rc;
    SJob *J = nullptr;
    {
        std::lock_guard<std::mutex> lock(S->mu);
        auto it = S->jobs.find(id);
        if (it == S->jobs.end()) return fail(MSL_E_INVALID, "msl_server_poll: unknown job id");
        J = it->second;
        if (J->state != kDone) return MSL_NOT_READY;
    }
    if (h_result) fill_result(*J, h_result);
    if (h_final) {
        const msl_job &s = J->spec;
        if (!s.keep_final || !J->holds_memory) return fail(MSL_E_STATE, "msl_server_poll: final state not kept");
        size_t nbytes = 0;
        const void *src = nullptr;
        if (s.kind == MSL_JOB_MAPFOLD_U32) {
            nbytes = (size_t)s.n * 4;
            src = J->u;
        } else {
            nbytes = (s.kind == MSL_JOB_BLUR2D ? (size_t)s.d0 * s.d1 : (size_t)s.d0 * s.d1 * s.d2) * 4;
            src = (s.steps & 1) ? J->b : J->a;
        }
        MSL_REQUIRE(final_bytes >= nbytes, "msl_server_poll: final buffer too small");
        MSL_HIP_TRY(hipMemcpyAsync(h_final, src, nbytes, hipMemcpyDeviceToHost, ctx().stream));
        MSL_HIP_TRY(hipStreamSynchronize(ctx().stream));
    }
    std::lock_guard<std::mutex> lock(S->mu);
    release_job_memory(*S, *J);
    S->jobs.erase(id);
    delete J;
    return MSL_OK;
}

int msl_server_info(const void *server, int64_t *h_rounds, size_t *h_peak_bytes, double *h_mean_active,
                    int64_t *h_done_jobs) {
    MSL_REQUIRE(server != nullptr, "msl_server_info: null server");
    const Server *S = static_cast<const Server *>(server);
    if (h_rounds) *h_rounds = S->round_no;
    if (h_peak_bytes) *h_peak_bytes = S->peak_bytes;
    if (h_mean_active) *h_mean_active = S->round_no ? S->active_sum / (double)S->round_no : 0.0;
    if (h_done_jobs) *h_done_jobs = S->done_jobs;
    return MSL_OK;
}

// Replays an arrival trace in real time: job i is submitted once
// arrival_s[i] seconds have passed since the call began; returns when every
// job is done, with its result (times relative to the start of the call).
int msl_server_run_trace(void *server, const msl_job *jobs, const double *arrival_s, int njobs,
                         msl_job_result *h_results) {
    int rc = check_server(server, "msl_server_run_trace");
    if (rc) return rc;
    MSL_REQUIRE(njobs >= 0 && (njobs == 0 || (jobs && arrival_s && h_results)), "msl_server_run_trace: bad trace");
    Server *S = static_cast<Server *>(server);
    for (int i = 0; i < njobs; i++) {
        rc = validate_job(jobs[i], S->arena.capacity());
        if (rc) return rc;
        MSL_REQUIRE(i == 0 || arrival_s[i] >= arrival_s[i - 1], "msl_server_run_trace: arrivals must be sorted");
    }
    MSL_REQUIRE(!has_work(*S) && inflight_rounds(*S) == 0, "msl_server_run_trace: the server must be idle");
    std::vector<int64_t> ids(njobs);
    const int64_t done0 = S->done_jobs;
    const double t0 = now_s();
    int next = 0;
    for (;;) {
        const double t = now_s() - t0;
        while (next < njobs && arrival_s[next] <= t) {
            msl_job j = jobs[next];
            j.keep_final = 0;
            rc = submit(*S, j, t0 + arrival_s[next], &ids[next]);
            if (rc) return rc;
            next++;
        }
        rc = poll_rounds(*S);
        if (rc) return rc;
        if (S->done_jobs - done0 == njobs) break;
        int did = 0;
        rc = run_round(*S, &did);
        if (rc) return rc;
        if (did) continue;
        if (inflight_rounds(*S) > 0) {
            rc = wait_oldest(*S);  // nothing new to launch
            if (rc) return rc;
        } else if (next < njobs) {
            const double wait = arrival_s[next] - (now_s() - t0);
            if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
        } else if (has_work(*S)) {
            return fail(MSL_E_STATE, "msl_server_run_trace: pending jobs cannot be admitted");
        }
    }
    for (int i = 0; i < njobs; i++) {
        std::lock_guard<std::mutex> lock(S->mu);
        SJob *J = S->jobs[ids[i]];
        fill_result(*J, &h_results[i]);
        h_results[i].t_arrival -= t0;
        h_results[i].t_admit -= t0;
        h_results[i].t_done -= t0;
        release_job_memory(*S, *J);
        S->jobs.erase(ids[i]);
        delete J;
    }
    return MSL_OK;
}

// The comparison baseline: one job at a time in arrival order, the job's
// buffers from a preallocated region, every step one launch followed by a
// stream synchronisation, then a fold copied to the host.
int msl_serve_plain_loop(const msl_job *jobs, const double *arrival_s, int njobs, size_t hbm_budget_bytes,
                         msl_job_result *h_results) {
    int rc = ensure_ready();
    if (rc) return rc;
    MSL_REQUIRE(njobs >= 0 && (njobs == 0 || (jobs && arrival_s && h_results)), "msl_serve_plain_loop: bad trace");
    if (ctx().capturing) return fail(MSL_E_STATE, "msl_serve_plain_loop: not allowed during graph capture");
    size_t need = 0;
    for (int i = 0; i < njobs; i++) {
        rc = validate_job(jobs[i], hbm_budget_bytes);
        if (rc) return rc;
        need = std::max(need, job_bytes(jobs[i]));
    }
    char *base = nullptr;
    if (need) MSL_HIP_TRY(hipMalloc(&base, need));
    hipStream_t st = ctx().stream;
    const double t0 = now_s();
    rc = MSL_OK;
    for (int i = 0; i < njobs && rc == MSL_OK; i++) {
        const msl_job &s = jobs[i];
        while (now_s() - t0 < arrival_s[i]) std::this_thread::yield();
        msl_job_result &res = h_results[i];
        std::memset(&res, 0, sizeof res);
        res.id = i;
        res.kind = s.kind;
        res.t_arrival = arrival_s[i];
        res.t_admit = now_s() - t0;
        res.bytes = job_bytes(s);
        if (s.kind == MSL_JOB_MAPFOLD_U32) {
            uint32_t *u = reinterpret_cast<uint32_t *>(base);
            rc = msl_fill_u32(u, (size_t)s.n, s.seed);
            if (!rc) rc = msl_map_affine_u32(u, u, (size_t)s.n, s.mul, s.add);
            if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(MSL_E_HIP, "msl_serve_plain_loop: sync");
            if (!rc) rc = msl_fold_sum_u32_host(u, (size_t)s.n, &res.sum_u32);
        } else {
            const size_t cells = s.kind == MSL_JOB_BLUR2D ? (size_t)s.d0 * s.d1 : (size_t)s.d0 * s.d1 * s.d2;
            float *a = reinterpret_cast<float *>(base);
            float *b = reinterpret_cast<float *>(base + job_bytes(s) / 2);
            rc = msl_fill_uniform_f32(a, cells, s.seed, s.kind == MSL_JOB_BLUR2D ? kScale255 : kScaleUnit);
            for (int it = 0; it < s.steps && rc == MSL_OK; it++) {
                float *src = (it & 1) ? b : a;
                float *dst = (it & 1) ? a : b;
                rc = s.kind == MSL_JOB_BLUR2D ? mean_step(dst, src, s.d0, s.d1, s.r, s.nv, 0)
                                              : heat_step(dst, src, s.d0, s.d1, s.d2, s.nv, nullptr);
                if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(MSL_E_HIP, "msl_serve_plain_loop: sync");
            }
            if (!rc) rc = msl_fold_sum_f32_host((s.steps & 1) ? b : a, cells, &res.sum, nullptr);
        }
        res.t_done = now_s() - t0;
    }
    if (base) (void)hipFree(base);
    return rc;
}

}  // extern "C"

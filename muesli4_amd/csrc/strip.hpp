// Wavefront strip march for 2-D stencils of radius R (mean blur, Game of
// Life), shared by the single-grid kernels and the batched persistent kernel.
//
// One wavefront owns a strip of 64*C columns (C adjacent columns per lane)
// and a segment of rows, and marches down the segment:
//   1. the input row y+R (strip plus R halo columns each side) is loaded,
//      rows prefetched D ahead in registers, and staged in a per-wave LDS row;
//   2. each lane reads its C+2R window back from LDS and forms C horizontal
//      row sums (dx ascending, from zero);
//   3. row sums live in a register ring of 2R+1 entries (the loop is unrolled
//      over the ring period so every ring index is static) and each output is
//      the ordered sum of its 2R+1 row sums (dy ascending, from zero),
//      finished by the policy (mean: correctly rounded division by (2R+1)^2).
// This is exactly the operation order of the scalar definition (bit-exact)
// with 2(2R+1) adds per output instead of (2R+1)^2, one HBM read of each
// input element (halo rows/columns come from L2) and one write per output.
//
// Row access modes:
//   kAligned   - cols % C == 0 and 4C-byte aligned rows: a C-vector is wholly
//                inside or outside the grid, one load per vector (the vector
//                at c is inside exactly when 0 <= c and c + C <= cols);
//   kUnaligned - C = 4, any cols >= 4: the load is clamped into the row (a
//                dword-aligned 16-byte load) and each column picks its value
//                from the loaded quad by the clamping shift, so no lane
//                branches;
//   kScalar    - one clamped load per column (cols < C).
#pragma once

#include "common.hpp"

namespace mslhip {

enum RowMode { kAligned = 0, kUnaligned = 1, kScalar = 2 };

template <typename T, int C>
struct VecN {
    T v[C];
};

template <typename T, int C>
struct NativeVecN;
template <typename T>
struct NativeVecN<T, 1> {
    typedef T type;
};
template <>
struct NativeVecN<float, 4> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct NativeVecN<int32_t, 4> {
    typedef int32_t type __attribute__((ext_vector_type(4)));
};

template <typename T, int C>
__device__ __forceinline__ VecN<T, C> to_vecn(const typename NativeVecN<T, C>::type &x) {
    VecN<T, C> r;
    if constexpr (C == 1) {
        r.v[0] = x;
    } else {
        r.v[0] = x.x;
        r.v[1] = x.y;
        r.v[2] = x.z;
        r.v[3] = x.w;
    }
    return r;
}

template <typename T, int C>
__device__ __forceinline__ typename NativeVecN<T, C>::type to_native(const VecN<T, C> &q) {
    if constexpr (C == 1) {
        return q.v[0];
    } else {
        return typename NativeVecN<T, C>::type{q.v[0], q.v[1], q.v[2], q.v[3]};
    }
}

// Loads of C consecutive elements: p aligned to 4C bytes / to 4 bytes.
template <typename T, int C>
__device__ __forceinline__ VecN<T, C> load_aligned(const T *p) {
    return to_vecn<T, C>(*reinterpret_cast<const typename NativeVecN<T, C>::type *>(p));
}
template <typename T, int C>
__device__ __forceinline__ VecN<T, C> load_dword_aligned(const T *p) {
    typename NativeVecN<T, C>::type x;
    __builtin_memcpy(&x, p, sizeof(x));
    return to_vecn<T, C>(x);
}

// Reads of C columns [c, c+C) of row yy are split into the load itself
// (load_vec: the row and column are clamped so the address is always valid)
// and the fix-up (fix_vec: columns outside the grid become nv, and in
// kUnaligned mode each column picks its element of the clamped vector).
// The march issues the load D rows ahead and applies the fix-up only when
// the row is used, so the prefetched loads are never waited on early.
// The row base is wave-uniform and the in-row offset a 32-bit byte offset
// (rows of fewer than 2^30 cells), so a load takes a scalar base and one
// per-lane offset.
template <typename T, int C, int MODE>
__device__ __forceinline__ VecN<T, C> load_vec(const T *__restrict__ in, int yy, int c, int rows, int cols) {
    const int yc = min(max(yy, 0), rows - 1);
    const T *rowp = in + (size_t)yc * cols;
    if constexpr (MODE == kAligned || MODE == kUnaligned) {
        const unsigned off = (unsigned)min(max(c, 0), cols - C) * (unsigned)sizeof(T);
        const T *p = reinterpret_cast<const T *>(reinterpret_cast<const char *>(rowp) + off);
        if constexpr (MODE == kAligned)
            return load_aligned<T, C>(p);
        else
            return load_dword_aligned<T, C>(p);
    } else {
        VecN<T, C> q;
#pragma unroll
        for (int k = 0; k < C; k++) q.v[k] = rowp[min(max(c + k, 0), cols - 1)];
        return q;
    }
}

template <typename T, int C, int MODE>
__device__ __forceinline__ VecN<T, C> fix_vec(const VecN<T, C> &v, int yy, int c, int rows, int cols, T nv) {
    const bool rowok = yy >= 0 && yy < rows;
    VecN<T, C> q;
    if constexpr (MODE == kAligned) {
        const bool ok = rowok && c >= 0 && c + C <= cols;  // wholly inside or wholly outside
#pragma unroll
        for (int k = 0; k < C; k++) q.v[k] = ok ? v.v[k] : nv;
    } else if constexpr (MODE == kUnaligned) {
        // [cc, cc+C) lies inside the row; column c+k sits at index k+s of the
        // loaded vector whenever it is inside the grid.
        const int cc = min(max(c, 0), cols - C);
        const int s = c - cc;
#pragma unroll
        for (int k = 0; k < C; k++) {
            const int j = k + s;
            T pick = v.v[0];
#pragma unroll
            for (int jj = 1; jj < C; jj++) pick = (j == jj) ? v.v[jj] : pick;
            const int col = c + k;
            q.v[k] = (rowok && col >= 0 && col < cols) ? pick : nv;
        }
    } else {
#pragma unroll
        for (int k = 0; k < C; k++) {
            const int col = c + k;
            q.v[k] = (rowok && col >= 0 && col < cols) ? v.v[k] : nv;
        }
    }
    return q;
}

// Stores the lane's C outputs of row orow starting at column x0 (aligned
// rows: non-temporal stores when NT, see stream_stores).
template <typename T, int C, int MODE, bool NT>
__device__ __forceinline__ void write_vec(T *__restrict__ orow, int x0, int cols, const VecN<T, C> &o) {
    typedef typename NativeVecN<T, C>::type V;
    if constexpr (MODE == kAligned) {
        if (x0 < cols) {
            // scalar row base plus a 32-bit per-lane byte offset
            V *p = reinterpret_cast<V *>(reinterpret_cast<char *>(orow) + (unsigned)x0 * (unsigned)sizeof(T));
            if constexpr (NT)
                __builtin_nontemporal_store(to_native<T, C>(o), p);
            else
                *p = to_native<T, C>(o);
        }
    } else if constexpr (MODE == kUnaligned) {
        if (x0 + C <= cols) {
            const V x = to_native<T, C>(o);
            __builtin_memcpy(orow + x0, &x, sizeof(x));
        } else {
#pragma unroll
            for (int k = 0; k < C; k++)
                if (x0 + k < cols) orow[x0 + k] = o.v[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < C; k++)
            if (x0 + k < cols) orow[x0 + k] = o.v[k];
    }
}

// ---- per-stencil policies -------------------------------------------------
template <int R_>
struct MeanPolicy {
    static constexpr int R = R_;
    static constexpr bool kCenter = false;
    typedef float T;
    float denom;  // (2R+1)^2
    float rcp;    // RN(1 / denom)
    __device__ __host__ static MeanPolicy make() {
        MeanPolicy p;
        p.denom = (float)((2 * R + 1) * (2 * R + 1));
        p.rcp = 1.0f / p.denom;
        return p;
    }
    // Row sum of window w (index RP-R+c+d holds column x0+c+d-R), dx
    // ascending.  The definition starts the sum at +0; dropping that first
    // add changes only a window of all -0 values (-0 instead of +0), and
    // every row sum enters a column sum that itself starts at +0, where
    // -0 and +0 give the same result.  So the outputs are unchanged.
    template <int RP>
    __device__ __forceinline__ float rowsum(const float *w, int c) const {
        float s = w[RP - R + c];
#pragma unroll
        for (int d = 1; d <= 2 * R; d++) s = __fadd_rn(s, w[RP - R + c + d]);
        return s;
    }
    // Correctly rounded (0 + colsum) / denom (identical to IEEE division of
    // the definition's column sum, which starts at +0: the march sums the
    // row sums from the first one, and the +0 only matters for a sum of -0).
    __device__ __forceinline__ float finish(float colsum, float /*center*/) const {
        return div_plus0(colsum, denom, rcp);
    }
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
    __device__ __forceinline__ int32_t acc(int32_t a, int32_t b) const { return a + b; }
};

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

template <int R, int C>
struct StripConsts {
    static_assert(C == 1 || C == 4, "1 or 4 columns per lane");
    static constexpr int RP = (R + C - 1) / C * C;  // halo padded to whole vectors
    static constexpr int SW = 64 * C;               // strip width
    static constexpr int LDSW = SW + 2 * RP;        // elements per staged row
    static constexpr int HL = RP / C;               // halo vectors per side
    static constexpr int NQ = 2 * HL + 1;           // window vectors per lane
    static constexpr int PER = 2 * R + 1;           // rows in one output's window
    static_assert(2 * HL <= 64, "halo wider than a wavefront");
    // Rows requested ahead of use: larger where the per-row work is too
    // short to hide latency (see ring_slots for the unroll this implies).
    // With one column per lane the rows are narrow and the whole window is
    // kept in flight.
    static constexpr int D =
        C == 1 ? PER
               : (R == 1 ? 6 : (PER <= 7 ? PER : (R > 10 ? 1 : (PER % 5 == 0 ? 5 : (PER % 3 == 0 || HL >= 3 ? 3 : 1)))));
};

// Sum of the values one wave stores, kept as a compensated (s, c) pair per
// lane (TwoSum per element, in row order).
struct LaneSum {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void push(float x) {
        float t, e;
        two_sum(s, x, t, e);
        s = t;
        c = __fadd_rn(c, e);
    }
};

// Row-sum ring size for a window of PER rows and a prefetch depth of D rows:
// the march loop is unrolled over lcm(ring size, D) rows so that every ring
// and queue slot is a compile-time register index.  When neither of PER and
// D divides the other, the ring is rounded up to a multiple of D (a few spare
// slots instead of an unroll of PER * D rows, which would overflow the
// instruction cache).
constexpr int ring_slots(int PER, int D) {
    return (D % PER == 0 || PER % D == 0) ? PER : D * ((PER + D - 1) / D);
}

// March one wave down its (strip, segment): input rows y0-R .. yend-1+R,
// output rows y0 .. yend-1 of the strip starting at column xs.  One code
// path for every tile: reads are clamped (load_vec / fix_vec), stores are masked per
// lane, and the loop is unrolled over exactly one ring period so the code
// stays small (the instruction cache is shared by two CUs) while every
// register index is a compile-time constant.  Row t+D is requested when row t
// is staged (DQ > 0 overrides the default prefetch depth).  With kSum, every
// stored value is also pushed into *sum.
template <typename Pol, int C, int MODE, bool kSum, int DQ, bool UP, bool NT, int kTail>
__device__ __forceinline__ void strip_march_dir(typename Pol::T *__restrict__ out,
                                                const typename Pol::T *__restrict__ in, int rows, int cols,
                                                typename Pol::T nv, int xs, int y0, int yend, const Pol &pol,
                                                typename Pol::T (*lds)[StripConsts<Pol::R, C>::LDSW], LaneSum *sum) {
    typedef typename Pol::T T;
    typedef StripConsts<Pol::R, C> K;
    constexpr int R = Pol::R;
    constexpr int RP = K::RP;
    constexpr int NQ = K::NQ;
    constexpr int PER = K::PER;
    constexpr int D = DQ > 0 ? DQ : K::D;
    constexpr int HL = K::HL;
    // the loop body is unrolled over UL = lcm(P, D) rows so that both the
    // ring slot (t mod P) and the queue slot (t mod D) are static
    constexpr int P = ring_slots(PER, D);
    constexpr int UL = P * D / gcd_c(P, D);

    const int lane = threadIdx.x & 63;
    const int x0 = xs + C * lane;
    // input rows y0-R .. yend-1+R (a wave-uniform scalar when kTail ends the
    // march at it)
    const int total = kTail ? __builtin_amdgcn_readfirstlane((yend - y0) + 2 * R) : (yend - y0) + 2 * R;
    // input row of march step t: ascending from y0-R, or descending from yend-1+R
    const int ybase = UP ? yend - 1 + R : y0 - R;
    constexpr int ystep = UP ? -1 : 1;
    // Halo vectors: left ones for lanes [0, HL), right ones for lanes
    // [HL, 2HL); every other lane re-reads and re-stages its own centre
    // vector, so no lane diverges.
    const bool has_halo = lane < 2 * HL;
    const int hx = has_halo ? (lane < HL ? xs - RP + C * lane : xs + K::SW + C * (lane - HL)) : x0;
    const int hslot = has_halo ? (lane < HL ? C * lane : RP + K::SW + C * (lane - HL)) : RP + C * lane;

    VecN<T, C> ring[P];  // row sums of the last P rows (the last PER are used)
    VecN<T, C> cring[Pol::kCenter ? P : 1];
    VecN<T, C> qc[D], qh[D];  // raw loads of rows t .. t+D-1 (centre and halo vectors)
#pragma unroll
    for (int d = 0; d < D; d++) {
        qc[d] = load_vec<T, C, MODE>(in, ybase + ystep * d, x0, rows, cols);
        qh[d] = load_vec<T, C, MODE>(in, ybase + ystep * d, hx, rows, cols);
    }

    // One march step: stage input row t, request row t+D, form its row sums
    // and, once 2R+1 rows are in the ring, finish output row t-R.  uu = t mod UL
    // is a compile-time constant after unrolling, so every ring and queue
    // index is static.
    auto step = [&](const int t, const int uu) __attribute__((always_inline)) {
        const int u = uu % P;  // ring slot of row t (tb is a multiple of P)
        const int qs = uu % D;   // queue slot of row t (tb is a multiple of D)
        T *buf = lds[t & 1];
        typedef typename NativeVecN<T, C>::type V;
        const int yy = ybase + ystep * t;
        *reinterpret_cast<V *>(buf + RP + C * lane) =
            to_native<T, C>(fix_vec<T, C, MODE>(qc[qs], yy, x0, rows, cols, nv));
        *reinterpret_cast<V *>(buf + hslot) = to_native<T, C>(fix_vec<T, C, MODE>(qh[qs], yy, hx, rows, cols, nv));
        // request row t+D into the slot just staged
        qc[qs] = load_vec<T, C, MODE>(in, yy + ystep * D, x0, rows, cols);
        qh[qs] = load_vec<T, C, MODE>(in, yy + ystep * D, hx, rows, cols);
        wave_lds_sync();
        T w[C * NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const VecN<T, C> v = to_vecn<T, C>(*reinterpret_cast<const V *>(buf + C * lane + C * q));
#pragma unroll
            for (int k = 0; k < C; k++) w[C * q + k] = v.v[k];
        }
#pragma unroll
        for (int c = 0; c < C; c++) ring[u].v[c] = pol.template rowsum<RP>(w, c);
        if constexpr (Pol::kCenter) {
#pragma unroll
            for (int c = 0; c < C; c++) cring[u].v[c] = w[RP + c];
        }
        const int y = yy - ystep * R;  // output row completed by row t
        if (t >= 2 * R && (UP ? y >= y0 : y < yend)) {
            VecN<T, C> o;
#pragma unroll
            for (int c = 0; c < C; c++) {
                // rows y-R .. y+R in ascending order: the oldest ring entry
                // first when marching down, the newest first when marching
                // up (from the first row sum: see MeanPolicy::finish)
                T acc = ring[UP ? u : (u + P - (PER - 1)) % P].v[c];
#pragma unroll
                for (int k = 1; k < PER; k++)
                    acc = pol.acc(acc, ring[UP ? (u + P - k) % P : (u + P - (PER - 1) + k) % P].v[c]);
                T centre = T(0);
                if constexpr (Pol::kCenter) centre = cring[(u + P - R) % P].v[c];
                o.v[c] = pol.finish(acc, centre);
            }
            write_vec<T, C, MODE, NT>(out + (size_t)y * cols, x0, cols, o);
            if constexpr (kSum) {
#pragma unroll
                for (int c = 0; c < C; c++)
                    if (x0 + c < cols) sum->push(o.v[c]);
            }
        }
    };
    // kTail 0: whole ring periods; the steps past the march only load, stage
    // and sum rows that no output uses.  kTail 1: every step checks the end
    // (the batch and serving items, whose short segments would lose 6-30 % of
    // their work to those steps).  kTail 2: whole periods without checks, then
    // one checked period for the rest (long single-grid segments that do not
    // fill whole periods).
    int tb = 0;
    if constexpr (kTail == 2) {
        for (; tb + UL <= total; tb += UL) {
#pragma unroll
            for (int uu = 0; uu < UL; uu++) step(tb + uu, uu);
        }
        if (tb < total) {
#pragma unroll
            for (int uu = 0; uu < UL; uu++) {
                if (tb + uu >= total) break;
                step(tb + uu, uu);
            }
        }
    } else {
        for (; tb < total; tb += UL) {
#pragma unroll
            for (int uu = 0; uu < UL; uu++) {
                if constexpr (kTail == 1)
                    if (tb + uu >= total) break;
                step(tb + uu, uu);
            }
        }
    }
}


// March in the given direction (up != 0: from the last row upwards).  The
// result is the same either way; alternating the direction of vertically
// adjacent segments makes both readers of the 2R halo rows between them
// read those rows at the same time (both at the start or both at the end
// of their marches), so the second read is served by L2.
template <typename Pol, int C, int MODE, bool kSum = false, int DQ = 0, bool NT = true, int kTail = 0>
__device__ __forceinline__ void strip_march(typename Pol::T *__restrict__ out, const typename Pol::T *__restrict__ in,
                                            int rows, int cols, typename Pol::T nv, int xs, int y0, int yend,
                                            const Pol &pol, typename Pol::T (*lds)[StripConsts<Pol::R, C>::LDSW],
                                            LaneSum *sum = nullptr, int up = 0) {
    if (up)
        strip_march_dir<Pol, C, MODE, kSum, DQ, true, NT, kTail>(out, in, rows, cols, nv, xs, y0, yend, pol, lds, sum);
    else
        strip_march_dir<Pol, C, MODE, kSum, DQ, false, NT, kTail>(out, in, rows, cols, nv, xs, y0, yend, pol, lds, sum);
}

// Two generations of the radius-R mean blur in one march (four columns per
// lane, aligned rows).  The wave computes generation 1 on the 64 * 4 columns
// [xs - RP, xs - RP + 256) exactly as strip_march_dir does (same staging,
// same row sums, same division), turns every generation-1 value outside the
// grid into nv, stages each completed generation-1 row in a second LDS row
// and forms generation 2 from it by the same row-sum / column-sum order.
// Lanes HL .. 63-HL, columns [xs, xs + 256 - 2RP), store generation 2; the
// outer lanes only feed their neighbours' windows.  Input rows y0-2R ..
// yend-1+2R are read once; each cell of both generations is computed with
// the operation order of one generation per launch, so the result is
// bit-identical to two single-generation marches.  MODE: kAligned or
// kUnaligned rows (as in strip_march_dir).
// Rows requested ahead by the two-generation march: each row carries twice
// the work of a single-generation row, so deeper queues pay where the
// single-generation rule prefetches one row (measured on 8192^2: r = 5
// 5.1 -> 6.1 TB/s with 5 rows, r = 6 5.0 -> 5.8 with 2).
constexpr int strip2_depth(int R, int d1) { return R == 5 ? 5 : (R == 6 ? 2 : d1); }

// kSum: every stored value of generation 2 is also pushed into *sum (rows in
// march order, the lane's columns ascending).
template <int R, int MODE, bool UP, bool NT, bool kSum = false>
__device__ __forceinline__ void strip2_march_dir(float *__restrict__ out, const float *__restrict__ in, int rows,
                                                 int cols, float nv, int xs, int y0, int yend,
                                                 const MeanPolicy<R> &pol,
                                                 float (*lds)[StripConsts<R, 4>::LDSW], LaneSum *sum = nullptr) {
    constexpr int C = 4;
    typedef StripConsts<R, C> K;
    constexpr int RP = K::RP;
    constexpr int NQ = K::NQ;
    constexpr int PER = K::PER;
    constexpr int HL = K::HL;
    constexpr int D = strip2_depth(R, K::D);
    constexpr int P = ring_slots(PER, D);
    constexpr int UL = P * D / gcd_c(P, D);
    typedef NativeVecN<float, C>::type V;

    const int lane = threadIdx.x & 63;
    const int xg = xs - RP;         // first generation-1 column of the wave
    const int x1 = xg + C * lane;   // this lane's generation-1 / output columns
    const int total = (yend - y0) + 4 * R;  // input rows y0-2R .. yend-1+2R
    const int ybase = UP ? yend - 1 + 2 * R : y0 - 2 * R;
    constexpr int ystep = UP ? -1 : 1;
    const bool has_halo = lane < 2 * HL;
    const int hx = has_halo ? (lane < HL ? xg - RP + C * lane : xg + K::SW + C * (lane - HL)) : x1;
    const int hslot = has_halo ? (lane < HL ? C * lane : RP + K::SW + C * (lane - HL)) : RP + C * lane;
    bool col_in[C];  // generation-1 column inside the grid
#pragma unroll
    for (int c = 0; c < C; c++) col_in[c] = (unsigned)(x1 + c) < (unsigned)cols;
    const bool store_lane = lane >= HL && lane < 64 - HL && x1 < cols;

    VecN<float, C> ring1[P], ring2[P];
    VecN<float, C> qc[D], qh[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        qc[d] = load_vec<float, C, MODE>(in, ybase + ystep * d, x1, rows, cols);
        qh[d] = load_vec<float, C, MODE>(in, ybase + ystep * d, hx, rows, cols);
    }

    for (int tb = 0; tb < total; tb += UL) {
#pragma unroll
        for (int uu = 0; uu < UL; uu++) {
            const int t = tb + uu;
            const int u = uu % P;
            const int qs = uu % D;
            float *buf = lds[t & 1];
            const int yy = ybase + ystep * t;
            *reinterpret_cast<V *>(buf + RP + C * lane) =
                to_native<float, C>(fix_vec<float, C, MODE>(qc[qs], yy, x1, rows, cols, nv));
            *reinterpret_cast<V *>(buf + hslot) =
                to_native<float, C>(fix_vec<float, C, MODE>(qh[qs], yy, hx, rows, cols, nv));
            qc[qs] = load_vec<float, C, MODE>(in, yy + ystep * D, x1, rows, cols);
            qh[qs] = load_vec<float, C, MODE>(in, yy + ystep * D, hx, rows, cols);
            wave_lds_sync();
            {
                float w[C * NQ];
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const VecN<float, C> v = to_vecn<float, C>(*reinterpret_cast<const V *>(buf + C * lane + C * q));
#pragma unroll
                    for (int k = 0; k < C; k++) w[C * q + k] = v.v[k];
                }
#pragma unroll
                for (int c = 0; c < C; c++) ring1[u].v[c] = pol.template rowsum<RP>(w, c);
            }
            if (t < 2 * R) continue;
            // generation-1 row y1 is complete: finish it, stage it (nv outside the grid)
            const int y1 = yy - ystep * R;
            const bool row1 = y1 >= 0 && y1 < rows;
            VecN<float, C> v1;
#pragma unroll
            for (int c = 0; c < C; c++) {
                float acc = ring1[UP ? u : (u + P - (PER - 1)) % P].v[c];
#pragma unroll
                for (int k = 1; k < PER; k++)
                    acc = pol.acc(acc, ring1[UP ? (u + P - k) % P : (u + P - (PER - 1) + k) % P].v[c]);
                v1.v[c] = (row1 && col_in[c]) ? pol.finish(acc, 0.0f) : nv;
            }
            float *gbuf = lds[2 + (t & 1)];
            *reinterpret_cast<V *>(gbuf + RP + C * lane) = to_native<float, C>(v1);
            wave_lds_sync();
            {
                float w[C * NQ];
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const VecN<float, C> v = to_vecn<float, C>(*reinterpret_cast<const V *>(gbuf + C * lane + C * q));
#pragma unroll
                    for (int k = 0; k < C; k++) w[C * q + k] = v.v[k];
                }
#pragma unroll
                for (int c = 0; c < C; c++) ring2[u].v[c] = pol.template rowsum<RP>(w, c);
            }
            if (t < 4 * R) continue;
            // output row y2: in [y0, yend) for 4R <= t < total; the unrolled
            // steps past the end of the march (t >= total) store nothing
            const int y2 = y1 - ystep * R;
            VecN<float, C> o;
#pragma unroll
            for (int c = 0; c < C; c++) {
                float acc = ring2[UP ? u : (u + P - (PER - 1)) % P].v[c];
#pragma unroll
                for (int k = 1; k < PER; k++)
                    acc = pol.acc(acc, ring2[UP ? (u + P - k) % P : (u + P - (PER - 1) + k) % P].v[c]);
                o.v[c] = pol.finish(acc, 0.0f);
            }
            if (store_lane && (UP ? y2 >= y0 : y2 < yend)) {
                write_vec<float, C, MODE, NT>(out + (size_t)y2 * cols, x1, cols, o);
                if constexpr (kSum) {
#pragma unroll
                    for (int c = 0; c < C; c++)
                        if (x1 + c < cols) sum->push(o.v[c]);
                }
            }
        }
    }
}

}  // namespace mslhip

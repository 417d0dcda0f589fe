// Work items of batched 2-D mean blurs (the batch engine and the serving
// engine): radius and row-mode specialisations of the strip march, chosen at
// run time by wave-uniform branches.
#pragma once

#include "strip.hpp"

namespace mslhip {

constexpr int kBatchMaxR = 10;
constexpr int kBatchMaxR4 = 8;  // largest radius run with 4 columns per lane
constexpr int kWidthClasses = 4;
constexpr int kKindAligned4 = 0;    // C = 4, aligned rows
constexpr int kKindUnaligned4 = 1;  // C = 4, dword-aligned rows
constexpr int kKindNarrow1 = 2;     // C = 1

// Prefetch depth per (R, C) in the batched kernels: three rows beyond
// R = 2 (the row-sum ring is rounded up to a multiple of 3 where 2R+1 is
// not one), which keeps the kernel holding every specialisation within 168
// VGPRs (three 4-wave workgroups per CU); narrow strips keep the whole
// window in flight.
template <int R, int C>
struct BatchDepth {
    static constexpr int PER = 2 * R + 1;
    static constexpr int D = C == 1 ? PER : (PER <= 5 ? PER : 3);
};

inline int width_class(int cols) {
    if (cols < 96) return 0;
    if (cols < 192) return 1;
    if (cols < 384) return 2;
    return 3;
}

// One work item: rows [y0, y1) of the strip starting at column xs of a
// rows x cols grid, read from src and written to dst, with the row mode of
// `kind`, marching upwards when `up`; radii above kBatchMaxR4 always run
// with one column per lane.
template <int R, bool kFold>
__device__ __forceinline__ void blur_item(const float *src, float *dst, int rows, int cols, float nv, int kind, int xs,
                                          int y0, int y1, float *ldsw, LaneSum *ls, int up) {
    const MeanPolicy<R> pol = MeanPolicy<R>::make();
    if constexpr (R > kBatchMaxR4) {
        strip_march<MeanPolicy<R>, 1, kAligned, kFold, BatchDepth<R, 1>::D, true, true>(
            dst, src, rows, cols, nv, xs, y0, y1, pol, reinterpret_cast<float(*)[StripConsts<R, 1>::LDSW]>(ldsw), ls, up);
    } else if (kind == kKindAligned4) {
        strip_march<MeanPolicy<R>, 4, kAligned, kFold, BatchDepth<R, 4>::D, true, true>(
            dst, src, rows, cols, nv, xs, y0, y1, pol, reinterpret_cast<float(*)[StripConsts<R, 4>::LDSW]>(ldsw), ls, up);
    } else if (kind == kKindUnaligned4) {
        strip_march<MeanPolicy<R>, 4, kUnaligned, kFold, BatchDepth<R, 4>::D, true, true>(
            dst, src, rows, cols, nv, xs, y0, y1, pol, reinterpret_cast<float(*)[StripConsts<R, 4>::LDSW]>(ldsw), ls, up);
    } else {
        strip_march<MeanPolicy<R>, 1, kAligned, kFold, BatchDepth<R, 1>::D, true, true>(
            dst, src, rows, cols, nv, xs, y0, y1, pol, reinterpret_cast<float(*)[StripConsts<R, 1>::LDSW]>(ldsw), ls, up);
    }
}

// Radius switch (r is wave-uniform, so the branch is a scalar jump).
template <bool kFold>
__device__ __forceinline__ void blur_item_r(int r, const float *src, float *dst, int rows, int cols, float nv,
                                            int kind, int xs, int y0, int y1, float *ldsw, LaneSum *ls, int up) {
    switch (r) {
        case 1: blur_item<1, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 2: blur_item<2, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 3: blur_item<3, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 4: blur_item<4, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 5: blur_item<5, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 6: blur_item<6, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 7: blur_item<7, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 8: blur_item<8, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        case 9: blur_item<9, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
        default: blur_item<10, kFold>(src, dst, rows, cols, nv, kind, xs, y0, y1, ldsw, ls, up); break;
    }
}

// Combines the lanes' compensated sums in a fixed butterfly (every lane ends
// with the same pair).
__device__ __forceinline__ void wave_pair_reduce(LaneSum &ls) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const float os = __shfl_xor(ls.s, m), oc = __shfl_xor(ls.c, m);
        float s, e;
        two_sum(ls.s, os, s, e);
        ls.c = __fadd_rn(__fadd_rn(ls.c, oc), e);
        ls.s = s;
    }
}

constexpr int kBatchLdsw = StripConsts<kBatchMaxR4, 4>::LDSW > StripConsts<kBatchMaxR, 1>::LDSW
                                 ? StripConsts<kBatchMaxR4, 4>::LDSW
                                 : StripConsts<kBatchMaxR, 1>::LDSW;


}  // namespace mslhip

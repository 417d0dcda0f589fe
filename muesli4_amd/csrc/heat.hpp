// 7-point heat march of one wavefront over a 64 x 8 column tile and a
// segment of z-planes, shared by the single-cube kernel and the batched
// serving kernel.
//
// Lane l owns column x = x0 + l and keeps, for a ring of kP planes, the kTY
// cells of its column plus the two y-halo cells (rows y0-1 and y0+kTY) and
// one x-halo cell per row (lane 0 reads x0-1, lane 63 reads x0+64, every
// other lane re-reads its own cell).  x neighbours come from the adjacent
// lanes by DPP, y neighbours from the lane's own registers, z neighbours from
// the plane ring; each plane is requested one step before its use.  Each cell
//   v = 0 + in(x-1) + in(x+1) + in(y-1) + in(y+1) + in(z-1) + in(z+1); v / 6
// is computed in exactly that order (examples/heatdiffusion.cpp:55-65).
#pragma once

#include "strip.hpp"

namespace mslhip {

constexpr int kTX = 64;  // x extent of a wave tile (one cell per lane)
constexpr int kTY = 8;   // y extent of a wave tile (cells per lane)
constexpr int kP = 4;    // plane ring: z-1, z, z+1 and the requested z+2

// Reads are branch-free: addresses are clamped into the cube so no load is
// predicated, and the neutral value replaces the cells outside it when a
// plane is first used, not when it is requested, so requested planes stay
// in flight.  In-plane byte offsets are 32-bit (planes of fewer than 2^30
// cells).

// Whole-wave lane shifts (DPP wave_shr:1 / wave_shl:1): lane i receives lane
// i-1 (resp. i+1); the lane whose source is outside the wave keeps 'edge'.
__device__ __forceinline__ float from_left(float v, float edge) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float v, float edge) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// March the tile (x0 = 64 * xtile, rows y0 .. y0+kTY-1) over planes
// z0 .. z1-1.  kResidual: fold max |v - in| into *dmax.  kSum: push every
// stored value into *sum.
template <bool kResidual, bool kSum, bool NT = true>
__device__ __forceinline__ void heat7_march(float *__restrict__ out, const float *__restrict__ in, int nx, int ny,
                                            int nz, float nv, int x0, int y0, int z0, int z1, float *dmax,
                                            LaneSum *sum) {
    const int lane = threadIdx.x & 63;
    const int x = x0 + lane;
    const int hxl = lane == 0 ? x - 1 : (lane == 63 ? x + 1 : x);  // x-halo column

    // A lane's cells sit at the same (x, y) in every plane: their in-plane
    // byte offsets (clamped into the plane, so loads are never predicated)
    // and in-plane validity are computed once; per plane only the
    // wave-uniform plane base and z validity change, so a load is one
    // instruction with a scalar base and the lane's 32-bit offset.
    const int xc = min(max(x, 0), nx - 1), hxc = min(max(hxl, 0), nx - 1);
    const bool xok = (unsigned)x < (unsigned)nx, hxok = (unsigned)hxl < (unsigned)nx;
    unsigned coff[kTY + 2], hoff[kTY];
    unsigned cok = 0, hok = 0;  // bit j: row j of col / hal lies in the plane
#pragma unroll
    for (int j = 0; j < kTY + 2; j++) {
        const int y = y0 - 1 + j;
        coff[j] = (unsigned)(min(max(y, 0), ny - 1) * nx + xc) * (unsigned)sizeof(float);
        cok |= (xok && (unsigned)y < (unsigned)ny) ? 1u << j : 0u;
    }
#pragma unroll
    for (int j = 0; j < kTY; j++) {
        const int y = y0 + j;
        hoff[j] = (unsigned)(min(max(y, 0), ny - 1) * nx + hxc) * (unsigned)sizeof(float);
        hok |= (hxok && (unsigned)y < (unsigned)ny) ? 1u << j : 0u;
    }
    const size_t plane = (size_t)nx * ny;

    // slot(zz) = (zz - z0 + 1) & (kP-1); col[j] is row y0-1+j
    float col[kTY + 2][kP];
    float hal[kTY][kP];
    auto load_plane = [&](int zz, int slot) {
        const char *p = reinterpret_cast<const char *>(in + (size_t)min(max(zz, 0), nz - 1) * plane);
#pragma unroll
        for (int j = 0; j < kTY + 2; j++) col[j][slot] = *reinterpret_cast<const float *>(p + coff[j]);
#pragma unroll
        for (int j = 0; j < kTY; j++) hal[j][slot] = *reinterpret_cast<const float *>(p + hoff[j]);
    };
    auto fix_plane = [&](int zz, int slot) {
        const bool zok = (unsigned)zz < (unsigned)nz;  // wave-uniform
#pragma unroll
        for (int j = 0; j < kTY + 2; j++) col[j][slot] = (zok && ((cok >> j) & 1u)) ? col[j][slot] : nv;
#pragma unroll
        for (int j = 0; j < kTY; j++) hal[j][slot] = (zok && ((hok >> j) & 1u)) ? hal[j][slot] : nv;
    };
#pragma unroll
    for (int k = 0; k < kP - 1; k++) load_plane(z0 - 1 + k, k);
    fix_plane(z0 - 1, 0);
    fix_plane(z0, 1);

    // Whole blocks of kP planes; planes past z1 are computed, not stored.
    for (int zb = z0; zb < z1; zb += kP) {
#pragma unroll
        for (int u = 0; u < kP; u++) {
            const int z = zb + u;
            const int sb = u & (kP - 1);        // z-1
            const int sc = (u + 1) & (kP - 1);  // z
            const int sa = (u + 2) & (kP - 1);  // z+1
            // request plane z+2 into the slot of z-2 (free since step z-1);
            // plane z+1 is used for the first time now
            load_plane(z + 2, (u + 3) & (kP - 1));
            fix_plane(z + 1, sa);
            char *op = reinterpret_cast<char *>(out + (size_t)z * plane);
#pragma unroll
            for (int j = 0; j < kTY; j++) {
                const float c = col[j + 1][sc];
                // the sum without its leading "0 +", which div6_plus0 restores
                float v = from_left(c, hal[j][sc]);           // x-1
                v = __fadd_rn(v, from_right(c, hal[j][sc]));  // x+1
                v = __fadd_rn(v, col[j][sc]);                 // y-1
                v = __fadd_rn(v, col[j + 2][sc]);             // y+1
                v = __fadd_rn(v, col[j + 1][sb]);             // z-1
                v = __fadd_rn(v, col[j + 1][sa]);             // z+1
                v = div6_plus0(v);                            // (0 + v) / 6, correctly rounded
                // row j+1 of col is output row y0+j: in the plane exactly
                // when its cell is, and then coff is its unclamped offset
                if (z < z1 && ((cok >> (j + 1)) & 1u)) {
                    float *o = reinterpret_cast<float *>(op + coff[j + 1]);
                    if constexpr (NT)
                        __builtin_nontemporal_store(v, o);
                    else
                        *o = v;
                    if constexpr (kResidual) {
                        const float d = fabsf(__fsub_rn(v, c));
                        *dmax = d > *dmax ? d : *dmax;
                    }
                    if constexpr (kSum) sum->push(v);
                }
            }
        }
    }
}

// Two generations per pass.  The wave's 64 lanes cover columns x0-2 ..
// x0+61 and it keeps input rows y0-2 .. y0+kT2Y+1; generation 1 is computed
// on rows y0-1 .. y0+kT2Y (exact on lanes 1 .. 62), generation 2 from it on
// rows y0 .. y0+kT2Y-1, exact on lanes 2 .. 61 (columns x0 .. x0+59), which
// are stored: the two edge lanes on each side serve as x-halo, so no extra
// halo loads are needed.  Along z generation 1 of plane z is formed first,
// then generation 2 of plane z-1 from the ring of generation-1 planes.
// Generation-1 cells outside the cube are nv, as the second step reads them.
// Every cell is computed in the one-step order, so the result is
// bit-identical to two single steps.  kResidual: max |generation 2 -
// generation 1| into *dmax.
constexpr int kT2X = kTX - 4;  // output columns of a two-generation wave tile
// Output rows of a two-generation wave tile: 5 rows (90 VGPRs, five waves
// per SIMD) beat 8 (124, four) although the y halo grows from 4/8 to 4/5
// (512^3, 200 iterations: 8 rows 32.5 ms, 7 32.6, 6 31.8, 5 30.5, 4 35.1;
// 320^3: 8 rows 7.13 ms, 5 6.90).
constexpr int kT2Y = 5;

template <bool kResidual, bool NT = true>
__device__ __forceinline__ void heat7x2_march(float *__restrict__ out, const float *__restrict__ in, int nx, int ny,
                                              int nz, float nv, int x0, int y0, int z0, int z1, float *dmax) {
    constexpr int NI = kT2Y + 4;  // input rows y0-2 .. y0+kT2Y+1
    constexpr int N1 = kT2Y + 2;  // generation-1 rows y0-1 .. y0+kT2Y
    const int lane = threadIdx.x & 63;
    const int x = x0 - 2 + lane;
    const int xc = min(max(x, 0), nx - 1);
    const bool xok = (unsigned)x < (unsigned)nx;
    unsigned ioff[NI];
    unsigned iok = 0;  // bit j: input row j lies in the plane
#pragma unroll
    for (int j = 0; j < NI; j++) {
        const int y = y0 - 2 + j;
        ioff[j] = (unsigned)(min(max(y, 0), ny - 1) * nx + xc) * (unsigned)sizeof(float);
        iok |= (xok && (unsigned)y < (unsigned)ny) ? 1u << j : 0u;
    }
    const size_t plane = (size_t)nx * ny;

    // input slot(zz) = (zz - z0 + 2) & 3; generation-1 slot(zz) = (zz - z0 + 1) & 3
    float inp[NI][kP], g1[N1][kP];
    auto load_plane = [&](int zz, int slot) {
        const char *p = reinterpret_cast<const char *>(in + (size_t)min(max(zz, 0), nz - 1) * plane);
#pragma unroll
        for (int j = 0; j < NI; j++) inp[j][slot] = *reinterpret_cast<const float *>(p + ioff[j]);
    };
    auto fix_plane = [&](int zz, int slot) {
        const bool zok = (unsigned)zz < (unsigned)nz;
#pragma unroll
        for (int j = 0; j < NI; j++) inp[j][slot] = (zok && ((iok >> j) & 1u)) ? inp[j][slot] : nv;
    };
    // one step of the 7-point rule on a lane's column: rows r-1, r, r+1 of
    // plane slots (below, centre, above), x neighbours from the adjacent lanes
    auto cell = [&](float(&a)[kP], float(&m)[kP], float(&b)[kP], float(&zl)[kP], float(&zh)[kP], int sb, int sc,
                    int sa) -> float {
        (void)sb;
        (void)sa;
        const float c = m[sc];
        float v = from_left(c, c);           // x-1 (the leading "0 +" is in div6_plus0)
        v = __fadd_rn(v, from_right(c, c));  // x+1
        v = __fadd_rn(v, a[sc]);             // y-1
        v = __fadd_rn(v, b[sc]);             // y+1
        v = __fadd_rn(v, zl[sb]);            // z-1
        v = __fadd_rn(v, zh[sa]);            // z+1
        return div6_plus0(v);
    };
    // generation 1 of plane zi from input slots (zi-1, zi, zi+1) into g1 slot d
    auto gen1 = [&](int zi, int sb, int sc, int sa, int d) {
        const bool zok = (unsigned)zi < (unsigned)nz;
#pragma unroll
        for (int j = 0; j < N1; j++) {
            const float v = cell(inp[j], inp[j + 1], inp[j + 2], inp[j + 1], inp[j + 1], sb, sc, sa);
            g1[j][d] = (zok && ((iok >> (j + 1)) & 1u)) ? v : nv;
        }
    };

#pragma unroll
    for (int j = 0; j < N1; j++) g1[j][3] = nv;  // read (unstored) by the first step
    load_plane(z0 - 2, 0);
    load_plane(z0 - 1, 1);
    load_plane(z0, 2);
    load_plane(z0 + 1, 3);
    fix_plane(z0 - 2, 0);
    fix_plane(z0 - 1, 1);
    fix_plane(z0, 2);
    gen1(z0 - 1, 0, 1, 2, 0);

    // step zi: generation 1 of plane zi, then generation 2 of plane zi-1
    for (int zb = z0; zb <= z1; zb += kP) {
#pragma unroll
        for (int u = 0; u < kP; u++) {
            const int zi = zb + u;
            load_plane(zi + 2, u & 3);
            fix_plane(zi + 1, (u + 3) & 3);
            gen1(zi, (u + 1) & 3, (u + 2) & 3, (u + 3) & 3, (u + 1) & 3);
            const int zo = zi - 1;
            const int tb = (u + 3) & 3, tc = u & 3, ta = (u + 1) & 3;  // g1 slots of zo-1, zo, zo+1
            char *op = reinterpret_cast<char *>(out + (size_t)min(max(zo, 0), nz - 1) * plane);
#pragma unroll
            for (int j = 0; j < kT2Y; j++) {
                const float v = cell(g1[j], g1[j + 1], g1[j + 2], g1[j + 1], g1[j + 1], tb, tc, ta);
                // output row y0+j is input row j+2: stored when in the cube
                if (zo >= z0 && zo < z1 && lane >= 2 && lane <= 61 && ((iok >> (j + 2)) & 1u)) {
                    float *o = reinterpret_cast<float *>(op + ioff[j + 2]);
                    if constexpr (NT)
                        __builtin_nontemporal_store(v, o);
                    else
                        *o = v;
                    if constexpr (kResidual) {
                        const float dd = fabsf(__fsub_rn(v, g1[j + 1][tc]));
                        *dmax = dd > *dmax ? dd : *dmax;
                    }
                }
            }
        }
    }
}

}  // namespace mslhip

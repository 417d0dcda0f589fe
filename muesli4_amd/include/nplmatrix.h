/*
 * nplmatrix.h -- read-only neighbourhood view handed to DM::mapStencil
 * functions, with the reference's name and call interface
 * (include/nplmatrix.h:56-168): operator()(x, y) with x the column and y the
 * row returns the neutral value outside [0,width) x [0,height).
 *
 * On the device the view first looks in the workgroup's LDS window (the
 * 2r+1 rows around the output row, plus r halo columns each side, staged
 * with nv outside the grid; row y lives in ring slot y & rowMask), and falls
 * back to the neutral value or global memory outside the window, so functors
 * that read beyond the declared radius still get correct values.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "muesli.h"

namespace msl {

template <typename T>
class NPLMatrix {
public:
    int width = 0;        // columns
    int height = 0;       // rows
    int stencilSize = 0;  // declared radius
    T neutralValue{};
    const T *data = nullptr;  // device grid, row-major
    // LDS window: columns [tileX, tileX+tileW), rows [tileY, tileY+tileH);
    // row y is stored at tile + (y & rowMask) * tileW
    const T *tile = nullptr;
    int tileX = 0, tileY = 0, tileW = 0, tileH = 0, rowMask = 0;

    MSL_USERFUNC NPLMatrix() {}

    MSL_USERFUNC T operator()(int x, int y) const {
        const int lx = x - tileX, ly = y - tileY;
        if ((unsigned)lx < (unsigned)tileW && (unsigned)ly < (unsigned)tileH) return detail::lds_read(tile, (y & rowMask) * tileW + lx);
        if ((unsigned)x >= (unsigned)width || (unsigned)y >= (unsigned)height) return neutralValue;
        return data[(size_t)y * width + x];
    }

    MSL_USERFUNC int coordinateToIndex(int x, int y) const { return y * width + x; }

    MSL_USERFUNC int2 indexToCoordinate(int i) const {
        int x = i % width;
        i /= width;
        int y = i % height;
        return int2{x, y};
    }
};

}  // namespace msl

/*
 * plcube.h -- read-only neighbourhood view handed to DC::mapStencil
 * functions, with the reference's name and call interface
 * (include/plcube.h:56-145): operator()(x, y, z) returns the neutral value
 * outside the cube; layout (z*height + y)*width + x.
 *
 * On the device reads come from the workgroup's LDS tile when the cell lies
 * inside it, else from global memory.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "muesli.h"

namespace msl {

template <typename T>
class PLCube {
public:
    int width = 0;   // columns (x)
    int height = 0;  // rows (y)
    int depth = 0;   // z
    int stencilSize = 0;
    T neutralValue{};
    const T *data = nullptr;
    const T *tile = nullptr;  // LDS tile, pitches tileW and tileW*tileH
    int tileX = 0, tileY = 0, tileZ = 0, tileW = 0, tileH = 0, tileD = 0;

    MSL_USERFUNC PLCube() {}

    MSL_USERFUNC const T &operator()(int x, int y, int z) const {
        if ((unsigned)x >= (unsigned)width || (unsigned)y >= (unsigned)height || (unsigned)z >= (unsigned)depth)
            return neutralValue;
        const int lx = x - tileX, ly = y - tileY, lz = z - tileZ;
        if ((unsigned)lx < (unsigned)tileW && (unsigned)ly < (unsigned)tileH && (unsigned)lz < (unsigned)tileD)
            return tile[(lz * tileH + ly) * tileW + lx];
        return data[((size_t)z * height + y) * width + x];
    }

    MSL_USERFUNC int coordinateToIndex(int x, int y, int z) const { return (z * height + y) * width + x; }

    MSL_USERFUNC int3 indexToCoordinate(int i) const {
        int x = i % width;
        i /= width;
        int y = i % height;
        int z = i / height;
        return int3{x, y, z};
    }
};

}  // namespace msl

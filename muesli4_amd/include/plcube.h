/*
 * plcube.h -- read-only neighbourhood view handed to DC::mapStencil
 * functions, with the reference's name and call interface
 * (include/plcube.h:56-145): operator()(x, y, z) returns the neutral value
 * outside the cube; layout (z*height + y)*width + x.
 *
 * On the device reads come from the workgroup's LDS window (the 2r+1 planes
 * around the output plane, each a tile plus halo staged with nv outside the
 * cube; plane z lives in ring slot z & planeMask) when the cell lies inside
 * it, else the neutral value or global memory.
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
    // LDS window: x in [tileX, tileX+tileW), y in [tileY, tileY+tileH),
    // z in [tileZ, tileZ+tileD); plane z at tile + (z & planeMask)*tileW*tileH
    const T *tile = nullptr;
    int tileX = 0, tileY = 0, tileZ = 0, tileW = 0, tileH = 0, tileD = 0, planeMask = 0;

    MSL_USERFUNC PLCube() {}

    MSL_USERFUNC T operator()(int x, int y, int z) const {
        const int lx = x - tileX, ly = y - tileY, lz = z - tileZ;
        if ((unsigned)lx < (unsigned)tileW && (unsigned)ly < (unsigned)tileH && (unsigned)lz < (unsigned)tileD)
            return detail::lds_read(tile, ((z & planeMask) * tileH + ly) * tileW + lx);
        if ((unsigned)x >= (unsigned)width || (unsigned)y >= (unsigned)height || (unsigned)z >= (unsigned)depth)
            return neutralValue;
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

/*
 * plmatrix.h -- neighbourhood view of the legacy DM::mapStencilMM skeleton,
 * with the reference's name and call interface (include/plmatrix.h:50-181):
 * get(row, col).
 *
 * Cells outside the grid read as the neutral value on all four sides; when
 * the skeleton was given a neutral-value functor, the cells of the border
 * band of width 'border' take the functor's value at that (row, col),
 * evaluated once per call into a padded border buffer.  (The reference reads
 * wrap-around or uninitialised values at the left and right edges.)
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "muesli.h"

namespace msl {

template <typename T>
class PLMatrix {
public:
    int rows = 0, cols = 0;
    int stencil_size = 0;
    T neutral_value{};
    const T *data = nullptr;        // device grid, row-major
    const T *border_data = nullptr;  // optional (rows+2b) x (cols+2b) padded values
    int border = 0;
    // LDS window: columns [tileCol, tileCol+tileCols), rows [tileRow,
    // tileRow+tileRows), staged with the values get() would return (grid,
    // border band or neutral value); row r at tile + (r & rowMask)*tileCols
    const T *tile = nullptr;
    int tileRow = 0, tileCol = 0, tileRows = 0, tileCols = 0, rowMask = 0;

    MSL_USERFUNC PLMatrix() {}

    // Value of a cell outside the grid: the border band (if any), else nv.
    MSL_USERFUNC T outside(int row, int col) const {
        if (border_data && row >= -border && row < rows + border && col >= -border && col < cols + border)
            return border_data[(size_t)(row + border) * (cols + 2 * border) + (col + border)];
        return neutral_value;
    }

    MSL_USERFUNC T get(int row, int col) const {
        const int lr = row - tileRow, lc = col - tileCol;
        if ((unsigned)lr < (unsigned)tileRows && (unsigned)lc < (unsigned)tileCols)
            return detail::lds_read(tile, (row & rowMask) * tileCols + lc);
        if ((unsigned)row >= (unsigned)rows || (unsigned)col >= (unsigned)cols) return outside(row, col);
        return data[(size_t)row * cols + col];
    }

    MSL_USERFUNC int getRows() const { return rows; }
    MSL_USERFUNC int getCols() const { return cols; }
    MSL_USERFUNC int getStencilSize() const { return stencil_size; }
};

}  // namespace msl

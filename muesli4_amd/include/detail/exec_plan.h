/*
 * exec_plan.h -- descriptor of a container's device slice, with the
 * reference's name and fields (include/detail/exec_plan.h:35-86).  With one
 * process and one GPU the slice is the whole container.
 */
#pragma once

#include <cstddef>

namespace msl {

template <typename T>
struct GPUExecutionPlan {
    int size = 0;
    int nLocal = 0;
    int first = 0;
    size_t bytes = 0;
    int firstRow = 0, firstCol = 0, gpuRows = 0, gpuCols = 0, gpuDepth = 0;
    T *d_Data = nullptr;
    T *h_Data = nullptr;
};

}  // namespace msl

/*
 * array.h -- fixed-size value array for user functors, with the reference's
 * name and interface (include/array.h:4-18): array<T, N> with data[N],
 * size() and operator[] usable on the host and the device.  The ACO
 * examples use it for cities and ant states.
 */
#pragma once

#include <cstddef>

#include "muesli.h"

template <typename T, unsigned int N>
struct array {
    T data[N];

    [[nodiscard]] int size() const { return (int)N; }
    MSL_USERFUNC T operator[](size_t i) const { return data[i]; }
    MSL_USERFUNC T &operator[](size_t i) { return data[i]; }
};

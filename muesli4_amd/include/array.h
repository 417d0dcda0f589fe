/*
 * array.h -- fixed-size value array used by example programs (name and
 * interface as in the reference include/array.h).
 */
#pragma once

#include <cstddef>

#include "muesli.h"

template <typename T, unsigned int N>
struct array {
    T data[N];
    int size() const { return N; }
    MSL_USERFUNC T operator[](size_t n) const { return data[n]; }
    MSL_USERFUNC T &operator[](size_t n) { return data[n]; }
};

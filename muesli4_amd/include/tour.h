/*
 * tour.h -- a tour of N stops plus its length, with the reference's name and
 * interface (include/tour.h:4-24), used by examples/aco_tsp_dataset.cpp.
 */
#pragma once

#include <cstddef>
#include <ostream>

#include "muesli.h"

template <typename T, unsigned int N>
struct tour {
    T data[N];
    double distance{};

    MSL_USERFUNC T operator[](size_t i) const { return data[i]; }
    MSL_USERFUNC T &operator[](size_t i) { return data[i]; }
    MSL_USERFUNC void setDist(double d) { distance = d; }
    MSL_USERFUNC double getDist() { return distance; }
    // prints the first two stops, as the reference does
    std::ostream &operator<<(std::ostream &os) {
        os << "(" << data[0] << ", " << data[1] << ")";
        return os;
    }
};

/*
 * vec3.h -- small 3-vector used by example programs (name and operators as
 * in the reference include/vec3.h so that those programs compile).
 */
#pragma once

#include "muesli.h"

template <typename T>
struct vec3 {
    T x, y, z;
    MSL_USERFUNC void operator+=(const vec3<T> &o) {
        x += o.x;
        y += o.y;
        z += o.z;
    }
    MSL_USERFUNC friend vec3<T> operator+(vec3<T> v, const vec3<T> &o) {
        v += o;
        return v;
    }
    MSL_USERFUNC void operator*=(const T &s) {
        x *= s;
        y *= s;
        z *= s;
    }
    MSL_USERFUNC friend vec3<T> operator*(vec3<T> v, const T &s) {
        v *= s;
        return v;
    }
    MSL_USERFUNC friend T operator*(const vec3<T> &a, const vec3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
};

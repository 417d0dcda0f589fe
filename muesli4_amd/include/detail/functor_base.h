/*
 * functor_base.h -- host-side knobs carried by stencil functors (tile width,
 * argument list), with the names of the reference
 * (include/detail/functor_base.h:52-203).  The device kernels choose their
 * own tile shapes; these values are stored and reported only.
 */
#pragma once

#include <vector>

#include "../argtype.h"

namespace msl {
namespace detail {

class FunctorBase {
public:
    FunctorBase() : tile_width(-1), local_indices(false) {}
    bool useLocalIndices() const { return local_indices; }
    void setLocalIndices(bool value) { local_indices = value; }
    MSL_USERFUNC int getTileWidth() const { return tile_width; }
    void setTileWidth(int value) {
        for (ArgumentType *arg : args) arg->setTileWidth(value);
        tile_width = value;
    }
    void notify() {
        for (ArgumentType *arg : args) arg->update();
    }
    void addArgument(ArgumentType *arg) {
        arg->setTileWidth(tile_width);
        args.push_back(arg);
    }
    int getSmemSize() const {
        int size = 0;
        for (ArgumentType *arg : args) size += arg->getSmemSize();
        return size;
    }
    virtual ~FunctorBase() {}

protected:
    int tile_width;
    bool local_indices;
    std::vector<ArgumentType *> args;
};

class MatrixFunctorBase : public FunctorBase {
public:
    void init(int nl, int ml, int fr, int fc) {
        nLocal = nl;
        mLocal = ml;
        firstRow = fr;
        firstCol = fc;
    }

protected:
    int nLocal = 0, mLocal = 0, firstRow = 0, firstCol = 0;
};

class ArrayFunctorBase : public FunctorBase {
public:
    void init(int nl, int f) {
        nLocal = nl;
        first = f;
    }

protected:
    int nLocal = 0, first = 0;
};

}  // namespace detail
}  // namespace msl

/*
 * functors.h -- the user-functor ("plugin") surface of the skeletons, with
 * the class and alias names of the reference (include/functors.h:57-410).
 *
 * Functor..Functor6 declare a virtual MSL_USERFUNC operator(); the skeleton
 * kernels call the derived class's operator with a qualified name, so no
 * virtual dispatch (and no host vtable pointer) is used on the device.
 * Stencil functions are plain MSL_USERFUNC functions passed as template
 * arguments: NPLMMapStencilFunctor<T> for DM::mapStencil and
 * DCMapStencilFunctor<T> for DC::mapStencil.
 */
#pragma once

#include <type_traits>
#include <utility>

#include "muesli.h"
#include "argtype.h"
#include "detail/functor_base.h"
#include "nplmatrix.h"
#include "plcube.h"
#include "plmatrix.h"

namespace msl {

template <typename T, typename R>
class Functor {
public:
    MSL_USERFUNC virtual R operator()(T value) const = 0;
    virtual ~Functor() {}
};

template <typename T1, typename T2, typename R>
class Functor2 {
public:
    MSL_USERFUNC virtual R operator()(T1 x, T2 y) const = 0;
    virtual ~Functor2() {}
};

template <typename T1, typename T2, typename T3, typename R>
class Functor3 {
public:
    MSL_USERFUNC virtual R operator()(T1 x, T2 y, T3 z) const = 0;
    virtual ~Functor3() {}
};

template <typename T1, typename T2, typename T3, typename T4, typename R>
class Functor4 {
public:
    MSL_USERFUNC virtual R operator()(T1 x, T2 y, T3 z, T4 v) const = 0;
    virtual ~Functor4() {}
};

template <typename T1, typename T2, typename T3, typename T4, typename T5, typename R>
class Functor5 {
public:
    MSL_USERFUNC virtual R operator()(T1 x, T2 y, T3 z, T4 v, T5 l) const = 0;
    virtual ~Functor5() {}
};

template <typename T1, typename T2, typename T3, typename T4, typename T5, typename T6, typename R>
class Functor6 {
public:
    MSL_USERFUNC virtual R operator()(T1 x, T2 y, T3 z, T4 v, T5 l, T6 m) const = 0;
    virtual ~Functor6() {}
};

template <typename T1, typename T2, typename T3, typename T4, typename R>
class Functor4P {
public:
    MSL_USERFUNC virtual void operator()(T1 x, T2 y, T3 *z, T4 *v) const = 0;
    virtual ~Functor4P() {}
};

template <typename T>
class DM;

namespace NVF {
// Neutral value provider for stencils: a constant for every padding cell.
class NeutralValueFunctor2 : public Functor2<int, int, int> {
public:
    NeutralValueFunctor2(int default_neutral) : default_neutral(default_neutral) {}
    MSL_USERFUNC int operator()(int /*row*/, int /*col*/) const { return default_neutral; }

private:
    int default_neutral;
};
}  // namespace NVF

// Base of map-stencil functors on matrices (stencil radius and tile knobs).
template <typename T, typename R, typename NeutralValueFunctor>
class MMapStencilFunctor : public detail::MatrixFunctorBase {
public:
    MMapStencilFunctor() : stencil_size(1) { this->setTileWidth(msl::DEFAULT_TILE_WIDTH); }
    int getStencilSize() { return stencil_size; }
    void setStencilSize(int value) { stencil_size = value; }
    virtual ~MMapStencilFunctor() {}

protected:
    int stencil_size;
};

// Base of the functors of DM::mapStencilMM: operator()(row, col, PLMatrix<T>*,
// ncol, nrow).  The shared-memory flag is a hint only; tiles are always
// staged in LDS when they fit.
template <typename T, typename R>
class DMMapStencilFunctor : public detail::MatrixFunctorBase {
public:
    DMMapStencilFunctor() : stencil_size(1) { this->setTileWidth(msl::DEFAULT_TILE_WIDTH); }
    int getStencilSize() { return stencil_size; }
    void setStencilSize(int value) { stencil_size = value; }
    bool getSharedMemory() { return shared_memory; }
    void setSharedMemory(bool value) { shared_memory = value; }
    virtual ~DMMapStencilFunctor() {}

protected:
    int stencil_size;
    bool shared_memory = false;
};

template <typename T, typename R, typename NeutralValueFunctor>
class DMNVFMapStencilFunctor : public detail::MatrixFunctorBase {
public:
    DMNVFMapStencilFunctor() : stencil_size(1) { this->setTileWidth(msl::DEFAULT_TILE_WIDTH); }
    int getStencilSize() { return stencil_size; }
    void setStencilSize(int value) { stencil_size = value; }
    bool getSharedMemory() { return shared_memory; }
    void setSharedMemory(bool value) { shared_memory = value; }
    virtual ~DMNVFMapStencilFunctor() {}

protected:
    int stencil_size;
    bool shared_memory = false;
};

template <typename T>
using DCMapStencilFunctor = T (*)(const PLCube<T> &cs, int x, int y, int z);

template <typename T>
using NPLMMapStencilFunctor = T (*)(const NPLMatrix<T> &pl, int x, int y);

}  // namespace msl

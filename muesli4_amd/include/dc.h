/*
 * dc.h -- msl::DC<T>, distributed cube (reference include/dc.h:80-557,
 * src/dc.cpp:49-592).  Layout (depth*rows + row)*cols + col; index
 * skeletons pass (col, row, depth): mapIndex*: f(x, y, z, v); zipIndex*:
 * f(x, y, z, a, b).  mapStencil<f>(result, r, nv): result(x,y,z) =
 * f(PLCube view, x, y, z) with nv outside the cube.  heat7: hand-tuned
 * library kernel for the 7-point heat step (no reference counterpart).
 */
#pragma once

#include <type_traits>

#include "da.h"
#include "ds.h"

namespace msl {

template <typename T>
class DC : public DS<T> {
public:
    DC() : DS<T>(), ncol(0), nrow(0), depth(0) {}
    // The reference rejects depths not divisible by processes x GPUs
    // (src/dc.cpp:52-55); with one process and one GPU every depth is valid.
    DC(int row, int col, int depth) : DS<T>(row * col * depth), ncol(col), nrow(row), depth(depth) {}
    DC(int row, int col, int depth, const T &v) : DS<T>(row * col * depth, v), ncol(col), nrow(row), depth(depth) {}
    DC(int row, int col, int depth, const T &v, bool rowComplete)
        : DS<T>(row * col * depth, v), ncol(col), nrow(row), depth(depth), rowComplete(rowComplete) {}
    DC(const DC<T> &other) : DS<T>(other), ncol(other.ncol), nrow(other.nrow), depth(other.depth) {}
    DC(DC<T> &&other) noexcept : DS<T>(std::move(other)), ncol(other.ncol), nrow(other.nrow), depth(other.depth) {}
    DC<T> &operator=(const DC<T> &other) {
        DS<T>::operator=(other);
        ncol = other.ncol;
        nrow = other.nrow;
        depth = other.depth;
        return *this;
    }
    DC<T> &operator=(DC<T> &&other) noexcept {
        ncol = other.ncol;
        nrow = other.nrow;
        depth = other.depth;
        DS<T>::operator=(std::move(other));
        return *this;
    }

    // ---- index skeletons (src/dc.cpp:385-550) -----------------------------
    template <typename MapIndexFunctor>
    void mapIndexInPlace(MapIndexFunctor &f) {
        this->updateDevice();
        launch_index_map(this->d_data, f);
        this->deviceWritten();
    }
    template <typename MapIndexFunctor>
    void mapIndex(MapIndexFunctor &f, DC<T> &b) {
        this->updateDevice();
        b.updateDevice();
        launch_index_map(b.getUserFunctionData(), f);
        this->deviceWritten();
    }
    template <typename T2, typename ZipIndexFunctor>
    void zipIndexInPlace(DC<T2> &b, ZipIndexFunctor &f) {
        this->updateDevice();
        b.updateDevice();
        launch_index_zip(this->d_data, b.getUserFunctionData(), f);
        this->deviceWritten();
    }
    template <typename T2, typename ZipIndexFunctor>
    void zipIndex(DC<T2> &b, DC<T2> &c, ZipIndexFunctor &f) {
        this->updateDevice();
        b.updateDevice();
        c.updateDevice();
        launch_index_zip(b.getUserFunctionData(), c.getUserFunctionData(), f);
        this->deviceWritten();
    }

    // ---- stencils (src/dc.cpp:559-592) ------------------------------------
    template <msl::DCMapStencilFunctor<T> f>
    void mapStencil(DC<T> &result, size_t stencilSize, T neutralValue) {
        if (result.ncol != ncol || result.nrow != nrow || result.depth != depth) throw detail::NotSameSizeException();
        this->updateDevice();
        if (this->n == 0) return;
        const int r = (int)stencilSize;
        const size_t lds = detail::march3d_lds_bytes(r, sizeof(T));
        if (lds <= detail::kMaxLdsBytes) {
            switch (r) {  // compile-time radius for the common small ones
                case 1: marchStencil<f, 1>(result, r, neutralValue, lds); break;
                case 2: marchStencil<f, 2>(result, r, neutralValue, lds); break;
                default: marchStencil<f, 0>(result, r, neutralValue, lds); break;
            }
        } else {
            detail::dc_stencil_global_kernel<T, f><<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                result.d_data, this->d_data, ncol, nrow, depth, r, neutralValue);
        }
        detail::check_launch("DC::mapStencil");
        result.deviceWritten();
    }
    template <typename MapStencilFunctor, typename NeutralValueFunctor>
    void mapStencilInPlace(MapStencilFunctor &, NeutralValueFunctor &) {
        throw detail::NotYetImplementedException();
    }

    // 7-point heat step v = (in(x-1)+in(x+1)+in(y-1)+in(y+1)+in(z-1)+in(z+1))/6,
    // summed left to right, via the hand-tuned kernel of libmsl_hip.so.
    void heat7(DC<T> &result, T nv) {
        static_assert(std::is_same<T, float>::value, "heat7 needs DC<float>");
        if (result.ncol != ncol || result.nrow != nrow || result.depth != depth) throw detail::NotSameSizeException();
        this->updateDevice();
        detail::check(msl_heat7_f32(result.d_data, this->d_data, ncol, nrow, depth, nv), "DC::heat7");
        result.deviceWritten();
    }

    // ---- element access -------------------------------------------------------
    T get(int row, int col, int d) const { return DS<T>::get((d * nrow + row) * ncol + col); }
    T get(int index) const { return DS<T>::get(index); }
    void set(int row, int col, int d, T value) { DS<T>::set((d * nrow + row) * ncol + col, value); }
    void set(int index, const T &value) { DS<T>::set(index, value); }
    int getCols() { return ncol; }
    int getRows() { return nrow; }
    int getDepth() { return depth; }

    // one row per line, planes separated by a dashed line; like the
    // reference (src/dc.cpp:290) it leaves std::cout at precision 2
    void show(const std::string &descr = std::string(), int limited = 0) {
        std::cout.precision(2);
        this->updateHost();
        const long count = (limited > 0 && limited < this->n) ? limited : this->n;
        printv("%s", detail::show_text(descr, this->localPartition, count, ncol, (long)ncol * nrow).c_str());
    }
    void printnCPU() {}

    int nlocalRows = 0;
    int ncol;
    int nrow;
    int depth;
    int firstRow = 0;
    bool rowComplete = true;

private:
    template <msl::DCMapStencilFunctor<T> f, int RC>
    void marchStencil(DC<T> &result, int r, T nv, size_t lds) {
        const detail::March3D m = detail::march3d_shape(ncol, nrow, depth, lds);
        const int mask = detail::march3d_ring_planes(r) - 1;
        detail::dc_stencil_kernel<T, f, RC><<<m.blocks, detail::kThreads, lds, detail::stream()>>>(
            result.d_data, this->d_data, ncol, nrow, depth, r, nv, m.xtiles, m.ytiles, m.zseg, mask);
    }
    template <typename In, typename F>
    void launch_index_map(const In *in, F &f) {
        if (this->n == 0) return;
        detail::map_index_3d_kernel<In, T, F>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(this->d_data, in, (size_t)this->n,
                                                                                   ncol, nrow, f);
        detail::check_launch("DC::mapIndex");
    }
    template <typename A, typename B, typename F>
    void launch_index_zip(const A *a, const B *b, F &f) {
        if (this->n == 0) return;
        detail::zip_index_3d_kernel<A, B, T, F>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(this->d_data, a, b,
                                                                                   (size_t)this->n, ncol, nrow, f);
        detail::check_launch("DC::zipIndex");
    }
};

}  // namespace msl

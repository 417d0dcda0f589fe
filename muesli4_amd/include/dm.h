/*
 * dm.h -- msl::DM<T>, row-major distributed matrix (reference
 * include/dm.h:81-608, src/dm.cpp:162-1591).
 *
 * Index skeletons pass (row, col): mapIndex*: f(row, col, v); zipIndex*:
 * f(row, col, a, b).  mapIndex(f, b) writes this = f(row, col, b) (the CPU
 * direction of the reference, as in DA and DC; the reference GPU path writes
 * b instead and skips the last element).
 *
 * mapStencil<f>(result, r, nv, shmem): result(x=col, y=row) = f(view, x, y)
 * with an NPLMatrix view over this matrix that returns nv outside the grid.
 * mapStencilMM(result, f, nv | nvf): the legacy functor form
 * f(row, col, PLMatrix*, ncol, nrow).
 * mapStencilMeanBlur / gameOfLife: hand-tuned library kernels for the two
 * stencils of the benchmark workloads (no reference counterpart).
 */
#pragma once

#include <type_traits>

#include "da.h"
#include "ds.h"

namespace msl {

template <typename T>
class DM : public DS<T> {
public:
    DM() : DS<T>(), ncol(0), nrow(0) {}
    DM(int row, int col) : DS<T>(row * col), ncol(col), nrow(row) { initDims(); }
    DM(int row, int col, const T &v) : DS<T>(row * col, v), ncol(col), nrow(row) { initDims(); }
    DM(int row, int col, const T &v, bool rowComplete)
        : DS<T>(row * col, v), ncol(col), nrow(row), rowComplete(rowComplete) {
        initDims();
    }
    DM(const DM<T> &other) : DS<T>(other), ncol(other.ncol), nrow(other.nrow), rowComplete(other.rowComplete) {
        initDims();
    }
    DM(DM<T> &&other) noexcept
        : DS<T>(std::move(other)), ncol(other.ncol), nrow(other.nrow), rowComplete(other.rowComplete) {
        initDims();
    }
    DM<T> &operator=(const DM<T> &other) {
        DS<T>::operator=(other);
        ncol = other.ncol;
        nrow = other.nrow;
        rowComplete = other.rowComplete;
        initDims();
        return *this;
    }
    DM<T> &operator=(DM<T> &&other) noexcept {
        ncol = other.ncol;
        nrow = other.nrow;
        rowComplete = other.rowComplete;
        DS<T>::operator=(std::move(other));
        initDims();
        return *this;
    }
    ~DM() {
        if (d_border) msl_free(d_border);
    }

    void setPointer(const T *pointer) { this->set(pointer); }

    // ---- index skeletons (src/dm.cpp:370-501) ------------------------------
    template <typename MapIndexFunctor>
    void mapIndexInPlace(MapIndexFunctor &f) {
        this->updateDevice();
        launch_index_map(this->d_data, f);
        this->deviceWritten();
    }
    template <typename MapIndexFunctor>
    void mapIndex(MapIndexFunctor &f, DM<T> &b) {
        this->updateDevice();
        b.updateDevice();
        launch_index_map(b.getUserFunctionData(), f);
        this->deviceWritten();
    }
    template <typename T2, typename ZipIndexFunctor>
    void zipIndexInPlace(DM<T2> &b, ZipIndexFunctor &f) {
        this->updateDevice();
        b.updateDevice();
        launch_index_zip(this->d_data, b.getUserFunctionData(), f);
        this->deviceWritten();
    }
    template <typename T2, typename ZipIndexFunctor>
    void zipIndex(DM<T2> &b, DM<T2> &c, ZipIndexFunctor &f) {
        this->updateDevice();
        b.updateDevice();
        c.updateDevice();
        launch_index_zip(b.getUserFunctionData(), c.getUserFunctionData(), f);
        this->deviceWritten();
    }

    // ---- stencils (src/dm.cpp:1535-1591) ----------------------------------
    template <msl::NPLMMapStencilFunctor<T> f>
    void mapStencil(DM<T> &result, size_t stencilSize, T neutralValue, bool /*shared_mem*/ = true) {
        checkSameShape(result);
        this->updateDevice();
        if (this->n == 0) return;
        const int r = (int)stencilSize;
        const size_t lds = detail::march2d_lds_bytes(r, sizeof(T));
        if (lds <= detail::kMaxLdsBytes) {
            const detail::March2D m = detail::march2d_shape(nrow, ncol);
            const int mask = detail::march2d_ring_rows(r) - 1;
            detail::dm_stencil_kernel<T, f><<<m.blocks, detail::kThreads, lds, detail::stream()>>>(
                result.d_data, this->d_data, nrow, ncol, r, neutralValue, m.strips_x, m.seg, mask);
        } else {
            detail::dm_stencil_global_kernel<T, f><<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                result.d_data, this->d_data, nrow, ncol, r, neutralValue);
        }
        detail::check_launch("DM::mapStencil");
        result.deviceWritten();
    }

    // Legacy form (src/dm.cpp:615-1331): f(row, col, PLMatrix<T>*, ncol, nrow)
    // with the constant neutral value outside the grid.
    template <typename T2, typename MapStencilFunctor>
    void mapStencilMM(DM<T2> &result, MapStencilFunctor &f, T neutral_value) {
        runStencilMM(result, f, neutral_value, nullptr, 0);
    }
    // Legacy form with a neutral-value functor nvf(row, col) for the cells of
    // the border band of width f.getStencilSize().
    template <typename T2, typename MapStencilFunctor, typename NeutralValueFunctor>
    void mapStencilMM(DM<T2> &result, MapStencilFunctor &f, NeutralValueFunctor &nvf) {
        const int r = f.getStencilSize();
        const size_t need = (size_t)(nrow + 2 * r) * (size_t)(ncol + 2 * r);
        if (need > borderCap) {
            if (d_border) msl_free(d_border);
            void *p = nullptr;
            detail::check(msl_malloc(&p, need * sizeof(T)), "DM::mapStencilMM border");
            d_border = static_cast<T *>(p);
            borderCap = need;
        }
        detail::border_kernel<T, NeutralValueFunctor>
            <<<detail::grid_for(need), detail::kThreads, 0, detail::stream()>>>(d_border, nrow, ncol, r, nvf);
        detail::check_launch("DM::mapStencilMM border");
        runStencilMM(result, f, (T)nvf(0, 0), d_border, r);
    }
    template <typename MapStencilFunctor, typename NeutralValueFunctor>
    void mapStencilInPlace(MapStencilFunctor &, NeutralValueFunctor &) {
        throw detail::NotYetImplementedException();
    }

    // ---- rotations (src/dm.cpp:1338-1533, one process) ---------------------
    // rotateRows(a): new row k = old row (k - a) mod nrow.  |a| == nrow and
    // a == 0 leave the matrix unchanged; |a| > nrow is reported and ignored,
    // as in the reference.
    void rotateRows(int a) {
        const int howmuch = a < 0 ? -a : a;
        if (howmuch == nrow || a == 0) return;
        if (!rowComplete) {
            throws(detail::RotateRowCompleteNotImplementedException());
            return;
        }
        if (howmuch > nlocalRows) {
            throws(detail::RotateRowManyNotImplementedException());
            return;
        }
        rotate(a < 0 ? a + nrow : a, 0);
    }
    // rotateCols(a): new (i, j) = old (i, (j + a) mod ncol).
    void rotateCols(int a) {
        if (ncol == 0) return;
        a %= ncol;
        if (a == 0) return;
        if (!rowComplete) {
            throws(detail::RotateColCompleteNotImplementedException());
            return;
        }
        rotate(0, a < 0 ? a + ncol : a);
    }

    // ---- library fast paths -------------------------------------------------
    // Mean blur of radius r (1..16), exact summation order of the scalar
    // definition, via the hand-tuned kernel of libmsl_hip.so.
    void mapStencilMeanBlur(DM<T> &result, int r, T nv) {
        static_assert(std::is_same<T, float>::value, "mapStencilMeanBlur needs DM<float>");
        checkSameShape(result);
        this->updateDevice();
        detail::check(msl_stencil2d_mean_f32(result.d_data, this->d_data, nrow, ncol, r, nv), "DM::mapStencilMeanBlur");
        result.deviceWritten();
    }
    // One Game of Life (B3/S23) step with dead cells outside the grid.
    void gameOfLife(DM<T> &result) {
        static_assert(std::is_same<T, int>::value || std::is_same<T, int32_t>::value, "gameOfLife needs DM<int>");
        checkSameShape(result);
        this->updateDevice();
        detail::check(msl_gol_i32(reinterpret_cast<int32_t *>(result.d_data),
                                  reinterpret_cast<const int32_t *>(this->d_data), nrow, ncol),
                      "DM::gameOfLife");
        result.deviceWritten();
    }

    // ---- element access -------------------------------------------------------
    MSL_USERFUNC T get2D(int row, int col) const {
#if defined(__HIP_DEVICE_COMPILE__)
        return this->d_data[(size_t)row * ncol + col];
#else
        return this->get(row * ncol + col);
#endif
    }
    void set2D(int row, int col, T value) { this->set(row * ncol + col, value); }
    int getnCol() { return ncol; }
    int getnRow() { return nrow; }

    void show(const std::string &descr = std::string(), int limited = 0) {
        this->updateHost();
        std::ostringstream s;
        if (!descr.empty()) s << descr << std::endl;
        const int rows = (limited > 0 && limited < nrow) ? limited : nrow;
        for (int i = 0; i < rows; i++) {
            s << "[";
            for (int j = 0; j < ncol; j++) s << this->localPartition[(size_t)i * ncol + j] << (j + 1 < ncol ? " " : "");
            s << "]" << std::endl;
        }
        printv("%s", s.str().c_str());
    }

    int ncol;
    int nrow;
    int nlocalRows = 0;
    int firstRow = 0;
    bool rowComplete = true;

private:
    T *d_border = nullptr;
    size_t borderCap = 0;

    void initDims() {
        nlocalRows = nrow;
        firstRow = 0;
    }
    template <typename U>
    void checkSameShape(const DM<U> &o) const {
        if (o.nrow != nrow || o.ncol != ncol) throw detail::NotSameSizeException();
    }
    void rotate(int row_shift, int col_shift) {
        if (this->n == 0) return;
        this->updateDevice();
        void *p = nullptr;
        detail::check(msl_malloc(&p, sizeof(T) * (size_t)this->n), "DM::rotate");
        T *fresh = static_cast<T *>(p);
        detail::rotate_kernel<T><<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
            fresh, this->d_data, nrow, ncol, row_shift, col_shift);
        detail::check_launch("DM::rotate");
        detail::check(msl_free(this->d_data), "DM::rotate");  // waits for the rotation
        this->d_data = fresh;
        this->deviceWritten();
    }
    template <typename T2, typename F>
    void runStencilMM(DM<T2> &result, F &f, T nv, const T *border, int border_w) {
        checkSameShape(result);
        this->updateDevice();
        if (this->n == 0) return;
        const int r = f.getStencilSize();
        const size_t lds = detail::march2d_lds_bytes(r, sizeof(T));
        if (lds <= detail::kMaxLdsBytes) {
            const detail::March2D m = detail::march2d_shape(nrow, ncol);
            const int mask = detail::march2d_ring_rows(r) - 1;
            detail::dm_stencil_mm_kernel<T, T2, F><<<m.blocks, detail::kThreads, lds, detail::stream()>>>(
                result.getUserFunctionData(), this->d_data, nrow, ncol, r, nv, border, border_w, m.strips_x, m.seg, mask,
                f);
        } else {
            detail::dm_stencil_mm_global_kernel<T, T2, F>
                <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                    result.getUserFunctionData(), this->d_data, nrow, ncol, r, nv, border, border_w, f);
        }
        detail::check_launch("DM::mapStencilMM");
        result.setCpuMemoryInSync(false);
    }
    template <typename In, typename F>
    void launch_index_map(const In *in, F &f) {
        if (this->n == 0) return;
        detail::map_index_2d_kernel<In, T, F>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(this->d_data, in, (size_t)this->n,
                                                                                   ncol, f);
        detail::check_launch("DM::mapIndex");
    }
    template <typename A, typename B, typename F>
    void launch_index_zip(const A *a, const B *b, F &f) {
        if (this->n == 0) return;
        detail::zip_index_2d_kernel<A, B, T, F>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(this->d_data, a, b,
                                                                                   (size_t)this->n, ncol, f);
        detail::check_launch("DM::zipIndex");
    }

    template <typename U>
    friend class DM;
};

}  // namespace msl

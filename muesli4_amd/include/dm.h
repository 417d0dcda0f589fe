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

    // ---- index skeletons with extra operands (src/dm.cpp:502-540,
    // 1757-1911): per-element operands are matrices of this shape, per-row
    // operands are DA<.> with one value per row.
    // this[k] = f(row, col, this[k], b[k], c[k])
    template <typename T2, typename T3, typename ZipIndexFunctor>
    void zipIndexInPlace3(DM<T2> &b, DM<T3> &c, ZipIndexFunctor &f) {
        checkSameShape(b);
        checkSameShape(c);
        this->updateDevice();
        b.updateDevice();
        c.updateDevice();
        if (this->n == 0) return;
        detail::zip3_index_2d_kernel<T, T2, T3, T, ZipIndexFunctor>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                this->d_data, this->d_data, b.getUserFunctionData(), c.getUserFunctionData(), (size_t)this->n, ncol, f);
        detail::check_launch("DM::zipIndexInPlace3");
        this->deviceWritten();
    }
    // this[k] = f(row, col, this[k], b[k], c[row])
    template <typename T2, typename T3, typename ZipIndexFunctor>
    void zipIndexInPlaceMA(DM<T2> &b, DA<T3> &c, ZipIndexFunctor &f) {
        checkSameShape(b);
        checkRowVector(c);
        this->updateDevice();
        b.updateDevice();
        c.updateDevice();
        if (this->n == 0) return;
        detail::zip_index_2d_row_kernel<T, T2, T3, T, ZipIndexFunctor>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                this->d_data, this->d_data, b.getUserFunctionData(), c.getUserFunctionData(), (size_t)this->n, ncol, f);
        detail::check_launch("DM::zipIndexInPlaceMA");
        this->deviceWritten();
    }
    // this[k] = f(row, col, this[k], c[row])
    template <typename T3, typename ZipIndexFunctor>
    void zipIndexInPlaceA(DA<T3> &c, ZipIndexFunctor &f) {
        checkRowVector(c);
        this->updateDevice();
        c.updateDevice();
        if (this->n == 0) return;
        detail::map_index_2d_row_kernel<T, T3, T, ZipIndexFunctor>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                this->d_data, this->d_data, c.getUserFunctionData(), (size_t)this->n, ncol, f);
        detail::check_launch("DM::zipIndexInPlaceA");
        this->deviceWritten();
    }
    // this[k] = f(row, col, this[k], b[row], c[row], d[row])
    template <typename T2, typename T3, typename T4, typename ZipIndexFunctor>
    void zipIndexInPlaceAAA(DA<T2> &b, DA<T3> &c, DA<T4> &d, ZipIndexFunctor &f) {
        checkRowVector(b);
        checkRowVector(c);
        checkRowVector(d);
        this->updateDevice();
        b.updateDevice();
        c.updateDevice();
        d.updateDevice();
        if (this->n == 0) return;
        detail::zip_index_2d_rows3_kernel<T, T2, T3, T4, T, ZipIndexFunctor>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                this->d_data, this->d_data, b.getUserFunctionData(), c.getUserFunctionData(), d.getUserFunctionData(),
                (size_t)this->n, ncol, f);
        detail::check_launch("DM::zipIndexInPlaceAAA");
        this->deviceWritten();
    }
    // this[k] = f(row, col, this[k], b) with b the device pointer of the
    // whole matrix b (the reference CPU form, src/dm.cpp:525-532; its GPU
    // kernel calls f(row, col, this, b) without storing a result)
    template <typename T2, typename ZipIndexFunctor>
    void crossZipIndexInPlace(DM<T2> &b, ZipIndexFunctor &f) {
        this->updateDevice();
        b.updateDevice();
        if (this->n == 0) return;
        detail::cross_zip_index_2d_kernel<T, T2, T, ZipIndexFunctor>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                this->d_data, this->d_data, b.getUserFunctionData(), (size_t)this->n, ncol, f);
        detail::check_launch("DM::crossZipIndexInPlace");
        this->deviceWritten();
    }

    // ---- row reductions and column writes (src/dm.cpp:1594-1871) ----------
    // c[row] = fold of row `row` with f(acc, x): one wave per row, each lane
    // folding its columns in order, then the 64 lane results in lane order.
    // The row's own elements only (the reference GPU path; its CPU path
    // starts from T(0)).
    template <typename ReduceFunctor>
    void reduceRows(DA<T> &c, ReduceFunctor &f) {
        checkRowVector(c);
        this->updateDevice();
        c.updateDevice();
        if (nrow == 0) return;
        if (ncol == 0) throw detail::SizeMismatchException();
        const int rows_per_block = detail::kThreads / 64;
        detail::reduce_rows_kernel<T, ReduceFunctor>
            <<<(nrow + rows_per_block - 1) / rows_per_block, detail::kThreads, 0, detail::stream()>>>(
                c.getUserFunctionData(), this->d_data, nrow, ncol, f);
        detail::check_launch("DM::reduceRows");
        c.setCpuMemoryInSync(false);
    }
    // c[row] = f(m[row][ncol-1], m[row][0], ... f(m[row][0], m[row][1], T2(0)))
    // in the reference's sequential order (src/dm.cpp:1731-1754).
    template <typename T2, typename ReduceFunctor>
    void reducetwoColumns(DA<T2> &c, ReduceFunctor &f) {
        checkRowVector(c);
        this->updateDevice();
        c.updateDevice();
        if (nrow == 0) return;
        detail::reduce_two_columns_kernel<T, T2, ReduceFunctor>
            <<<detail::grid_for(nrow), detail::kThreads, 0, detail::stream()>>>(c.getUserFunctionData(), this->d_data,
                                                                                nrow, ncol, f);
        detail::check_launch("DM::reducetwoColumns");
        c.setCpuMemoryInSync(false);
    }
    // this[row][col] = da[row] (src/dm.cpp:1856-1880)
    void setColumn(DA<T> &da, int col) {
        if (col < 0 || col >= ncol) throw detail::IndexOutOfBoundsException();
        if (da.getSize() != nrow) throw detail::SizeMismatchException();
        this->updateDevice();
        da.updateDevice();
        if (nrow == 0) return;
        detail::set_column_kernel<T><<<detail::grid_for(nrow), detail::kThreads, 0, detail::stream()>>>(
            this->d_data, da.getUserFunctionData(), nrow, ncol, col);
        detail::check_launch("DM::setColumn");
        this->deviceWritten();
    }
    // one process holds every partition
    void broadcastPartition(int /*partitionIndex*/) {}

    // ---- stencils (src/dm.cpp:1535-1591) ----------------------------------
    template <msl::NPLMMapStencilFunctor<T> f>
    void mapStencil(DM<T> &result, size_t stencilSize, T neutralValue, bool /*shared_mem*/ = true) {
        checkSameShape(result);
        this->updateDevice();
        if (this->n == 0) return;
        const int r = (int)stencilSize;
        const size_t lds = detail::march2d_lds_bytes(r, sizeof(T));
        if (lds <= detail::kMaxLdsBytes) {
            switch (r) {  // compile-time radius for the common small ones
                case 1: marchStencil<f, 1>(result, r, neutralValue, lds); break;
                case 2: marchStencil<f, 2>(result, r, neutralValue, lds); break;
                case 3: marchStencil<f, 3>(result, r, neutralValue, lds); break;
                case 4: marchStencil<f, 4>(result, r, neutralValue, lds); break;
                default: marchStencil<f, 0>(result, r, neutralValue, lds); break;
            }
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

    // one matrix row per line; like the reference (src/dm.cpp:271) it leaves
    // std::cout in fixed notation
    void show(const std::string &descr = std::string(), int limited = 0) {
        std::cout << std::fixed;
        this->updateHost();
        const long count = (limited > 0 && limited < this->n) ? limited : this->n;
        printv("%s", detail::show_text(descr, this->localPartition, count, ncol, 0).c_str());
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
    template <typename U>
    void checkRowVector(const DA<U> &v) const {
        if (v.getSize() != nrow) throw detail::SizeMismatchException();
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
            switch (r) {  // compile-time radius for the common small ones
                case 1: marchStencilMM<1>(result, f, r, nv, border, border_w, lds); break;
                case 2: marchStencilMM<2>(result, f, r, nv, border, border_w, lds); break;
                default: marchStencilMM<0>(result, f, r, nv, border, border_w, lds); break;
            }
        } else {
            detail::dm_stencil_mm_global_kernel<T, T2, F>
                <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                    result.getUserFunctionData(), this->d_data, nrow, ncol, r, nv, border, border_w, f);
        }
        detail::check_launch("DM::mapStencilMM");
        result.setCpuMemoryInSync(false);
    }
    template <msl::NPLMMapStencilFunctor<T> f, int RC>
    void marchStencil(DM<T> &result, int r, T nv, size_t lds) {
        const detail::March2D m = detail::march2d_shape(nrow, ncol);
        const int mask = detail::march2d_ring_rows(r) - 1;
        detail::dm_stencil_kernel<T, f, RC><<<m.blocks, detail::kThreads, lds, detail::stream()>>>(
            result.d_data, this->d_data, nrow, ncol, r, nv, m.strips_x, m.seg, mask);
    }
    template <int RC, typename T2, typename F>
    void marchStencilMM(DM<T2> &result, F &f, int r, T nv, const T *border, int border_w, size_t lds) {
        const detail::March2D m = detail::march2d_shape(nrow, ncol);
        const int mask = detail::march2d_ring_rows(r) - 1;
        // without a border band the cells off the grid are nv: the staging
        // then has no conditional load of the band
        if (border)
            detail::dm_stencil_mm_kernel<T, T2, F, RC, true><<<m.blocks, detail::kThreads, lds, detail::stream()>>>(
                result.getUserFunctionData(), this->d_data, nrow, ncol, r, nv, border, border_w, m.strips_x, m.seg,
                mask, f);
        else
            detail::dm_stencil_mm_kernel<T, T2, F, RC, false><<<m.blocks, detail::kThreads, lds, detail::stream()>>>(
                result.getUserFunctionData(), this->d_data, nrow, ncol, r, nv, nullptr, 0, m.strips_x, m.seg, mask, f);
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

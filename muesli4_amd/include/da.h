/*
 * da.h -- msl::DA<T>, one-dimensional distributed array (reference
 * include/da.h:80-236, src/da.cpp:230-459).  Index skeletons pass the global
 * index first: mapIndex*: f(i, v); zipIndex*: f(i, a, b).
 */
#pragma once

#include "ds.h"

namespace msl {

template <typename T>
class DM;

template <typename T>
class DA : public DS<T> {
public:
    DA() : DS<T>() {}
    DA(int size) : DS<T>(size) {}
    DA(int size, const T &v) : DS<T>(size, v) {}
    DA(const DA<T> &other) : DS<T>(other) {}
    DA(DA<T> &&other) noexcept : DS<T>(std::move(other)) {}
    DA<T> &operator=(const DA<T> &other) {
        DS<T>::operator=(other);
        return *this;
    }
    DA<T> &operator=(DA<T> &&other) noexcept {
        DS<T>::operator=(std::move(other));
        return *this;
    }

    // this[i] = f(i, this[i])  (src/da.cpp:230-255)
    template <typename MapIndexFunctor>
    void mapIndexInPlace(MapIndexFunctor &f) {
        this->updateDevice();
        launch_index_map(this->d_data, f);
        this->deviceWritten();
    }
    // this[i] = f(i, b[i])  (src/da.cpp:326-355)
    template <typename MapIndexFunctor>
    void mapIndex(MapIndexFunctor &f, DA<T> &b) {
        this->updateDevice();
        b.updateDevice();
        launch_index_map(b.getUserFunctionData(), f);
        this->deviceWritten();
    }
    // this[i] = f(i, this[i], b[i])  (src/da.cpp:373-399)
    template <typename T2, typename ZipIndexFunctor>
    void zipIndexInPlace(DA<T2> &b, ZipIndexFunctor &f) {
        this->updateDevice();
        b.updateDevice();
        launch_index_zip(this->d_data, b.getUserFunctionData(), f);
        this->deviceWritten();
    }
    // this[i] = f(i, b[i], c[i])  (src/da.cpp:401-430; operand order as the
    // signature, the GPU order of the reference)
    template <typename T2, typename ZipIndexFunctor>
    void zipIndex(DA<T2> &b, DA<T2> &c, ZipIndexFunctor &f) {
        this->updateDevice();
        b.updateDevice();
        c.updateDevice();
        launch_index_zip(b.getUserFunctionData(), c.getUserFunctionData(), f);
        this->deviceWritten();
    }
    // this[i] = f(i, &rows[i][0], this[i]): the functor sees row i of the
    // matrix through a device pointer (src/da.cpp:256-289)
    template <typename MapIndexFunctor, typename T2>
    void mapIndexInPlaceDMRows(DM<T2> &rows, MapIndexFunctor &f) {
        if (this->n > rows.getnRow()) throw detail::SizeMismatchException();
        this->updateDevice();
        rows.updateDevice();
        if (this->n == 0) return;
        detail::map_index_dmrows_kernel<T, T2, T, MapIndexFunctor>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                this->d_data, this->d_data, rows.getUserFunctionData(), (size_t)this->n, rows.getnCol(), f);
        detail::check_launch("DA::mapIndexInPlaceDMRows");
        this->deviceWritten();
    }
    // this[i] = f(i, &rows[i][0], b[i], this[i])  (src/da.cpp:290-324)
    template <typename MapIndexFunctor, typename T1, typename T2>
    void zipIndexInPlaceDMRows(DA<T1> &b, DM<T2> &rows, MapIndexFunctor &f) {
        if (this->n > rows.getnRow() || b.getSize() != this->n) throw detail::SizeMismatchException();
        this->updateDevice();
        b.updateDevice();
        rows.updateDevice();
        if (this->n == 0) return;
        detail::zip_index_dmrows_kernel<T, T1, T2, T, MapIndexFunctor>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(
                this->d_data, this->d_data, b.getUserFunctionData(), rows.getUserFunctionData(), (size_t)this->n,
                rows.getnCol(), f);
        detail::check_launch("DA::zipIndexInPlaceDMRows");
        this->deviceWritten();
    }
    // this[i] = f(this[i], b[i], c[i])  (src/da.cpp:432-459)
    template <typename T2, typename T3, typename ZipFunctor>
    void zipInPlace3(DA<T2> &b, DA<T3> &c, ZipFunctor &f) {
        DS<T>::zipInPlace3(b, c, f);
    }

    // One process holds every partition: broadcasting or permuting partitions
    // leaves the data unchanged (src/da.cpp:165-225 with np = 1).
    void broadcastPartition(int /*partitionIndex*/) {}
    template <typename Functor>
    inline void permutePartition(Functor & /*f*/) {}


private:
    template <typename In, typename F>
    void launch_index_map(const In *in, F &f) {
        if (this->n == 0) return;
        detail::map_index_1d_kernel<In, T, F>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(this->d_data, in, (size_t)this->n,
                                                                                   this->firstIndex, f);
        detail::check_launch("DA::mapIndex");
    }
    template <typename A, typename B, typename F>
    void launch_index_zip(const A *a, const B *b, F &f) {
        if (this->n == 0) return;
        detail::zip_index_1d_kernel<A, B, T, F>
            <<<detail::grid_for(this->n), detail::kThreads, 0, detail::stream()>>>(this->d_data, a, b, (size_t)this->n,
                                                                                   this->firstIndex, f);
        detail::check_launch("DA::zipIndex");
    }
};

}  // namespace msl

/*
 * ds.h -- msl::DS<T>, the device-resident distributed-structure base of DA,
 * DM and DC (reference include/ds.h:69-642, src/ds.cpp:42-954).
 *
 * One process, one GPU: the n elements live in one device buffer that the
 * object owns; localPartition is a host mirror (allocated with the object,
 * as in the reference, because programs read it directly) that is refreshed
 * lazily.  Skeletons run as stream-ordered kernels; a result is valid for
 * every later skeleton, gather() or get() without an explicit sync.
 * Semantics follow the reference GPU path with its inconsistencies resolved:
 * map(f, b) writes this = f(b); zip(b, c, f) writes this = f(b, c) in
 * signature order; fold is deterministic for a given device.
 */
#pragma once

#include <cstring>
#include <sstream>
#include <string>
#include <utility>

#include "detail/exec_plan.h"
#include "detail/skeleton_kernels.h"
#include "functors.h"
#include "muesli.h"

namespace msl {

template <typename T>
class DS {
public:
    // ---- construction (src/ds.cpp:63-131) ---------------------------------
    DS() : n(0) { init(); }
    explicit DS(int size) : n(size) { init(); }
    DS(int elements, const T &v) : n(elements) {
        init();
        fill(v);
    }
    DS(const DS<T> &other) : n(other.n) {
        init();
        copyFrom(other);
    }
    DS(DS<T> &&other) noexcept { steal(other); }
    DS<T> &operator=(const DS<T> &other) {
        if (this != &other) {
            release();
            n = other.n;
            init();
            copyFrom(other);
        }
        return *this;
    }
    DS<T> &operator=(DS<T> &&other) noexcept {
        if (this != &other) {
            release();
            steal(other);
        }
        return *this;
    }
    virtual ~DS() { release(); }

    // ---- fill (src/ds.cpp:289-321) -----------------------------------------
    void fill(const T &element) {
        if (n > 0) {
            detail::fill_kernel<T><<<detail::grid_for(n), detail::kThreads, 0, detail::stream()>>>(d_data, (size_t)n,
                                                                                                  element);
            detail::check_launch("DS::fill");
        }
        deviceWritten();
    }
    // Copies the values (the reference adopts the pointer on one GPU; a copy
    // keeps ownership unambiguous).
    void fill(T *const values) {
        if (n > 0) std::memcpy(localPartition, values, sizeof(T) * (size_t)n);
        hostWritten();
        updateDevice();
    }
    void fill(T (*f)(int)) {
        for (int i = 0; i < n; i++) localPartition[i] = f(i + firstIndex);
        hostWritten();
        updateDevice();
    }
    template <typename F>
    void fill(const F &f) {
        for (int i = 0; i < n; i++) localPartition[i] = f(i + firstIndex);
        hostWritten();
        updateDevice();
    }
    void fill_root_init(T *const values) { fill(values); }

    // ---- map (src/ds.cpp:612-663) ------------------------------------------
    template <typename MapFunctor>
    void mapInPlace(MapFunctor &f) {
        updateDevice();
        launch_map(d_data, d_data, f);
        deviceWritten();
    }
    // this[i] = f(b[i])
    template <typename F>
    void map(F &f, DS<T> &b) {
        updateDevice();
        b.updateDevice();
        launch_map(d_data, b.d_data, f);
        deviceWritten();
    }

    // ---- zip (src/ds.cpp:667-727) ------------------------------------------
    // this[i] = f(this[i], b[i])
    template <typename T2, typename ZipFunctor>
    void zipInPlace(DS<T2> &b, ZipFunctor &f) {
        updateDevice();
        b.updateDevice();
        launch_zip(d_data, d_data, b.getUserFunctionData(), f);
        deviceWritten();
    }
    // this[i] = f(b[i], c[i])
    template <typename T2, typename ZipFunctor>
    void zip(DS<T2> &b, DS<T2> &c, ZipFunctor &f) {
        updateDevice();
        b.updateDevice();
        c.updateDevice();
        launch_zip(d_data, b.getUserFunctionData(), c.getUserFunctionData(), f);
        deviceWritten();
    }
    // this[i] = f(this[i], b[i], c[i]) (not implemented by the reference).
    template <typename T2, typename T3, typename ZipFunctor>
    void zipInPlace3(DS<T2> &b, DS<T3> &c, ZipFunctor &f) {
        updateDevice();
        b.updateDevice();
        c.updateDevice();
        if (n > 0) {
            detail::zip3_kernel<T, T2, T3, T, ZipFunctor>
                <<<detail::grid_for(n), detail::kThreads, 0, detail::stream()>>>(
                    d_data, d_data, b.getUserFunctionData(), c.getUserFunctionData(), (size_t)n, f);
            detail::check_launch("DS::zipInPlace3");
        }
        deviceWritten();
    }

    // ---- fold (src/ds.cpp:739-954) -----------------------------------------
    // Device fold; the combination order is fixed (per-thread index order,
    // then a pairwise tree), so the result is deterministic for a device.
    // The flag of the reference (final fold on the CPU) is accepted and
    // ignored: the final combine runs on the device and stores the result
    // straight into the library's pinned result slot (no copy launch).
    // (Folding the partials in the last block of the first launch instead
    // needs an agent-scope release per block, an L2 write-back each, which
    // measured slower: 44.5 -> 53.5 us on 8192^2 floats.)
    template <typename FoldFunctor>
    T fold(FoldFunctor &f, bool /*final_fold_on_cpu*/ = true) {
        if (n == 0) throw detail::NotSameSizeException();
        updateDevice();
        const int blocks = detail::stream_grid(n, 2);
        ensureFoldScratch(blocks);
        void *slot = nullptr;
        size_t slot_bytes = 0;
        detail::check(msl_result_slot(&slot, &slot_bytes), "DS::fold");
        T *result = sizeof(T) <= slot_bytes ? static_cast<T *>(slot) : d_part + blocks;
        detail::fold_stage1_kernel<T, FoldFunctor>
            <<<blocks, detail::kThreads, 0, detail::stream()>>>(d_data, (size_t)n, d_part, d_has, f,
                                                                detail::aligned16(d_data));
        detail::check_launch("DS::fold");
        detail::fold_stage2_kernel<T, FoldFunctor>
            <<<1, detail::kThreads, 0, detail::stream()>>>(d_part, d_has, blocks, result, d_has + blocks, f);
        detail::check_launch("DS::fold");
        T r{};
        if (result == slot) {
            detail::check(msl_synchronize(), "DS::fold");
            std::memcpy(static_cast<void *>(&r), slot, sizeof(T));
        } else {
            detail::check(msl_memcpy_d2h(&r, result, sizeof(T)), "DS::fold");
        }
        return r;
    }
    // Host fold in index order (src/ds.cpp:922-954 without its data race).
    template <typename FoldFunctor>
    T foldCPU(FoldFunctor &f) {
        updateHost();
        T r = localPartition[0];
        for (int i = 1; i < n; i++) r = f(r, localPartition[i]);
        return r;
    }

    // ---- gather / host access (src/ds.cpp:263-591) ------------------------
    // Returns the host mirror (valid until the next skeleton modifies it).
    T *gather() {
        updateHost();
        return localPartition;
    }
    void gather(msl::DS<T> &ds) {
        updateDevice();
        if (ds.n != n) throw detail::NotSameSizeException();
        detail::check(msl_memcpy_d2d(ds.d_data, d_data, sizeof(T) * (size_t)n), "DS::gather");
        ds.deviceWritten();
    }
    void gather(T *&templatepointer) {
        updateHost();
        std::memcpy(templatepointer, localPartition, sizeof(T) * (size_t)n);
    }

    // the host mirror, refreshed from the device first (src/ds.cpp:263-267)
    T *getLocalPartition() {
        updateHost();
        return localPartition;
    }
    T *getnCPUPartition() { return localPartition; }
    void setLocalPartition(T *elements) {
        std::memcpy(localPartition, elements, sizeof(T) * (size_t)n);
        hostWritten();
    }
    T get(int index) const {
        if (index < 0 || index >= n) throw detail::IndexOutOfBoundsException();
        if (hostValid) return localPartition[index];
        T v{};
        detail::check(msl_memcpy_d2h(&v, d_data + index, sizeof(T)), "DS::get");
        return v;
    }
    // Writes the host mirror; the device copy is refreshed once, by the next
    // skeleton or updateDevice() (the reference copies each element both ways).
    void set(int globalIndex, const T &v) {
        if (globalIndex < 0 || globalIndex >= n) throw detail::IndexOutOfBoundsException();
        updateHost();
        localPartition[globalIndex] = v;
        hostWritten();
    }
    void set(const T *pointer) {
        std::memcpy(localPartition, pointer, sizeof(T) * (size_t)n);
        hostWritten();
    }
    int getSize() const { return n; }
    int getLocalSize() const { return nLocal; }
    int getFirstIndex() const { return firstIndex; }
    // false marks the host mirror stale (device newer), true marks it current.
    void setCpuMemoryInSync(bool b) {
        if (b) {
            hostValid = true;
        } else {
            hostValid = false;
            deviceValid = true;
        }
        cpuMemoryInSync = hostValid;
    }
    bool isLocal(int index) const { return index >= firstIndex && index < firstIndex + nLocal; }
    T getLocal(int localIndex) { return get(localIndex); }
    void setLocal(int localIndex, const T &v) { set(localIndex, v); }
    T operator[](int index) { return get(index); }

    GPUExecutionPlan<T> *getExecPlans() {
        plan.d_Data = d_data;
        plan.h_Data = localPartition;
        plan.size = n;
        plan.nLocal = n;
        plan.first = 0;
        plan.bytes = sizeof(T) * (size_t)n;
        return &plan;
    }
    GPUExecutionPlan<T> getExecPlan(int /*device*/) { return *getExecPlans(); }
    void setCopyDistribution() { dist = COPY; }
    void setDistribution() { dist = DIST; }

    // ---- residency (src/ds.cpp:441-484) ------------------------------------
    void updateDevice(int forceupdate = 0) {
        if ((forceupdate || !deviceValid) && n > 0) {
            detail::check(msl_memcpy_h2d(d_data, localPartition, sizeof(T) * (size_t)n), "DS::updateDevice");
        }
        deviceValid = true;
    }
    void updateHost() {
        if (!hostValid && n > 0) {
            detail::check(msl_memcpy_d2h(localPartition, d_data, sizeof(T) * (size_t)n), "DS::updateHost");
        }
        hostValid = true;
        cpuMemoryInSync = true;
    }
    void freeDevice() {}
    void printTime() {}

    MSL_USERFUNC T *getUserFunctionData() { return d_data; }

    void showLocal(const std::string &descr) {
        updateHost();
        printv("%s", detail::show_local_text(descr, localPartition, n).c_str());
    }
    // limited > 0 prints the first `limited` elements only
    void show(const std::string &descr = std::string(), int limited = 0) {
        updateHost();
        const long count = (limited > 0 && limited < n) ? limited : n;
        printv("%s", detail::show_text(descr, localPartition, count, 0, 0).c_str());
    }
    void printLocal() { show(); }
    long getnCPU() { return 0; }

    // ---- attributes with the reference's names ------------------------------
    int id = 0;
    T *localPartition = nullptr;  // host mirror
    int n = 0;
    int nLocal = 0;
    int firstIndex = 0;
    int np = 1;
    bool cpuMemoryInSync = true;
    Distribution dist = DIST;
    int ng = 1;
    long nGPU = 0;
    long nCPU = 0;

protected:
    T *d_data = nullptr;  // device buffer (owned)
    bool hostValid = true;
    bool deviceValid = true;
    GPUExecutionPlan<T> plan;
    // fold scratch: partials and "has value" flags, one extra slot each
    T *d_part = nullptr;
    int *d_has = nullptr;
    int foldCap = 0;

    void init() {
        nLocal = n;
        nGPU = n;
        if (n > 0) {
            void *p = nullptr;
            detail::check(msl_malloc(&p, sizeof(T) * (size_t)n), "DS: device allocation");
            d_data = static_cast<T *>(p);
            localPartition = new T[(size_t)n];
        }
        hostValid = deviceValid = true;
        cpuMemoryInSync = true;
    }
    void release() {
        if (d_data) msl_free(d_data);
        if (d_part) msl_free(d_part);
        if (d_has) msl_free(d_has);
        delete[] localPartition;
        d_data = nullptr;
        d_part = nullptr;
        d_has = nullptr;
        foldCap = 0;
        localPartition = nullptr;
    }
    void steal(DS<T> &o) {
        n = o.n;
        nLocal = o.nLocal;
        nGPU = o.nGPU;
        firstIndex = o.firstIndex;
        d_data = o.d_data;
        localPartition = o.localPartition;
        hostValid = o.hostValid;
        deviceValid = o.deviceValid;
        cpuMemoryInSync = o.cpuMemoryInSync;
        d_part = o.d_part;
        d_has = o.d_has;
        foldCap = o.foldCap;
        o.d_data = nullptr;
        o.localPartition = nullptr;
        o.d_part = nullptr;
        o.d_has = nullptr;
        o.foldCap = 0;
        o.n = o.nLocal = 0;
    }
    void copyFrom(const DS<T> &other) {
        if (n == 0) return;
        DS<T> &o = const_cast<DS<T> &>(other);
        o.updateDevice();
        detail::check(msl_memcpy_d2d(d_data, o.d_data, sizeof(T) * (size_t)n), "DS copy");
        deviceWritten();
    }
    void deviceWritten() {
        deviceValid = true;
        hostValid = false;
        cpuMemoryInSync = false;
    }
    void hostWritten() {
        hostValid = true;
        deviceValid = false;
        cpuMemoryInSync = true;
    }
    void ensureFoldScratch(int blocks) {
        if (blocks + 1 <= foldCap) return;
        if (d_part) msl_free(d_part);
        if (d_has) msl_free(d_has);
        void *p = nullptr;
        detail::check(msl_malloc(&p, sizeof(T) * (size_t)(blocks + 1)), "DS::fold scratch");
        d_part = static_cast<T *>(p);
        detail::check(msl_malloc(&p, sizeof(int) * (size_t)(blocks + 1)), "DS::fold scratch");
        d_has = static_cast<int *>(p);
        foldCap = blocks + 1;
    }
    template <typename F>
    void launch_map(T *out, const T *in, F &f) {
        if (n == 0) return;
        const bool quad = detail::aligned16(out) && detail::aligned16(in);
        detail::map_kernel<T, T, F>
            <<<quad ? detail::stream_grid(n, 1) : detail::grid_for(n), detail::kThreads, 0, detail::stream()>>>(
                out, in, (size_t)n, f, quad);
        detail::check_launch("DS::map");
    }
    template <typename A, typename B, typename F>
    void launch_zip(T *out, const A *a, const B *b, F &f) {
        if (n == 0) return;
        const bool quad = detail::aligned16(out) && detail::aligned16(a) && detail::aligned16(b);
        detail::zip_kernel<A, B, T, F>
            <<<quad ? detail::stream_grid(n, 1) : detail::grid_for(n), detail::kThreads, 0, detail::stream()>>>(
                out, a, b, (size_t)n, f, quad);
        detail::check_launch("DS::zip");
    }

    template <typename U>
    friend class DS;
};

}  // namespace msl

/*
 * muesli.h -- runtime of the MI355X skeleton library (drop-in for Muesli's
 * muesli.h, reference include/muesli.h:127-644 and src/muesli.cpp:63-294).
 *
 * One process drives one GPU: initSkeletons initialises the device context
 * of libmsl_hip.so, every skeleton is ordered on its single stream, and the
 * MPI-era globals keep their names with single-process values (proc_id 0,
 * num_total_procs 1).  Programs written against the reference compile
 * unchanged with hipcc; MPI_Wtime is provided when no MPI header is used.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "msl_hip.h"
#include "detail/exception.h"

#define MSL_USERFUNC __host__ __device__
#define MSL_GPUFUNC __device__
#define MSL_CPUFUNC __host__

#ifndef MPI_VERSION
// Single-process stand-ins for the two MPI calls example programs make
// directly (timing and the rank query); the library itself uses no MPI.
inline double MPI_Wtime() {
    using clock = std::chrono::steady_clock;
    return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}
#endif

// Device overload of std::rand() for user functors that call it inside an
// MSL_USERFUNC (examples/meanblur2.cpp:59-64 initialise their cubes this way).
// Each call takes the next value of a per-module counter, hashed with
// SplitMix64 into [0, RAND_MAX]: the multiset of values is fixed for a given
// number of calls, the element each value lands on is not (as with an
// unseeded host rand() shared by threads).
static __device__ unsigned long long msl_device_rand_counter = 0;
namespace std {
__device__ inline int rand() noexcept {
    unsigned long long z = atomicAdd(&msl_device_rand_counter, 1ull) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (int)(z % ((unsigned long long)RAND_MAX + 1ull));
}
}  // namespace std

#ifndef _OPENMP
#ifndef omp_get_thread_num
#define omp_get_thread_num() 0
#endif
#endif

namespace msl {

typedef int ProcessorNo;
enum Distribution { DIST, COPY };

// Global configuration (reference include/muesli.h:149-181).  Values that
// described MPI ranks or CPU/GPU splits are kept for source compatibility;
// cpu_fraction is accepted and ignored (all elements live on the GPU).
class Muesli {
public:
    static inline int proc_id = 0;
    static inline int proc_entrance = 0;
    static inline int running_proc_no = 0;
    static inline int num_total_procs = 1;
    static inline int num_local_procs = 1;
    static inline double start_time = 0.0;
    static inline char *program_name = nullptr;
    static inline int distribution_mode = 0;
    static inline int task_group_size = 256;
    static inline int num_conc_kernels = 16;
    static inline int num_threads = 1;
    static inline int num_runs = 1;
    static inline int num_gpus = 1;
    static inline int reps = 1;
    static inline double cpu_fraction = 0.0;
    static inline int max_gpus = 1;
    static inline int threads_per_block = 256;
    static inline int tpb_x = 16;
    static inline int tpb_y = 16;
    static inline bool debug = false;
    static inline bool use_timer = false;
    static inline int elem_per_thread = 1;
};

static const int MYTAG = 1;
static const int MYADULTTAG = 18;
static const int STOPTAG = 2;
static const int TERMINATION_TEST = 3;
static const int RANDOM_DISTRIBUTION = 1;
static const int CYCLIC_DISTRIBUTION = 2;
static const int DEFAULT_DISTRIBUTION = CYCLIC_DISTRIBUTION;
static const int UNDEFINED = -1;
static const int DEFAULT_TASK_GOUP_SIZE = 256;
static const int DEFAULT_NUM_CONC_KERNELS = 16;
static const int DEFAULT_NUM_RUNS = 1;
static const int DEFAULT_TILE_WIDTH = 16;

namespace detail {

// Raises the C-ABI status as a C++ exception carrying msl_last_error().
inline void check(int rc, const char *what) {
    if (rc != MSL_OK) {
        const char *m = msl_last_error();
        throw RuntimeError(std::string(what) + ": " + (m ? m : "") + " (status " + std::to_string(rc) + ")");
    }
}

inline hipStream_t stream() {
    void *s = msl_stream();
    if (!s) check(MSL_E_STATE, "msl_stream");
    return (hipStream_t)s;
}

inline int num_cus() {
    static int cus = 0;
    if (!cus) {
        check(msl_device_info(&cus, nullptr, nullptr, nullptr, 0), "msl_device_info");
        if (cus <= 0) cus = 256;
    }
    return cus;
}

inline void check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw RuntimeError(std::string(what) + ": " + hipGetErrorString(e));
}

// Reads one element of a view's LDS window.  In device code the pointer is
// re-typed as an LDS (address space 3) pointer: a generic pointer would let
// the compiler merge the window read with the view's global-memory fallback
// into one flat load per access.  Class types keep the generic read (their
// copy constructors take generic references).
template <typename T>
MSL_USERFUNC __forceinline__ T lds_read(const T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_arithmetic<T>::value) return *(const __attribute__((address_space(3))) T *)p;
#endif
    return *p;
}

// Element idx of an LDS array: the index is applied to the 32-bit LDS
// pointer, so the address is 32-bit integer arithmetic (indexing the generic
// pointer first would compute it in 64 bits).
template <typename T>
MSL_USERFUNC __forceinline__ T lds_read(const T *base, int idx) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_arithmetic<T>::value) return ((const __attribute__((address_space(3))) T *)base)[idx];
#endif
    return base[idx];
}

struct Timing {
    double start = 0.0, split = 0.0;
};
inline Timing &timing() {
    static Timing t;
    return t;
}

}  // namespace detail

// ---- runtime entry points (reference src/muesli.cpp:63-276) ---------------
inline void initSkeletons(int argc, char **argv, bool debug = false) {
    detail::check(msl_init(0), "initSkeletons");
    Muesli::proc_id = 0;
    Muesli::num_total_procs = 1;
    Muesli::num_local_procs = 1;
    Muesli::num_gpus = 1;
    Muesli::max_gpus = 1;
    Muesli::program_name = argc > 0 ? argv[0] : nullptr;
    Muesli::debug = debug;
    Muesli::start_time = MPI_Wtime();
}

// Waits for all outstanding skeletons.  The device context stays alive until
// the process ends so that containers destroyed afterwards can still free.
inline void terminateSkeletons() { detail::check(msl_synchronize(), "terminateSkeletons"); }

inline void syncStreams() { detail::check(msl_synchronize(), "syncStreams"); }

inline bool isRootProcess() { return true; }

inline void printv(const char *format, ...) {
    va_list argp;
    va_start(argp, format);
    vprintf(format, argp);
    va_end(argp);
}

inline void setNumThreads(int n) { Muesli::num_threads = n; }
inline void setNumRuns(int n) { Muesli::num_runs = n; }
// Only one GPU is used; other requests are recorded but do not change that.
inline void setNumGpus(int n) { Muesli::num_gpus = n > 0 ? 1 : 1; }
inline void setThreadsPerBlock(int n) { Muesli::threads_per_block = n; }
inline void setThreadsPerBlock(int x, int y) {
    Muesli::tpb_x = x;
    Muesli::tpb_y = y;
}
inline void setNumConcurrentKernels(int n) { Muesli::num_conc_kernels = n; }
inline void setTaskGroupSize(int n) { Muesli::task_group_size = n; }
inline void setDebug(bool v) { Muesli::debug = v; }
inline void setReps(int v) { Muesli::reps = v; }

inline void startTiming() {
    syncStreams();
    detail::timing().start = detail::timing().split = MPI_Wtime();
}
inline double splitTime(int /*run*/) {
    syncStreams();
    double now = MPI_Wtime();
    double r = now - detail::timing().split;
    detail::timing().split = now;
    return r;
}
inline double stopTiming() {
    syncStreams();
    return MPI_Wtime() - detail::timing().start;
}
inline void printTimeToFile(const char *id, const char *file_name) {
    std::ofstream f(file_name, std::ios::app);
    f << id << ";" << stopTiming() << "\n";
}

inline void fail_exit() {
    syncStreams();
    std::exit(EXIT_FAILURE);
}
inline void throws(const detail::Exception &e) { std::cout << Muesli::proc_id << ": " << e << std::endl; }

inline void barrier() {}

template <typename T>
T getNegativeInfinity() {
    if (std::numeric_limits<T>::has_infinity) return -std::numeric_limits<T>::infinity();
    return std::numeric_limits<T>::min();
}
template <typename T>
T getPositiveInfinity() {
    if (std::numeric_limits<T>::has_infinity) return std::numeric_limits<T>::infinity();
    return std::numeric_limits<T>::max();
}

template <typename C1, typename C2>
inline C1 proj1_2(C1 a, C2) {
    return a;
}
template <typename C1, typename C2>
inline C2 proj2_2(C1, C2 b) {
    return b;
}

namespace detail {
// Text of the containers' show(): "descr: " and a line break when descr is
// given, then "[v0 v1 ... vlast]", a line break, and an empty line.  After
// every row_len values the separator is "\n " instead of " ", and after
// every plane_len values "\n-------------\n " (0 disables either) -- the
// layout of src/ds.cpp:511-534, src/da.cpp:136-158, src/dm.cpp:266-296 and
// src/dc.cpp:284-319, so that dumps of the two implementations diff.
template <typename T>
std::string show_text(const std::string &descr, const T *v, long count, long row_len, long plane_len) {
    std::ostringstream s;
    if (!descr.empty()) s << descr << ": " << std::endl;
    s << "[";
    for (long i = 0; i < count; i++) {
        s << v[i];
        if (i + 1 == count) break;
        if (plane_len > 0 && (i + 1) % plane_len == 0)
            s << "\n-------------\n ";
        else if (row_len > 0 && (i + 1) % row_len == 0)
            s << "\n ";
        else
            s << " ";
    }
    s << "]" << std::endl << std::endl;
    return s.str();
}

// Text of showLocal(): "descr: [v0 v1 ... vlast ]" and a line break
// (src/ds.cpp:493-508).
template <typename T>
std::string show_local_text(const std::string &descr, const T *v, long count) {
    std::ostringstream s;
    if (!descr.empty()) s << descr << ": ";
    s << "[";
    for (long i = 0; i < count; i++) s << v[i] << " ";
    s << "]" << std::endl;
    return s.str();
}
}  // namespace detail

template <typename T>
inline void show(T *a, int size) {
    std::ostringstream s;
    s << "[";
    for (int i = 0; i < size; i++) s << a[i] << (i + 1 < size ? " " : "");
    s << "]" << std::endl;
    printv("%s", s.str().c_str());
}

// ---- per-thread random numbers for user functors (reference
// include/muesli.h:609-641) ---------------------------------------------------
// One counter-based SplitMix64 stream per (seed, index), the same on the
// host and the device: generateRandomState(seed, i) starts the stream of
// element i, randDouble draws uniformly from [min, max) and randInt from
// [min, max].  (The reference draws from curand on the device and from the
// unseeded host rand() on the CPU; neither sequence is reproduced here.)
struct RandomState {
    unsigned long long s;
};
typedef RandomState MSL_RANDOM_STATE;

namespace detail {
MSL_USERFUNC inline unsigned long long splitmix_next(RandomState &st) {
    unsigned long long z = (st.s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace detail

MSL_USERFUNC inline MSL_RANDOM_STATE generateRandomState(size_t seed, size_t someIndex) {
    RandomState st{(unsigned long long)seed};
    st.s ^= (unsigned long long)someIndex * 0xD1B54A32D192ED03ull;  // decorrelate neighbouring indices
    detail::splitmix_next(st);
    return st;
}

// uniform in [0, 1) with 53 random bits
MSL_USERFUNC inline double randUnit(MSL_RANDOM_STATE &state) {
    return (double)(detail::splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

MSL_USERFUNC inline double randDouble(double min, double max, MSL_RANDOM_STATE &state) {
    return min + randUnit(state) * (max - min);
}

MSL_USERFUNC inline int randInt(int min, int max, MSL_RANDOM_STATE &state) {
    const int v = min + (int)(randUnit(state) * ((double)max - (double)min + 1.0));
    return v > max ? max : v;
}

}  // namespace msl

/*
 * skeleton_bench.cpp -- throughput of the C++ skeleton API (msl::DA/DM/DC
 * with user functors) on the workloads of the reference's example programs,
 * beside the library's hand-tuned kernels for the same operations.
 *
 * Functors are written the way the reference examples write them:
 *   heat3D      examples/heatdiffusion.cpp:55-65 through DC::mapStencil<f>;
 *   GoLFunctor  examples/gameoflife.cpp:18-43 through DM::mapStencilMM;
 *   a mean blur of radius 1 through DM::mapStencil<f> (NPLMatrix);
 *   map / zip / fold functors as in examples/da_test.cpp.
 * Prints one JSON object; rates count 8 bytes per cell per stencil step
 * (one read, one write), 8 B/element for map, 12 for zip, 4 for fold.
 * Times are wall clock over back-to-back calls, except *_device: HIP events
 * on the library stream around one call (the fold's wall time also holds
 * the host round trip of its result), and *_kernels: HIP events around the
 * fold's two kernels alone.
 */
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "dc.h"
#include "dm.h"
#include "functors.h"

using namespace msl;

MSL_USERFUNC float heat3D(const PLCube<float> &c, int x, int y, int z) {
    float v = 0;
    v += c(x - 1, y, z);
    v += c(x + 1, y, z);
    v += c(x, y - 1, z);
    v += c(x, y + 1, z);
    v += c(x, y, z - 1);
    v += c(x, y, z + 1);
    v /= 6;
    return v;
}

MSL_USERFUNC float blur1(const NPLMatrix<float> &m, int x, int y) {
    float total = 0.0f;
    for (int dy = -1; dy <= 1; dy++) {
        float row = 0.0f;
        for (int dx = -1; dx <= 1; dx++) row += m(x + dx, y + dy);
        total += row;
    }
    return total / 9.0f;
}

class GoLFunctor : public DMMapStencilFunctor<int, int> {
public:
    MSL_USERFUNC int operator()(int row, int col, PLMatrix<int> *in, int, int) const {
        const int s = in->get(row - 1, col) + in->get(row - 1, col - 1) + in->get(row - 1, col + 1) +
                      in->get(row + 1, col) + in->get(row + 1, col - 1) + in->get(row + 1, col + 1) +
                      in->get(row, col - 1) + in->get(row, col + 1);
        const int live = in->get(row, col);
        int next = 0;
        if (live == 1 && (s == 2 || s == 3)) next = 1;
        if (live == 0 && s == 3) next = 1;
        return next;
    }
};

class Affine : public Functor<float, float> {
public:
    MSL_USERFUNC float operator()(float x) const override { return x * 1.5f + 3.f; }
};
class Add : public Functor2<float, float, float> {
public:
    MSL_USERFUNC float operator()(float x, float y) const override { return x + y; }
};

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device time of one call (HIP events on the library stream around it),
// median of reps: for calls that end with a host copy (fold), whose wall
// time also holds the host round trip.
template <typename F>
static double device_timed(F f, int reps) {
    f();
    std::vector<float> ms;
    for (int i = 0; i < reps; i++) {
        float t = 0.0f;
        msl_timer_start();
        f();
        msl_timer_stop(&t);
        ms.push_back(t);
    }
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2] * 1e-3;
}

template <typename F>
static double timed(F f, int reps) {
    f();  // warm-up: code object, LDS configuration
    msl_synchronize();
    const double t0 = now();
    for (int i = 0; i < reps; i++) f();
    msl_synchronize();
    return (now() - t0) / reps;
}

int main(int argc, char **argv) {
    msl::initSkeletons(argc, argv);
    std::vector<std::string> out;
    auto rate = [&](const char *name, double bytes, double t) {
        char buf[256];
        snprintf(buf, sizeof buf, "\"%s\": {\"ms\": %.4f, \"GBs\": %.1f}", name, t * 1e3, bytes / t / 1e9);
        out.push_back(buf);
    };

    {  // DC<float> 512^3 heat: user functor vs library kernel
        const int n = 512;
        const double cells = (double)n * n * n;
        DC<float> a(n, n, n), b(n, n, n, 0.0f);
        msl_fill_uniform_f32(a.getUserFunctionData(), (size_t)cells, 1, 1.0f / 16777216.0f);
        a.setCpuMemoryInSync(false);
        rate("dc_heat3D_functor_512", 8 * cells, timed([&] {
                 a.mapStencil<heat3D>(b, 1, 0.0f);
                 b.mapStencil<heat3D>(a, 1, 0.0f);
             }, 5) / 2);
        rate("dc_heat7_library_512", 8 * cells, timed([&] {
                 a.heat7(b, 0.0f);
                 b.heat7(a, 0.0f);
             }, 5) / 2);
    }
    {  // DM<int> 1024^2 Game of Life: legacy functor, library kernel
        const int n = 1024;
        const double cells = (double)n * n;
        DM<int> a(n, n), b(n, n, 0);
        msl_fill_bernoulli_i32(reinterpret_cast<int32_t *>(a.getUserFunctionData()), (size_t)cells, 2,
                               (uint64_t)(0.3 * 18446744073709551616.0));
        a.setCpuMemoryInSync(false);
        GoLFunctor gol;
        rate("dm_gol_mapStencilMM_functor_1024", 8 * cells, timed([&] {
                 a.mapStencilMM(b, gol, 0);
                 b.mapStencilMM(a, gol, 0);
             }, 50) / 2);
        rate("dm_gol_library_1024", 8 * cells, timed([&] {
                 a.gameOfLife(b);
                 b.gameOfLife(a);
             }, 50) / 2);
    }
    {  // DM<float> 8192^2: blur r=1 functor vs library; map / zip / fold functors
        const int n = 8192;
        const double cells = (double)n * n;
        DM<float> a(n, n), b(n, n, 0.0f), c(n, n, 0.0f);
        msl_fill_uniform_f32(a.getUserFunctionData(), (size_t)cells, 3, 255.0f / 16777216.0f);
        a.setCpuMemoryInSync(false);
        rate("dm_blur_r1_functor_8192", 8 * cells, timed([&] {
                 a.mapStencil<blur1>(b, 1, 0.0f);
                 b.mapStencil<blur1>(a, 1, 0.0f);
             }, 5) / 2);
        rate("dm_blur_r1_library_8192", 8 * cells, timed([&] {
                 a.mapStencilMeanBlur(b, 1, 0.0f);
                 b.mapStencilMeanBlur(a, 1, 0.0f);
             }, 5) / 2);
        Affine aff;
        Add add;
        rate("dm_map_functor_8192", 8 * cells, timed([&] { c.map(aff, a); }, 10));
        rate("dm_zip_functor_8192", 12 * cells, timed([&] { c.zip(a, b, add); }, 10));
        rate("dm_fold_functor_8192", 4 * cells, timed([&] { (void)a.fold(add); }, 10));
        rate("dm_fold_functor_8192_device", 4 * cells, device_timed([&] { (void)a.fold(add); }, 10));
        {  // the two kernels DS::fold launches, with the result left in device memory
            const float *d = a.getUserFunctionData();
            const int blocks = detail::stream_grid((size_t)cells, 2);
            float *part = nullptr;
            int *has = nullptr;
            msl_malloc((void **)&part, sizeof(float) * (blocks + 1));
            msl_malloc((void **)&has, sizeof(int) * (blocks + 1));
            rate("dm_fold_functor_8192_kernels", 4 * cells, device_timed([&] {
                     detail::fold_stage1_kernel<float, Add><<<blocks, detail::kThreads, 0, detail::stream()>>>(
                         d, (size_t)cells, part, has, add, detail::aligned16(d));
                     detail::fold_stage2_kernel<float, Add><<<1, detail::kThreads, 0, detail::stream()>>>(
                         part, has, blocks, part + blocks, has + blocks, add);
                 }, 10));
            msl_synchronize();
            msl_free(part);
            msl_free(has);
        }
    }
    printf("{");
    for (size_t i = 0; i < out.size(); i++) printf("%s%s", i ? ", " : "", out[i].c_str());
    printf("}\n");
    msl::terminateSkeletons();
    return 0;
}

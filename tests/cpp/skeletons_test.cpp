/*
 * skeletons_test.cpp -- drives the msl::DA / DM / DC class templates (the
 * C++ skeleton API compiled by hipcc over libmsl_hip.so) the way user
 * programs do: functor classes, function-pointer stencils, extended lambdas.
 *
 * Usage: skeletons_test <dir>
 *   Reads the seeded inputs written by tests/test_gpu_cpp_api.py into <dir>
 *   and writes each skeleton's output next to them; the Python side compares
 *   the outputs with the CPU oracle.  Closed-form checks (the expectations of
 *   the reference's own example programs, examples/da_test.cpp,
 *   examples/dm_test.cpp, examples/dc_test.cpp, examples/ds_test.cpp and the
 *   all-ones addStencil of examples/dc_map_stencil_test.cpp) run in-process;
 *   every failure prints "FAIL <name>" and makes the exit status 1.
 *
 * File format: int32 dims[3] (x, y, z; unused = 1) followed by the elements.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "da.h"
#include "dc.h"
#include "dm.h"
#include "functors.h"

using namespace msl;

static int g_failures = 0;

static void check(bool ok, const std::string &name) {
    printf("%s %s\n", ok ? "OK  " : "FAIL", name.c_str());
    if (!ok) g_failures++;
}

template <typename T>
static std::vector<T> read_grid(const std::string &path, int dims[3]) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path.c_str());
        std::exit(2);
    }
    f.read(reinterpret_cast<char *>(dims), 3 * sizeof(int32_t));
    std::vector<T> v((size_t)dims[0] * dims[1] * dims[2]);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) {
        fprintf(stderr, "short read %s\n", path.c_str());
        std::exit(2);
    }
    return v;
}

template <typename T>
static void write_grid(const std::string &path, const T *data, int x, int y, int z) {
    std::ofstream f(path, std::ios::binary);
    const int32_t dims[3] = {x, y, z};
    f.write(reinterpret_cast<const char *>(dims), sizeof(dims));
    f.write(reinterpret_cast<const char *>(data), (std::streamsize)((size_t)x * y * z * sizeof(T)));
}

// ---- user functors, in the styles the reference examples use ----------------
class Triple : public Functor<int, int> {
public:
    MSL_USERFUNC int operator()(int x) const override { return 3 * x; }
};
class Add : public Functor2<int, int, int> {
public:
    MSL_USERFUNC int operator()(int x, int y) const override { return x + y; }
};
class Max : public Functor2<float, float, float> {
public:
    MSL_USERFUNC float operator()(float x, float y) const override { return x > y ? x : y; }
};
class Sum3 : public Functor3<int, int, int, int> {
public:
    MSL_USERFUNC int operator()(int x, int y, int z) const override { return x + y + z; }
};
class IndexPlus : public Functor2<int, int, int> {  // DA mapIndex: f(i, v)
public:
    MSL_USERFUNC int operator()(int i, int v) const override { return i + v; }
};
class IndexMulZip : public Functor3<int, int, int, int> {  // DA zipIndex: f(i, a, b)
public:
    MSL_USERFUNC int operator()(int i, int a, int b) const override { return i * a + b; }
};
class RowColLinear : public Functor3<int, int, float, float> {  // DM mapIndex: f(row, col, v)
public:
    explicit RowColLinear(int cols) : cols(cols) {}
    MSL_USERFUNC float operator()(int row, int col, float) const override { return (float)(row * cols + col); }

private:
    int cols;
};
class RowColZip : public Functor4<int, int, int, int, int> {  // DM zipIndex: f(row, col, a, b)
public:
    MSL_USERFUNC int operator()(int row, int col, int a, int b) const override { return row * 1000 + col + a - b; }
};
class CubeIndex : public Functor4<int, int, int, int, int> {  // DC mapIndex: f(col, row, depth, v)
public:
    MSL_USERFUNC int operator()(int x, int y, int z, int v) const override { return x + 100 * y + 10000 * z + v; }
};
class CubeZip : public Functor5<int, int, int, int, int, int> {  // DC zipIndex: f(col, row, depth, a, b)
public:
    MSL_USERFUNC int operator()(int x, int y, int z, int a, int b) const override { return (x + y + z) * a + b; }
};

// ---- functors of the extra DM / DA skeletons (the forms the ACO examples use)
class Zip3Idx : public Functor5<int, int, int, int, int, int> {  // f(row, col, a, b, c)
public:
    MSL_USERFUNC int operator()(int row, int col, int a, int b, int c) const override {
        return a + 2 * b + 3 * c + row * 1000 + col;
    }
};
class RowOperand : public Functor4<int, int, int, double, int> {  // f(row, col, a, v[row])
public:
    MSL_USERFUNC int operator()(int row, int col, int a, double v) const override { return a + (int)v * 10 + col; }
};
class MatRowOperand : public Functor5<int, int, int, int, double, int> {  // f(row, col, a, b, v[row])
public:
    MSL_USERFUNC int operator()(int row, int col, int a, int b, double v) const override {
        return a * b + (int)v + row;
    }
};
class ThreeRowOperands : public Functor6<int, int, int, int, double, float, int> {  // f(row, col, a, u, v, w)
public:
    MSL_USERFUNC int operator()(int row, int col, int a, int u, double v, float w) const override {
        return a + u * 100 + (int)v * 10 + (int)w + col;
    }
};
class CrossTranspose {  // f(row, col, a, whole): reads the transposed element
public:
    explicit CrossTranspose(int n) : n(n) {}
    MSL_USERFUNC int operator()(int row, int col, int a, int *whole) const { return a + whole[col * n + row]; }

private:
    int n;
};
class RowReduceSum : public Functor2<int, int, int> {
public:
    MSL_USERFUNC int operator()(int acc, int x) const override { return acc + x; }
};
class RowReduceMax : public Functor2<double, double, double> {
public:
    MSL_USERFUNC double operator()(double acc, double x) const override { return x > acc ? x : acc; }
};
class PairProduct : public Functor3<int, int, long long, long long> {  // f(x, next, acc)
public:
    MSL_USERFUNC long long operator()(int x, int nx, long long acc) const override {
        return acc * 3 + (long long)x * nx;
    }
};
class RowEnds {  // DA: f(i, &rows[i][0], v)
public:
    explicit RowEnds(int cols) : cols(cols) {}
    MSL_USERFUNC int operator()(int i, int *row, int v) const { return v + row[0] + row[cols - 1] + i; }

private:
    int cols;
};
class RowDot {  // DA: f(i, &rows[i][0], b[i], v)
public:
    explicit RowDot(int cols) : cols(cols) {}
    MSL_USERFUNC int operator()(int i, int *row, int b, int v) const {
        int s = 0;
        for (int c = 0; c < cols; c++) s += row[c] * (c + 1);
        return s * b + v - i;
    }

private:
    int cols;
};
class DrawInts {  // per-element random draws: f(i, v)
public:
    MSL_USERFUNC int operator()(int i, int) const {
        MSL_RANDOM_STATE st = generateRandomState(1234, i);
        int acc = 0;
        for (int k = 0; k < 5; k++) acc = acc * 31 + randInt(3, 17, st);
        return acc;
    }
};
class DrawDoubles {
public:
    MSL_USERFUNC double operator()(int i, double) const {
        MSL_RANDOM_STATE st = generateRandomState(99, i);
        return randDouble(-2.0, 5.0, st);
    }
};

// Game of Life in the legacy mapStencilMM form (examples/gameoflife.cpp:18-43).
class LifeMM : public DMMapStencilFunctor<int, int> {
public:
    MSL_USERFUNC int operator()(int row, int col, PLMatrix<int> *in, int, int) const {
        const int s = in->get(row - 1, col - 1) + in->get(row - 1, col) + in->get(row - 1, col + 1) +
                      in->get(row, col - 1) + in->get(row, col + 1) + in->get(row + 1, col - 1) +
                      in->get(row + 1, col) + in->get(row + 1, col + 1);
        const int alive = in->get(row, col);
        return (alive == 1 && (s == 2 || s == 3)) || (alive == 0 && s == 3) ? 1 : 0;
    }
};

// Border values for the neutral-value-functor form: nvf(row, col).
class BorderValue : public Functor2<int, int, float> {
public:
    MSL_USERFUNC float operator()(int row, int col) const override { return (float)(row * 7 + col * 3); }
};
// Four-neighbour sum (the Jacobi sweep shape, examples/jacobisolver.cpp:46-63)
// with radius 2 so the border band is two cells wide.
class Cross2 : public DMNVFMapStencilFunctor<float, float, BorderValue> {
public:
    Cross2() { this->setStencilSize(2); }
    MSL_USERFUNC float operator()(int row, int col, PLMatrix<float> *in, int, int) const {
        float s = 0.0f;
        s += in->get(row + 2, col);
        s += in->get(row - 2, col);
        s += in->get(row, col + 1);
        s += in->get(row, col - 2);
        return s;
    }
};

// Function-pointer stencils.
MSL_USERFUNC int life_npl(const NPLMatrix<int> &m, int x, int y) {
    int s = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
            if (dx != 0 || dy != 0) s += m(x + dx, y + dy);
    const int alive = m(x, y);
    return (alive == 1 && (s == 2 || s == 3)) || (alive == 0 && s == 3) ? 1 : 0;
}
// Mean blur of radius 3: row sums first (each from 0), then their sum,
// then one division by 49 -- the order of the library's mean-blur kernel.
MSL_USERFUNC float blur3_npl(const NPLMatrix<float> &m, int x, int y) {
    float total = 0.0f;
    for (int dy = -3; dy <= 3; dy++) {
        float row = 0.0f;
        for (int dx = -3; dx <= 3; dx++) row += m(x + dx, y + dy);
        total += row;
    }
    return total / 49.0f;
}
MSL_USERFUNC float heat_plc(const PLCube<float> &c, int x, int y, int z) {
    float v = 0.0f;
    v += c(x - 1, y, z);
    v += c(x + 1, y, z);
    v += c(x, y - 1, z);
    v += c(x, y + 1, z);
    v += c(x, y, z - 1);
    v += c(x, y, z + 1);
    v /= 6;
    return v;
}
// Box sum of radius 2, dx outermost, then dy, then dz (the loop order of
// examples/dc_map_stencil_test.cpp:78-88).
MSL_USERFUNC float box2_plc(const PLCube<float> &c, int x, int y, int z) {
    float v = 0.0f;
    for (int dx = -2; dx <= 2; dx++)
        for (int dy = -2; dy <= 2; dy++)
            for (int dz = -2; dz <= 2; dz++) v += c(x + dx, y + dy, z + dz);
    return v;
}
MSL_USERFUNC float box1_plc(const PLCube<float> &c, int x, int y, int z) {
    float v = 0.0f;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) v += c(x + dx, y + dy, z + dz);
    return v;
}

// Reads two cells away under a declared radius of 1: outside the staged
// window the views fall back to global memory (or nv off the grid).
MSL_USERFUNC float reach2_npl(const NPLMatrix<float> &m, int x, int y) {
    float v = m(x - 2, y);
    v += m(x + 2, y) * 2.0f;
    v += m(x, y - 2) * 3.0f;
    v += m(x, y + 2) * 4.0f;
    v -= m(x + 1, y + 1);
    return v;
}
MSL_USERFUNC float reach2_plc(const PLCube<float> &c, int x, int y, int z) {
    float v = c(x + 2, y, z);
    v += c(x, y - 2, z) * 2.0f;
    v += c(x, y, z + 2) * 3.0f;
    v += c(x - 1, y + 1, z - 1);
    return v;
}

// ---- closed-form checks --------------------------------------------------------
static void closed_forms() {
    const int n = 1000;
    // DA map / zip / fold / index skeletons (the expectations of da_test.cpp).
    DA<int> a(n, 2), b(n, 5), c(n, 0);
    Triple triple;
    Add add;
    a.mapInPlace(triple);  // 6
    check(a.get(0) == 6 && a.get(n - 1) == 6, "DA mapInPlace");
    c.map(triple, b);  // 15
    int *g = c.gather();
    bool ok = true;
    for (int i = 0; i < n; i++) ok &= g[i] == 15;
    check(ok, "DA map");
    c.zip(a, b, add);  // 11
    g = c.gather();
    ok = true;
    for (int i = 0; i < n; i++) ok &= g[i] == 11;
    check(ok, "DA zip");
    c.zipInPlace(a, add);  // 17
    check(c.fold(add, true) == 17 * n, "DA zipInPlace + fold");
    Sum3 sum3;
    c.zipInPlace3(a, b, sum3);  // 17+6+5 = 28
    check(c.get(n / 2) == 28, "DA zipInPlace3");
    IndexPlus ip;
    c.mapIndexInPlace(ip);  // 28 + i
    g = c.gather();
    ok = true;
    for (int i = 0; i < n; i++) ok &= g[i] == 28 + i;
    check(ok, "DA mapIndexInPlace");
    c.mapIndex(ip, b);  // i + 5
    g = c.gather();
    ok = true;
    for (int i = 0; i < n; i++) ok &= g[i] == 5 + i;
    check(ok, "DA mapIndex");
    // getLocalPartition refreshes the host mirror after a device skeleton
    // (src/ds.cpp:263-267)
    c.mapInPlace(triple);  // 3 * (i + 5)
    int *lp = c.getLocalPartition();
    ok = true;
    for (int i = 0; i < n; i++) ok &= lp[i] == 3 * (5 + i);
    check(ok, "DA getLocalPartition after mapInPlace");
    c.mapIndex(ip, b);  // i + 5 again
    IndexMulZip iz;
    c.zipIndexInPlace(b, iz);  // i*(5+i) + 5
    g = c.gather();
    ok = true;
    for (int i = 0; i < n; i++) ok &= g[i] == i * (5 + i) + 5;
    check(ok, "DA zipIndexInPlace");
    c.zipIndex(a, b, iz);  // i*6 + 5
    g = c.gather();
    ok = true;
    for (int i = 0; i < n; i++) ok &= g[i] == 6 * i + 5;
    check(ok, "DA zipIndex");
    // extended lambda, as examples/da_test.cpp:97 passes one
    auto twice = [] __host__ __device__(int x) { return 2 * x; };
    c.mapInPlace(twice);
    check(c.get(10) == 2 * (6 * 10 + 5), "DA lambda map");
    // set() then a skeleton sees the new value; get() after a skeleton sees it
    c.set(3, 1000);
    c.mapInPlace(triple);
    check(c.get(3) == 3000 && c.get(4) == 3 * 2 * 29, "DA set/get residency");

    // fold of a ragged length (not a multiple of the block) and of one element
    DA<int> odd(777, 1);
    check(odd.fold(add) == 777, "DA fold ragged");
    DA<int> one(1, 42);
    check(one.fold(add) == 42, "DA fold single");
    DA<float> mx(12345, 0.5f);
    mx.set(12000, 9.25f);
    Max fmax;
    check(mx.fold(fmax) == 9.25f, "DA fold max");
    // many folds in a row on two large containers, each result read from the
    // pinned result slot right after the previous one: integer sums exact
    {
        const int big = (1 << 24) + 13;
        DA<int> u(big, 1), w(big, 0);
        w.mapIndexInPlace(ip);  // i
        w.mapInPlace(triple);   // 3i
        const int expect_u = big;
        // 3 * big * (big - 1) / 2 modulo 2^32, as the device's wrapping int adds give it
        const int expect_w = (int)(unsigned)(3ull * (unsigned long long)big * (unsigned long long)(big - 1) / 2ull);
        bool all = true;
        for (int k = 0; k < 40; k++) {
            all &= u.fold(add) == expect_u;
            all &= w.fold(add) == expect_w;
        }
        check(all, "DA fold: 80 folds in a row, exact");
    }

    // DM index skeletons (dm_test.cpp) and rotations
    const int rows = 37, cols = 53;
    DM<float> lin(rows, cols, 0.0f);
    RowColLinear rcl(cols);
    lin.mapIndexInPlace(rcl);
    ok = true;
    for (int r = 0; r < rows; r++)
        for (int q = 0; q < cols; q++) ok &= lin.get2D(r, q) == (float)(r * cols + q);
    check(ok, "DM mapIndexInPlace (row, col)");
    DM<int> m1(rows, cols, 4), m2(rows, cols, 1), m3(rows, cols, 0);
    RowColZip rz;
    m3.zipIndex(m1, m2, rz);
    ok = true;
    for (int r = 0; r < rows; r++)
        for (int q = 0; q < cols; q++) ok &= m3.get2D(r, q) == r * 1000 + q + 3;
    check(ok, "DM zipIndex (row, col)");

    // DC index skeletons (dc_test.cpp): f(col, row, depth, v)
    const int cr = 9, cc = 11, cd = 7;
    DC<int> cube(cr, cc, cd, 1), cube2(cr, cc, cd, 2), cube3(cr, cc, cd, 0);
    CubeIndex ci;
    cube.mapIndexInPlace(ci);
    ok = true;
    for (int z = 0; z < cd; z++)
        for (int y = 0; y < cr; y++)
            for (int x = 0; x < cc; x++) ok &= cube.get(y, x, z) == x + 100 * y + 10000 * z + 1;
    check(ok, "DC mapIndexInPlace (col, row, depth)");
    CubeZip cz;
    cube3.zipIndex(cube, cube2, cz);
    ok = true;
    for (int z = 0; z < cd; z++)
        for (int y = 0; y < cr; y++)
            for (int x = 0; x < cc; x++) ok &= cube3.get(y, x, z) == (x + y + z) * cube.get(y, x, z) + 2;
    check(ok, "DC zipIndex");

    // all-ones 3-D box stencil of radius 1: 27 inside, 18 on faces, 12 on
    // edges, 8 at corners (nv = 0)
    DC<float> ones(10, 12, 9, 1.0f), boxed(10, 12, 9, 0.0f);
    ones.mapStencil<box1_plc>(boxed, 1, 0.0f);
    ok = true;
    for (int z = 0; z < 9; z++)
        for (int y = 0; y < 10; y++)
            for (int x = 0; x < 12; x++) {
                const int faces = (x == 0 || x == 11) + (y == 0 || y == 9) + (z == 0 || z == 8);
                const float want = faces == 0 ? 27.f : faces == 1 ? 18.f : faces == 2 ? 12.f : 8.f;
                ok &= boxed.get(y, x, z) == want;
            }
    check(ok, "DC mapStencil all-ones 27/18/12/8");

    // empty containers are legal and skeletons on them do nothing
    DA<int> empty(0);
    empty.mapInPlace(triple);
    check(empty.getSize() == 0, "DA empty");
}

// ---- the extra DM / DA skeletons and the random-number API -------------------
static void extra_skeletons() {
    const int rows = 23, cols = 67;
    DM<int> a(rows, cols), b(rows, cols), c(rows, cols);
    std::vector<int> ha(rows * cols), hb(rows * cols), hc(rows * cols);
    for (int k = 0; k < rows * cols; k++) {
        ha[k] = (k * 7) % 13 - 6;
        hb[k] = (k * 5) % 11;
        hc[k] = k % 9;
    }
    a.fill(ha.data());
    b.fill(hb.data());
    c.fill(hc.data());
    DA<double> vd(rows);
    DA<int> vi(rows);
    DA<float> vf(rows);
    for (int r = 0; r < rows; r++) {
        vd.set(r, r * 0.5 + 1.0);
        vi.set(r, r - 4);
        vf.set(r, (float)(r % 5));
    }

    Zip3Idx z3;
    a.zipIndexInPlace3(b, c, z3);
    bool ok = true;
    for (int r = 0; r < rows; r++)
        for (int q = 0; q < cols; q++) {
            const int k = r * cols + q;
            ok &= a.get2D(r, q) == ha[k] + 2 * hb[k] + 3 * hc[k] + r * 1000 + q;
            ha[k] = ha[k] + 2 * hb[k] + 3 * hc[k] + r * 1000 + q;
        }
    check(ok, "DM zipIndexInPlace3");

    MatRowOperand mro;
    a.zipIndexInPlaceMA(b, vd, mro);
    ok = true;
    for (int r = 0; r < rows; r++)
        for (int q = 0; q < cols; q++) {
            const int k = r * cols + q;
            ha[k] = ha[k] * hb[k] + (int)(r * 0.5 + 1.0) + r;
            ok &= a.get2D(r, q) == ha[k];
        }
    check(ok, "DM zipIndexInPlaceMA");

    RowOperand ro;
    a.zipIndexInPlaceA(vd, ro);
    ok = true;
    for (int r = 0; r < rows; r++)
        for (int q = 0; q < cols; q++) {
            const int k = r * cols + q;
            ha[k] = ha[k] + (int)(r * 0.5 + 1.0) * 10 + q;
            ok &= a.get2D(r, q) == ha[k];
        }
    check(ok, "DM zipIndexInPlaceA");

    ThreeRowOperands tro;
    a.zipIndexInPlaceAAA(vi, vd, vf, tro);
    ok = true;
    for (int r = 0; r < rows; r++)
        for (int q = 0; q < cols; q++) {
            const int k = r * cols + q;
            ha[k] = ha[k] + (r - 4) * 100 + (int)(r * 0.5 + 1.0) * 10 + (r % 5) + q;
            ok &= a.get2D(r, q) == ha[k];
        }
    check(ok, "DM zipIndexInPlaceAAA");

    bool threw = false;
    try {
        DA<double> shortv(rows - 1);
        a.zipIndexInPlaceA(shortv, ro);
    } catch (const detail::SizeMismatchException &) {
        threw = true;
    }
    check(threw, "DM zipIndexInPlaceA size mismatch");

    // crossZipIndexInPlace on a square matrix: this += transpose(b)
    const int sq = 40;
    DM<int> s1(sq, sq), s2(sq, sq);
    std::vector<int> h1(sq * sq), h2(sq * sq);
    for (int k = 0; k < sq * sq; k++) {
        h1[k] = k;
        h2[k] = 3 * k + 1;
    }
    s1.fill(h1.data());
    s2.fill(h2.data());
    CrossTranspose ct(sq);
    s1.crossZipIndexInPlace(s2, ct);
    ok = true;
    for (int r = 0; r < sq; r++)
        for (int q = 0; q < sq; q++) ok &= s1.get2D(r, q) == h1[r * sq + q] + h2[q * sq + r];
    check(ok, "DM crossZipIndexInPlace");

    // reduceRows: exact int sums and double max over rows of 1, 63, 64, 65 and
    // 300 columns (fewer, equal and more columns than a wave's lanes)
    for (int rc : {1, 63, 64, 65, 300}) {
        const int rr = 37;
        DM<int> m(rr, rc);
        DM<double> md(rr, rc);
        std::vector<int> hm(rr * rc);
        std::vector<double> hd(rr * rc);
        for (int k = 0; k < rr * rc; k++) {
            hm[k] = (k * 13) % 101 - 50;
            hd[k] = -1000.0 + ((k * 37) % 997) * 0.25;
        }
        m.fill(hm.data());
        md.fill(hd.data());
        DA<int> sums(rr);
        DA<double> maxes(rr);
        RowReduceSum rs;
        RowReduceMax rm;
        m.reduceRows(sums, rs);
        md.reduceRows(maxes, rm);
        ok = true;
        for (int r = 0; r < rr; r++) {
            int want = 0;
            double wmax = hd[r * rc];
            for (int q = 0; q < rc; q++) {
                want += hm[r * rc + q];
                wmax = hd[r * rc + q] > wmax ? hd[r * rc + q] : wmax;
            }
            ok &= sums.get(r) == want && maxes.get(r) == wmax;
        }
        check(ok, "DM reduceRows cols=" + std::to_string(rc));
    }

    // reducetwoColumns: sequential f(x, next, acc) from 0, wrapping to column 0
    {
        const int rr = 9, rc = 6;
        DM<int> m(rr, rc);
        std::vector<int> hm(rr * rc);
        for (int k = 0; k < rr * rc; k++) hm[k] = (k % 7) - 3;
        m.fill(hm.data());
        DA<long long> out(rr);
        PairProduct pp;
        m.reducetwoColumns(out, pp);
        ok = true;
        for (int r = 0; r < rr; r++) {
            long long acc = 0;
            for (int q = 0; q < rc; q++) acc = acc * 3 + (long long)hm[r * rc + q] * hm[r * rc + (q + 1) % rc];
            ok &= out.get(r) == acc;
        }
        check(ok, "DM reducetwoColumns");
    }

    // setColumn and its argument checks
    {
        DM<int> m(rows, cols, 7);
        DA<int> col(rows);
        for (int r = 0; r < rows; r++) col.set(r, 100 + r);
        m.setColumn(col, 5);
        ok = true;
        for (int r = 0; r < rows; r++)
            for (int q = 0; q < cols; q++) ok &= m.get2D(r, q) == (q == 5 ? 100 + r : 7);
        check(ok, "DM setColumn");
        bool bad_col = false, bad_len = false;
        try {
            m.setColumn(col, cols);
        } catch (const detail::IndexOutOfBoundsException &) {
            bad_col = true;
        }
        try {
            DA<int> shortc(rows - 2);
            m.setColumn(shortc, 0);
        } catch (const detail::SizeMismatchException &) {
            bad_len = true;
        }
        check(bad_col && bad_len, "DM setColumn argument checks");
    }

    // DA with matrix rows
    {
        const int na = 30, mc = 11;
        DM<int> m(na + 3, mc);  // more rows than array elements is allowed
        std::vector<int> hm((na + 3) * mc);
        for (int k = 0; k < (na + 3) * mc; k++) hm[k] = (k * 3) % 17;
        m.fill(hm.data());
        DA<int> x(na, 5), y(na);
        for (int i = 0; i < na; i++) y.set(i, i % 4);
        RowEnds re(mc);
        x.mapIndexInPlaceDMRows(m, re);
        ok = true;
        std::vector<int> hx(na);
        for (int i = 0; i < na; i++) {
            hx[i] = 5 + hm[i * mc] + hm[i * mc + mc - 1] + i;
            ok &= x.get(i) == hx[i];
        }
        check(ok, "DA mapIndexInPlaceDMRows");
        RowDot rd(mc);
        x.zipIndexInPlaceDMRows(y, m, rd);
        ok = true;
        for (int i = 0; i < na; i++) {
            int sdot = 0;
            for (int q = 0; q < mc; q++) sdot += hm[i * mc + q] * (q + 1);
            ok &= x.get(i) == sdot * (i % 4) + hx[i] - i;
        }
        check(ok, "DA zipIndexInPlaceDMRows");
    }

    // random numbers: the device draws equal the host draws, and stay in range
    {
        const int nr = 5000;
        DA<int> ri(nr, 0);
        DA<double> rdv(nr, 0.0);
        DrawInts di;
        DrawDoubles dd;
        ri.mapIndexInPlace(di);
        rdv.mapIndexInPlace(dd);
        ok = true;
        bool in_range = true;
        int lo = 100, hi = -1;
        for (int i = 0; i < nr; i++) {
            ok &= ri.get(i) == di(i, 0) && rdv.get(i) == dd(i, 0.0);
            MSL_RANDOM_STATE st = generateRandomState(1234, i);
            for (int k = 0; k < 5; k++) {
                const int v = randInt(3, 17, st);
                lo = v < lo ? v : lo;
                hi = v > hi ? v : hi;
            }
            in_range &= rdv.get(i) >= -2.0 && rdv.get(i) < 5.0;
        }
        check(ok, "random numbers: device == host");
        check(in_range && lo == 3 && hi == 17, "random numbers: ranges");
    }

    // show() layouts (compared with the reference's text by the Python side)
    {
        DA<int> sa(5);
        DM<int> sm(2, 3);
        DC<int> sc(2, 2, 2);
        for (int i = 0; i < 5; i++) sa.set(i, i * 10);
        for (int i = 0; i < 6; i++) sm.set(i, i + 1);
        for (int i = 0; i < 8; i++) sc.set(i, i);
        fflush(stdout);
        sa.show("sa");
        sa.show("", 3);
        sa.showLocal("sal");
        sm.show("sm");
        sc.show("sc");
        fflush(stdout);
    }
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    msl::initSkeletons(argc, argv);
    try {
        closed_forms();
        extra_skeletons();

        int d[3];
        // ---- Game of Life: legacy functor, function pointer, library kernel
        std::vector<int> life = read_grid<int>(dir + "/gol_in.bin", d);
        const int lc = d[0], lr = d[1], steps = 20;
        DM<int> g1(lr, lc), g2(lr, lc, 0);
        g1.fill(life.data());
        LifeMM lifemm;
        for (int s = 0; s < steps; s++) {
            if (s % 2 == 0) g1.mapStencilMM(g2, lifemm, 0);
            else g2.mapStencilMM(g1, lifemm, 0);
        }
        write_grid(dir + "/gol_mm.bin", g1.gather(), lc, lr, 1);
        DM<int> h1(lr, lc), h2(lr, lc, 0);
        h1.fill(life.data());
        for (int s = 0; s < steps; s++) {
            if (s % 2 == 0) h1.mapStencil<life_npl>(h2, 1, 0);
            else h2.mapStencil<life_npl>(h1, 1, 0);
        }
        write_grid(dir + "/gol_npl.bin", h1.gather(), lc, lr, 1);
        DM<int> k1(lr, lc), k2(lr, lc, 0);
        k1.fill(life.data());
        for (int s = 0; s < steps; s++) {
            if (s % 2 == 0) k1.gameOfLife(k2);
            else k2.gameOfLife(k1);
        }
        write_grid(dir + "/gol_fast.bin", k1.gather(), lc, lr, 1);
        Add add;
        check(k1.fold(add) == g1.fold(add), "GoL fold(sum) agrees across forms");

        // ---- 2-D mean blur r = 3, nv = 0.5: function pointer and library kernel
        std::vector<float> img = read_grid<float>(dir + "/blur_in.bin", d);
        const int bc = d[0], br = d[1];
        DM<float> src(br, bc), dst(br, bc, 0.0f), dst2(br, bc, 0.0f);
        src.fill(img.data());
        src.mapStencil<blur3_npl>(dst, 3, 0.5f);
        write_grid(dir + "/blur_npl.bin", dst.gather(), bc, br, 1);
        src.mapStencilMeanBlur(dst2, 3, 0.5f);
        write_grid(dir + "/blur_fast.bin", dst2.gather(), bc, br, 1);
        DM<float> reach(br, bc, 0.0f);
        src.mapStencil<reach2_npl>(reach, 1, 0.5f);
        write_grid(dir + "/reach_npl.bin", reach.gather(), bc, br, 1);

        // ---- legacy functor with a border-value functor, radius 2
        DM<float> cr(br, bc, 0.0f);
        BorderValue bv;
        Cross2 cross;
        src.mapStencilMM(cr, cross, bv);
        write_grid(dir + "/cross_nvf.bin", cr.gather(), bc, br, 1);

        // ---- rotations (dm_test.cpp:130-144)
        DM<float> rot(src);
        rot.rotateRows(3);
        rot.rotateCols(-5);
        write_grid(dir + "/rot_r3_cm5.bin", rot.gather(), bc, br, 1);
        rot.rotateRows(-7);
        rot.rotateCols(bc + 2);
        write_grid(dir + "/rot_rm7_c2.bin", rot.gather(), bc, br, 1);

        // ---- 3-D heat, 5 steps: function pointer and library kernel; box r = 2
        std::vector<float> cube = read_grid<float>(dir + "/heat_in.bin", d);
        const int nx = d[0], ny = d[1], nz = d[2];
        DC<float> c1(ny, nx, nz), c2(ny, nx, nz, 0.0f);
        c1.fill(cube.data());
        for (int s = 0; s < 5; s++) {
            if (s % 2 == 0) c1.mapStencil<heat_plc>(c2, 1, 0.0f);
            else c2.mapStencil<heat_plc>(c1, 1, 0.0f);
        }
        write_grid(dir + "/heat_plc.bin", c2.gather(), nx, ny, nz);
        DC<float> f1(ny, nx, nz), f2(ny, nx, nz, 0.0f);
        f1.fill(cube.data());
        for (int s = 0; s < 5; s++) {
            if (s % 2 == 0) f1.heat7(f2, 0.0f);
            else f2.heat7(f1, 0.0f);
        }
        write_grid(dir + "/heat_fast.bin", f2.gather(), nx, ny, nz);
        DC<float> box(ny, nx, nz, 0.0f);
        c1.fill(cube.data());
        c1.mapStencil<box2_plc>(box, 2, 0.25f);
        write_grid(dir + "/box2.bin", box.gather(), nx, ny, nz);
        DC<float> reach3(ny, nx, nz, 0.0f);
        c1.mapStencil<reach2_plc>(reach3, 1, 0.75f);
        write_grid(dir + "/reach_plc.bin", reach3.gather(), nx, ny, nz);

        // ---- float fold of the blur input through a user functor
        DA<float> flat((int)img.size());
        flat.fill(img.data());
        auto fadd = [] __host__ __device__(float x, float y) { return x + y; };
        const float fs = flat.fold(fadd);
        write_grid(dir + "/fold_sum.bin", &fs, 1, 1, 1);
    } catch (const msl::detail::Exception &e) {
        fprintf(stderr, "exception: %s\n", e.what());
        return 3;
    }
    msl::terminateSkeletons();
    printf("%d failures\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}

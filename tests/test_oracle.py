"""Pins the oracle (oracle/refcpu.c) with checks that do not restate it.

Each test uses an independent source of truth: published test vectors,
closed forms, textbook Game-of-Life patterns, numpy's IEEE float32
arithmetic, math.fsum, or a brute-force loop that is written differently
from the oracle (explicit neighbour lists, no shared helper).
"""
import math

import numpy as np
import pytest

from oracle import refcpu as R

# Published SplitMix64 outputs for seed 0 (Vigna's reference generator):
# state starts at 0 and each call adds the golden gamma before mixing.
SPLITMIX64_SEED0 = [16294208416658607535, 7960286522194355700, 487617019471545679]


def test_hash64_known_vectors():
    assert [R.hash64(0, i) for i in range(3)] == SPLITMIX64_SEED0


def _np_hash(seed, n):
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(n, dtype=np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def test_fills_match_numpy_restatement():
    n = 10007
    h = _np_hash(42, n)
    u = (h >> np.uint64(40)).astype(np.float32)
    np.testing.assert_array_equal(R.fill_uniform_f32(n, 42), u * R.SCALE_255)
    x = R.fill_uniform_f32(n, 42)
    assert x.min() >= 0.0 and x.max() < 255.0
    np.testing.assert_array_equal(R.fill_bernoulli_i32(n, 7), (_np_hash(7, n) < np.uint64(R.BERNOULLI_03)).astype(np.int32))
    b = R.fill_bernoulli_i32(200000, 11)
    assert abs(b.mean() - 0.3) < 0.005
    np.testing.assert_array_equal(R.fill_u32(n, 3), (_np_hash(3, n) >> np.uint64(32)).astype(np.uint32))


def test_elementwise_match_numpy_ieee():
    x = R.fill_uniform_f32(4099, 1)
    y = R.fill_uniform_f32(4099, 2)
    np.testing.assert_array_equal(R.map_affine_f32(x, 1.5, 3.0), x * np.float32(1.5) + np.float32(3.0))
    np.testing.assert_array_equal(R.zip_add_f32(x, y), x + y)
    u = R.fill_u32(4099, 5)
    np.testing.assert_array_equal(R.map_affine_u32(u, 3, 1), u * np.uint32(3) + np.uint32(1))
    m = R.mapindex_lin_f32(37, 53)
    np.testing.assert_array_equal(m, np.arange(37 * 53, dtype=np.int64).astype(np.float32).reshape(37, 53))


def test_mapindex_rounding_beyond_2_24():
    # 8192 x 8192 indices exceed 2^24: the int->float conversion rounds to
    # nearest even; check a band of large rows against numpy's conversion.
    rows, cols = 8192, 8192
    m = R.mapindex_lin_f32(rows, cols)
    tail = np.arange((rows - 3) * cols, rows * cols, dtype=np.int64).astype(np.float32)
    np.testing.assert_array_equal(m[-3:].ravel(), tail)
    assert m[-1, -1] == np.float32(67108864.0)  # 67108863 rounds up


def test_folds_against_fsum_and_closed_forms():
    x = R.fill_uniform_f32(100003, 9)
    assert R.fold_sum_f64(x) == pytest.approx(math.fsum(x.astype(np.float64)), rel=1e-15)
    # sequential float32 sum == numpy's sequential accumulate in float32
    small = x[:5000]
    assert R.fold_sum_f32_seq(small) == float(np.add.accumulate(small, dtype=np.float32)[-1])
    # stagnation: adding 1.0f stops changing at 2^24
    ones = np.ones(2 ** 24 + 1000, np.float32)
    assert R.fold_sum_f32_seq(ones) == 16777216.0
    assert R.fold_sum_f64(ones) == float(2 ** 24 + 1000)
    u = R.fill_u32(70001, 4)
    assert R.fold_sum_u32(u) == int(u.astype(np.uint64).sum() % (1 << 32))
    c = R.fill_bernoulli_i32(70001, 4)
    assert R.fold_sum_i32(c) == int(c.sum())
    assert R.fold_max_f32(x) == float(x.max())
    y = R.fill_uniform_f32(100003, 10)
    assert R.max_abs_diff_f32(x, y) == float(np.abs(x - y).max())


# ---- 2-D mean blur ---------------------------------------------------------
def _brute_mean(a, r, nv):
    rows, cols = a.shape
    out = np.empty_like(a)
    f = np.float32
    for y in range(rows):
        for x in range(cols):
            s = f(0.0)
            for yy in range(y - r, y + r + 1):
                rs = f(0.0)
                for xx in range(x - r, x + r + 1):
                    v = a[yy, xx] if (0 <= yy < rows and 0 <= xx < cols) else f(nv)
                    rs = f(rs + v)
                s = f(s + rs)
            out[y, x] = f(s / f((2 * r + 1) ** 2))
    return out


@pytest.mark.parametrize("r", [1, 2, 3])
def test_mean_blur_brute_force(r):
    a = R.fill_uniform_f32(9 * 13, 5 + r).reshape(9, 13)
    np.testing.assert_array_equal(R.stencil2d_mean_f32(a, r, 0.0), _brute_mean(a, r, 0.0))
    np.testing.assert_array_equal(R.stencil2d_mean_f32(a, r, 7.25), _brute_mean(a, r, 7.25))


@pytest.mark.parametrize("r", [1, 5, 10])
def test_mean_blur_constant_field_closed_form(r):
    rows, cols = 31, 40
    a = np.ones((rows, cols), np.float32)
    # nv equal to the field: exactly 1 everywhere (sums of ones are exact)
    assert np.all(R.stencil2d_mean_f32(a, r, 1.0) == 1.0)
    # nv = 0: value = (#rows in range * #cols in range) / (2r+1)^2
    out = R.stencil2d_mean_f32(a, r, 0.0)
    ys = np.arange(rows)
    xs = np.arange(cols)
    ny = np.minimum(ys + r, rows - 1) - np.maximum(ys - r, 0) + 1
    nx = np.minimum(xs + r, cols - 1) - np.maximum(xs - r, 0) + 1
    expect = (ny[:, None] * nx[None, :]).astype(np.float32) / np.float32((2 * r + 1) ** 2)
    np.testing.assert_array_equal(out, expect)


def test_mean_blur_is_not_transposed():
    # x = column: a single hot cell in a wide grid spreads r columns each way
    a = np.zeros((5, 11), np.float32)
    a[2, 5] = 9.0
    out = R.stencil2d_mean_f32(a, 1, 0.0)
    assert np.count_nonzero(out) == 9 and out[1:4, 4:7].min() == 1.0


# ---- Game of Life -------------------------------------------------------------
def _step(a, n=1):
    for _ in range(n):
        a = R.gol_i32(a)
    return a


def test_gol_block_still_life():
    a = np.zeros((6, 6), np.int32)
    a[2:4, 2:4] = 1
    np.testing.assert_array_equal(_step(a, 5), a)


def test_gol_blinker_period_two():
    a = np.zeros((5, 5), np.int32)
    a[2, 1:4] = 1
    b = _step(a)
    expect = np.zeros_like(a)
    expect[1:4, 2] = 1
    np.testing.assert_array_equal(b, expect)
    np.testing.assert_array_equal(_step(a, 2), a)


def test_gol_glider_moves_diagonally():
    a = np.zeros((12, 12), np.int32)
    for (y, x) in [(0, 1), (1, 2), (2, 0), (2, 1), (2, 2)]:
        a[y + 1, x + 1] = 1
    b = _step(a, 4)
    np.testing.assert_array_equal(b, np.roll(np.roll(a, 1, axis=0), 1, axis=1))


def test_gol_brute_force_random():
    a = R.fill_bernoulli_i32(14 * 19, 3).reshape(14, 19)
    rows, cols = a.shape
    expect = np.zeros_like(a)
    for y in range(rows):
        for x in range(cols):
            n = 0
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if (dy or dx) and 0 <= y + dy < rows and 0 <= x + dx < cols:
                        n += a[y + dy, x + dx]
            expect[y, x] = 1 if (n == 3 or (a[y, x] == 1 and n == 2)) else 0
    np.testing.assert_array_equal(R.gol_i32(a), expect)


def test_gol_edges_are_dead_outside():
    # a blinker touching the border loses its off-grid cells
    a = np.zeros((4, 4), np.int32)
    a[0, 0:3] = 1
    b = R.gol_i32(a)
    expect = np.zeros_like(a)
    expect[0:2, 1] = 1
    np.testing.assert_array_equal(b, expect)


# ---- 3-D heat -----------------------------------------------------------------
def test_heat7_brute_force():
    nz, ny, nx = 5, 6, 7
    a = R.fill_uniform_f32(nz * ny * nx, 12, R.SCALE_UNIT).reshape(nz, ny, nx)
    f = np.float32
    expect = np.empty_like(a)
    nv = f(0.5)

    def g(z, y, x):
        return a[z, y, x] if (0 <= z < nz and 0 <= y < ny and 0 <= x < nx) else nv

    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                v = f(0.0)
                for (dz, dy, dx) in [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]:
                    v = f(v + g(z + dz, y + dy, x + dx))
                expect[z, y, x] = f(v / f(6.0))
    np.testing.assert_array_equal(R.heat7_f32(a, 0.5), expect)
    out, m = R.heat7_residual_f32(a, 0.5)
    np.testing.assert_array_equal(out, expect)
    assert m == float(np.abs(expect - a).max())


def test_heat7_constant_field():
    a = np.ones((6, 7, 8), np.float32)
    assert np.all(R.heat7_f32(a, 1.0) == 1.0)
    out = R.heat7_f32(a, 0.0)
    assert out[3, 3, 3] == 1.0
    assert out[0, 3, 3] == np.float32(5.0) / np.float32(6.0)
    assert out[0, 0, 3] == np.float32(4.0) / np.float32(6.0)
    assert out[0, 0, 0] == np.float32(3.0) / np.float32(6.0)


def test_boxsum_all_ones_closed_form():
    # examples/dc_map_stencil_test.cpp: interior 27, face 18, edge 12, corner 8
    a = np.ones((16, 10, 10), np.float32)
    out = R.stencil3d_boxsum_f32(a, 1, 0.0)
    assert out[5, 5, 5] == 27 and out[0, 5, 5] == 18 and out[0, 0, 5] == 12 and out[0, 0, 0] == 8
    vals, counts = np.unique(out, return_counts=True)
    assert set(vals.tolist()) == {8.0, 12.0, 18.0, 27.0}


# ---- FP64 drift references -------------------------------------------------------
def test_mean_blur_f64_brute_force():
    a = R.fill_uniform_f32(7 * 10, 31).reshape(7, 10).astype(np.float64)
    r, nv = 2, 0.75
    rows, cols = a.shape
    expect = np.empty_like(a)
    for y in range(rows):
        for x in range(cols):
            tot = 0.0
            for dy in range(-r, r + 1):
                row = 0.0
                for dx in range(-r, r + 1):
                    yy, xx = y + dy, x + dx
                    row += a[yy, xx] if (0 <= yy < rows and 0 <= xx < cols) else nv
                tot += row
            expect[y, x] = tot / (2 * r + 1) ** 2
    np.testing.assert_array_equal(R.stencil2d_mean_f64(a, r, nv), expect)


def test_heat7_f64_brute_force_and_closed_form():
    nz, ny, nx = 4, 5, 6
    a = R.fill_uniform_f32(nz * ny * nx, 32, R.SCALE_UNIT).reshape(nz, ny, nx).astype(np.float64)

    def g(z, y, x):
        return a[z, y, x] if (0 <= z < nz and 0 <= y < ny and 0 <= x < nx) else 0.0

    expect = np.empty_like(a)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                v = 0.0
                for (dz, dy, dx) in [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]:
                    v += g(z + dz, y + dy, x + dx)
                expect[z, y, x] = v / 6.0
    np.testing.assert_array_equal(R.heat7_f64(a, 0.0), expect)
    ones = np.ones((5, 5, 5))
    out = R.heat7_f64(ones, 0.0)
    assert out[2, 2, 2] == 1.0 and out[0, 0, 0] == 0.5


# ---- numpy restatements on grids with a real interior ---------------------------
# The oracle's stencil loops run row-wise with the bounds tests split off the
# interior; these restatements build each cell's sum from shifted slices of a
# padded copy (numpy float32/float64 elementwise IEEE adds in the same order),
# so interior cells at every radius and all four borders are covered.
def _np_mean(a, r, nv):
    dt = a.dtype.type
    rows, cols = a.shape
    p = np.pad(a, r, constant_values=dt(nv))
    s = np.zeros_like(a)
    for dy in range(2 * r + 1):
        rs = np.zeros_like(a)
        for dx in range(2 * r + 1):
            rs = rs + p[dy:dy + rows, dx:dx + cols]
        s = s + rs
    return s / dt((2 * r + 1) ** 2)


def _np_heat(a, nv):
    dt = a.dtype.type
    p = np.pad(a, 1, constant_values=dt(nv))
    v = np.zeros_like(a)
    for sl in [(slice(1, -1), slice(1, -1), slice(0, -2)), (slice(1, -1), slice(1, -1), slice(2, None)),
               (slice(1, -1), slice(0, -2), slice(1, -1)), (slice(1, -1), slice(2, None), slice(1, -1)),
               (slice(0, -2), slice(1, -1), slice(1, -1)), (slice(2, None), slice(1, -1), slice(1, -1))]:
        v = v + p[sl]
    return v / dt(6)


@pytest.mark.parametrize("r", [1, 4, 5, 10, 16])
@pytest.mark.parametrize("shape", [(61, 97), (2, 130), (130, 3)])
def test_mean_blur_numpy_restatement(r, shape):
    rows, cols = shape
    a = R.fill_uniform_f32(rows * cols, 90 + r).reshape(rows, cols)
    a[::7, ::5] = -0.0
    for nv in (0.0, -0.0, 2.5):
        np.testing.assert_array_equal(R.stencil2d_mean_f32(a, r, nv).view(np.uint32),
                                      _np_mean(a, r, nv).view(np.uint32))
        a64 = a.astype(np.float64)
        np.testing.assert_array_equal(R.stencil2d_mean_f64(a64, r, nv), _np_mean(a64, r, nv))


@pytest.mark.parametrize("shape", [(23, 31, 47), (1, 9, 70), (12, 1, 5), (7, 6, 1)])
def test_heat7_numpy_restatement(shape):
    nz, ny, nx = shape
    a = R.fill_uniform_f32(nz * ny * nx, 77, R.SCALE_UNIT).reshape(shape)
    a[::3, ::2, ::5] = -0.0
    for nv in (0.0, -0.0, 0.25):
        expect = _np_heat(a, nv)
        np.testing.assert_array_equal(R.heat7_f32(a, nv).view(np.uint32), expect.view(np.uint32))
        out, m = R.heat7_residual_f32(a, nv)
        np.testing.assert_array_equal(out.view(np.uint32), expect.view(np.uint32))
        assert m == float(np.abs(expect - a).max())
        a64 = a.astype(np.float64)
        np.testing.assert_array_equal(R.heat7_f64(a64, nv), _np_heat(a64, nv))


def test_iterate_helpers_equal_repeated_steps():
    a = R.fill_uniform_f32(40 * 50, 5).reshape(40, 50)
    ref = a
    for _ in range(5):
        ref = R.stencil2d_mean_f32(ref, 2, 1.0)
    np.testing.assert_array_equal(R.mean_blur_iterate_f32(a, 2, 5, 1.0), ref)
    h = R.fill_uniform_f32(6 * 7 * 8, 6, R.SCALE_UNIT).reshape(6, 7, 8)
    ref, res = h, []
    for it in range(7):
        if (it + 1) % 3 == 0:
            ref, m = R.heat7_residual_f32(ref, 0.0)
            res.append(m)
        else:
            ref = R.heat7_f32(ref, 0.0)
    got, got_res = R.heat7_iterate_f32(h, 7, 3)
    np.testing.assert_array_equal(got, ref)
    assert got_res == res
    ref64 = h.astype(np.float64)
    for _ in range(4):
        ref64 = R.heat7_f64(ref64)
    np.testing.assert_array_equal(R.heat7_iterate_f64(h, 4), ref64)

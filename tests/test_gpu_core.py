"""GPU parity: every C-ABI skeleton against the oracle on the same seeded inputs.

Bar: bit-exact for integer, elementwise-float and stencil results (the HIP
kernels keep the oracle's operation order, no FMA contraction, IEEE
division); the compensated float fold is checked against the oracle's FP64
sum with a stated relative tolerance.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from muesli4_amd import hip
from muesli4_amd.hip import DeviceArray, call
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

# Compensated FP32 fold vs FP64 oracle: relative error bound used in tests.
FOLD_REL_TOL = 1e-6
# FP32 heat / blur iterations against FP64 recurrences: absolute drift bound
# on values in [0, 1) (heat) and [0, 255) (blur, scaled by 255).
DRIFT_ABS_TOL = 1e-5


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    hip.init(0)
    yield
    hip.synchronize()


def dev_uniform(n, seed, scale=R.SCALE_255):
    d = DeviceArray(n, np.float32)
    call("msl_fill_uniform_f32", d.ptr, n, seed, float(scale))
    return d


# ---- fills --------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 3, 4, 4097, 1000003])
def test_fills_bit_exact(n):
    note(f"fills n={n}")
    d = dev_uniform(n, 77)
    np.testing.assert_array_equal(d.to_host(), R.fill_uniform_f32(n, 77))
    b = DeviceArray(n, np.int32)
    call("msl_fill_bernoulli_i32", b.ptr, n, 5, R.BERNOULLI_03)
    np.testing.assert_array_equal(b.to_host(), R.fill_bernoulli_i32(n, 5))
    u = DeviceArray(n, np.uint32)
    call("msl_fill_u32", u.ptr, n, 6)
    np.testing.assert_array_equal(u.to_host(), R.fill_u32(n, 6))
    c = DeviceArray(n, np.float32)
    call("msl_fill_f32", c.ptr, n, 2.5)
    assert np.all(c.to_host() == np.float32(2.5))


# ---- elementwise --------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 5, 4096, 65539, 8192 * 8192])
def test_map_zip_bit_exact(n):
    note(f"map/zip n={n}")
    a = dev_uniform(n, 1)
    b = dev_uniform(n, 2)
    out = DeviceArray(n, np.float32)
    ha, hb = R.fill_uniform_f32(n, 1), R.fill_uniform_f32(n, 2)
    call("msl_map_affine_f32", out.ptr, a.ptr, n, 1.5, 3.0)
    np.testing.assert_array_equal(out.to_host(), R.map_affine_f32(ha, 1.5, 3.0))
    call("msl_zip_add_f32", out.ptr, a.ptr, b.ptr, n)
    np.testing.assert_array_equal(out.to_host(), R.zip_add_f32(ha, hb))
    # in place (mapInPlace)
    call("msl_map_affine_f32", a.ptr, a.ptr, n, 1.5, 3.0)
    np.testing.assert_array_equal(a.to_host(), R.map_affine_f32(ha, 1.5, 3.0))


def test_map_unaligned_pointer():
    note("map unaligned")
    n = 1001
    a = dev_uniform(n + 1, 3)
    out = DeviceArray(n + 1, np.float32)
    call("msl_map_affine_f32", out.ptr + 4, a.ptr + 4, n, 1.5, 3.0)
    got = out.to_host()[1:]
    np.testing.assert_array_equal(got, R.map_affine_f32(R.fill_uniform_f32(n + 1, 3)[1:], 1.5, 3.0))


def test_map_u32_wraps():
    note("map u32")
    n = 100003
    u = DeviceArray(n, np.uint32)
    call("msl_fill_u32", u.ptr, n, 9)
    out = DeviceArray(n, np.uint32)
    call("msl_map_affine_u32", out.ptr, u.ptr, n, 3, 1)
    np.testing.assert_array_equal(out.to_host(), R.map_affine_u32(R.fill_u32(n, 9), 3, 1))


@pytest.mark.parametrize("shape", [(1, 1), (3, 7), (37, 53), (8192, 8192)])
def test_mapindex_bit_exact(shape):
    note(f"mapindex {shape}")
    rows, cols = shape
    out = DeviceArray((rows, cols), np.float32)
    call("msl_mapindex_lin_f32", out.ptr, rows, cols)
    np.testing.assert_array_equal(out.to_host(), R.mapindex_lin_f32(rows, cols))


# ---- folds --------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 7, 1000, 1 << 20, 8192 * 8192 + 3])
def test_fold_sum_f32_vs_fp64(n):
    note(f"fold f32 n={n}")
    d = dev_uniform(n, 11)
    s64, s32 = C.c_double(), C.c_float()
    call("msl_fold_sum_f32_host", d.ptr, n, C.byref(s64), C.byref(s32))
    ref = R.fold_sum_f64(R.fill_uniform_f32(n, 11))
    rel = abs(s64.value - ref) / max(abs(ref), 1e-30)
    assert rel <= FOLD_REL_TOL, (s64.value, ref, rel)
    assert s32.value == np.float32(s64.value)
    # deterministic
    s64b = C.c_double()
    call("msl_fold_sum_f32_host", d.ptr, n, C.byref(s64b), None)
    assert s64b.value == s64.value


def test_fold_sum_f32_exact_where_fp32_stagnates():
    note("fold f32 ones")
    n = (1 << 24) + 1000
    d = DeviceArray(n, np.float32)
    call("msl_fill_f32", d.ptr, n, 1.0)
    s64 = C.c_double()
    call("msl_fold_sum_f32_host", d.ptr, n, C.byref(s64), None)
    assert s64.value == float(n)
    assert R.fold_sum_f32_seq(np.ones(n, np.float32)) == 16777216.0  # the plain FP32 order stagnates


def test_fold_sum_f32_device_result():
    note("fold f32 device result")
    n = 123457
    d = dev_uniform(n, 12)
    res = DeviceArray(1, np.float64)
    call("msl_fold_sum_f32", d.ptr, n, res.ptr)
    got = float(res.to_host()[0])
    ref = R.fold_sum_f64(R.fill_uniform_f32(n, 12))
    assert abs(got - ref) / ref <= FOLD_REL_TOL


@pytest.mark.parametrize("n", [1, 4099, 10 ** 7 + 1])
def test_int_folds_exact(n):
    note(f"int folds n={n}")
    u = DeviceArray(n, np.uint32)
    call("msl_fill_u32", u.ptr, n, 21)
    r = C.c_uint32()
    call("msl_fold_sum_u32_host", u.ptr, n, C.byref(r))
    assert r.value == R.fold_sum_u32(R.fill_u32(n, 21))
    b = DeviceArray(n, np.int32)
    call("msl_fill_bernoulli_i32", b.ptr, n, 22, R.BERNOULLI_03)
    s = C.c_int64()
    call("msl_fold_sum_i32_host", b.ptr, n, C.byref(s))
    assert s.value == R.fold_sum_i32(R.fill_bernoulli_i32(n, 22))
    f = dev_uniform(n, 23)
    m = C.c_float()
    call("msl_fold_max_f32_host", f.ptr, n, C.byref(m))
    assert m.value == R.fold_max_f32(R.fill_uniform_f32(n, 23))


def test_fold_max_rejects_empty():
    note("fold max empty")
    d = DeviceArray(4, np.float32)
    m = C.c_float()
    with pytest.raises(hip.MslError):
        call("msl_fold_max_f32_host", d.ptr, 0, C.byref(m))


# ---- 2-D stencils -------------------------------------------------------------
SHAPES_2D = [(1, 1), (1, 300), (300, 1), (7, 9), (37, 53), (64, 256), (130, 1001), (256, 260)]


@pytest.mark.parametrize("r", list(range(1, 11)))
def test_mean_blur_bit_exact_small(r):
    for (rows, cols) in SHAPES_2D:
        for nv in (0.0, 3.5):
            note(f"mean r={r} {rows}x{cols} nv={nv}")
            n = rows * cols
            a = dev_uniform(n, 100 + r)
            out = DeviceArray((rows, cols), np.float32)
            call("msl_stencil2d_mean_f32", out.ptr, a.ptr, rows, cols, r, nv)
            ref = R.stencil2d_mean_f32(R.fill_uniform_f32(n, 100 + r).reshape(rows, cols), r, nv)
            np.testing.assert_array_equal(out.to_host(), ref, err_msg=f"r={r} shape={rows}x{cols} nv={nv}")


# Signed zeros: the definition's sums start at +0, so a window of -0 values
# averages to +0, while -0 neutral values and mixed zeros elsewhere must come
# out exactly as in the oracle, through every kernel family (one generation,
# two per launch, six per launch at r = 1).
@pytest.mark.parametrize("r", [1, 2, 5, 10])
@pytest.mark.parametrize("iters", [1, 2, 6])
def test_mean_blur_signed_zeros(r, iters, monkeypatch):
    rows, cols = 70, 296
    rng = np.random.default_rng(17 + r)
    x = np.full((rows, cols), -0.0, np.float32)
    x[rng.random((rows, cols)) < 0.02] = 0.0
    x[rng.random((rows, cols)) < 0.002] = 1.5
    x[:, : cols // 2] = -0.0  # a half of pure -0 windows
    for gens in ("1", "2"):
        monkeypatch.setenv("MSL_MEAN_GENS", gens)
        note(f"mean r={r} signed zeros x{iters} gens={gens}")
        a = DeviceArray.from_host(x)
        b = DeviceArray((rows, cols), np.float32)
        ref = x.copy()
        for _ in range(iters):
            ref = R.stencil2d_mean_f32(ref, r, -0.0)
        inb = C.c_int()
        call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, rows, cols, r, -0.0, iters, C.byref(inb))
        got = (b if inb.value else a).to_host().reshape(rows, cols)
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_fast_division_matches_ieee_on_all_inputs():
    # every 32-bit pattern, every divisor (2r+1)^2 used by the blur kernels
    for r in range(1, 17):
        note(f"div selftest r={r}")
        bad = C.c_uint64(1)
        call("msl_selftest_div_odd_square", r, C.byref(bad))
        assert bad.value == 0, (r, bad.value)
    note("div selftest d=6 (heat)")
    bad = C.c_uint64(1)
    call("msl_selftest_div_six", C.byref(bad))
    assert bad.value == 0, bad.value


@pytest.mark.parametrize("r", [11, 13, 16])
def test_mean_blur_large_radius(r):
    rows, cols = 150, 333
    note(f"mean r={r} {rows}x{cols}")
    a = dev_uniform(rows * cols, r)
    out = DeviceArray((rows, cols), np.float32)
    call("msl_stencil2d_mean_f32", out.ptr, a.ptr, rows, cols, r, 1.0)
    ref = R.stencil2d_mean_f32(R.fill_uniform_f32(rows * cols, r).reshape(rows, cols), r, 1.0)
    np.testing.assert_array_equal(out.to_host(), ref)


def test_mean_blur_rejects_bad_radius():
    note("mean bad radius")
    a = DeviceArray(16, np.float32)
    b = DeviceArray(16, np.float32)
    for r in (0, 17, -1):
        with pytest.raises(hip.MslError):
            call("msl_stencil2d_mean_f32", b.ptr, a.ptr, 4, 4, r, 0.0)
    with pytest.raises(hip.MslError):
        call("msl_stencil2d_mean_f32", a.ptr, a.ptr, 4, 4, 1, 0.0)


# Grids above 128 MiB take the non-temporal-store kernels, smaller ones the
# ordinary-store kernels; both are covered.
@pytest.mark.parametrize("r,rows,cols", [(1, 8192, 8192), (5, 5632, 6144), (10, 5632, 6144), (5, 2048, 2048),
                                         (10, 2048, 2048)])
def test_mean_blur_8192_single_step(r, rows, cols):
    note(f"mean full r={r} {rows}x{cols}")
    n = rows * cols
    a = dev_uniform(n, 5)
    out = DeviceArray((rows, cols), np.float32)
    call("msl_stencil2d_mean_f32", out.ptr, a.ptr, rows, cols, r, 0.0)
    ref = R.stencil2d_mean_f32(R.fill_uniform_f32(n, 5).reshape(rows, cols), r, 0.0)
    np.testing.assert_array_equal(out.to_host(), ref)


@pytest.mark.parametrize("r,iters", [(1, 20), (5, 20), (10, 3)])
def test_mean_blur_iterate_graph(r, iters):
    rows, cols = 384, 520
    note(f"mean iterate r={r} iters={iters}")
    n = rows * cols
    a = dev_uniform(n, 8)
    b = DeviceArray((rows, cols), np.float32)
    inb = C.c_int()
    ref = R.fill_uniform_f32(n, 8).reshape(rows, cols)
    for _ in range(iters):
        ref = R.stencil2d_mean_f32(ref, r, 0.0)
    for rep in range(2):  # second call replays the cached graph
        call("msl_fill_uniform_f32", a.ptr, n, 8, float(R.SCALE_255))
        call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, rows, cols, r, 0.0, iters, C.byref(inb))
        got = (b if inb.value else a).to_host().reshape(rows, cols)
        np.testing.assert_array_equal(got, ref, err_msg=f"replay {rep}")


# Radius 1 iterates up to four generations per launch in registers; every
# count, shape (partial tiles) and neutral value is bit-exact.
@pytest.mark.parametrize("shape", [(3, 5), (17, 33), (100, 260), (33, 64), (1, 40), (40, 1), (70, 70)])
@pytest.mark.parametrize("iters", [2, 3, 4, 5, 8, 9])
def test_mean_blur_r1_iterate_multi_generation(shape, iters):
    rows, cols = shape
    n = rows * cols
    for nv in (0.0, 3.5):
        note(f"mean r=1 iterate {shape} x{iters} nv={nv}")
        a = dev_uniform(n, 55)
        b = DeviceArray((rows, cols), np.float32)
        ref = R.fill_uniform_f32(n, 55).reshape(rows, cols)
        for _ in range(iters):
            ref = R.stencil2d_mean_f32(ref, 1, nv)
        inb = C.c_int()
        call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, rows, cols, 1, nv, iters, C.byref(inb))
        assert inb.value == (iters & 1)
        np.testing.assert_array_equal((b if inb.value else a).to_host().reshape(rows, cols), ref)


@pytest.fixture
def mean_pairs(monkeypatch):
    """Force two generations per launch for radii 2..7 (the library's default
    on aligned grids above 64 MiB)."""
    monkeypatch.setenv("MSL_MEAN_GENS", "2")


# Radii 2..7 iterate two generations per launch on rows of 4 or more cells
# (aligned rows and odd widths): ragged strips (240 or 248 output columns per
# wave) and segments, odd counts (one single-generation launch keeps the
# result-buffer parity) and neutral values are bit-exact against the oracle.
@pytest.mark.parametrize("shape", [(1, 4), (5, 8), (9, 12), (37, 244), (100, 484), (70, 964), (300, 248),
                                   (6, 5), (11, 7), (37, 243), (100, 485), (70, 963), (3, 2)])
@pytest.mark.parametrize("r", [2, 3, 5, 7])
@pytest.mark.parametrize("iters", [2, 3, 4])
def test_mean_blur_iterate_two_generation_launches(shape, r, iters, mean_pairs):
    rows, cols = shape
    n = rows * cols
    for nv in (0.0, 3.5):
        note(f"mean r={r} iterate pairs {shape} x{iters} nv={nv}")
        a = dev_uniform(n, 66)
        b = DeviceArray((rows, cols), np.float32)
        ref = R.fill_uniform_f32(n, 66).reshape(rows, cols)
        for _ in range(iters):
            ref = R.stencil2d_mean_f32(ref, r, nv)
        inb = C.c_int()
        call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, rows, cols, r, nv, iters, C.byref(inb))
        assert inb.value == (iters & 1)
        np.testing.assert_array_equal((b if inb.value else a).to_host().reshape(rows, cols), ref)


# On large grids the pairs are the default (radii 2-4 from 32 MiB, 5-7 on
# grids above 128 MiB, which store with non-temporal stores), with the last
# strip partial (4096 = 16.5 strips of 248 columns, 4104 = 17.1 of 240).  The
# plan query shows that two-generation launches run; results are bit-exact
# against the oracle.
@pytest.mark.parametrize("r,rows,cols", [(2, 4160, 4096), (4, 2100, 4096), (5, 8200, 4104), (7, 8200, 4104),
                                         (3, 4160, 4095), (6, 8200, 4103)])
def test_mean_blur_iterate_large_grid_default_pairs(r, rows, cols, monkeypatch):
    monkeypatch.delenv("MSL_MEAN_GENS", raising=False)
    iters = 4
    n = rows * cols
    note(f"mean r={r} iterate {rows}x{cols} x{iters} default plan")
    a = dev_uniform(n, 67)
    b = DeviceArray((rows, cols), np.float32)
    plan = hip.mean_iterate_plan(a.ptr, b.ptr, rows, cols, r, iters)
    assert plan == [2, 2], plan
    inb = C.c_int()
    call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, rows, cols, r, 0.0, iters, C.byref(inb))
    assert inb.value == 0
    ref = R.mean_blur_iterate_f32(R.fill_uniform_f32(n, 67).reshape(rows, cols), r, iters)
    np.testing.assert_array_equal(a.to_host().reshape(rows, cols), ref)


# The headline configuration against the oracle: 8192^2 x 20 iterations of
# the mean blur at r = 1, 5 and 10 through the default plan (six generations
# per launch at r = 1, two at r = 5, one at r = 10), bit for bit.
@pytest.mark.parametrize("r,gens", [(1, 6), (5, 2), (10, 1)])
def test_mean_blur_8192_x20_vs_oracle(r, gens, monkeypatch):
    monkeypatch.delenv("MSL_MEAN_GENS", raising=False)
    rows = cols = 8192
    iters = 20
    n = rows * cols
    note(f"mean r={r} 8192^2 x{iters} vs oracle")
    a = dev_uniform(n, 900 + r)
    b = DeviceArray((rows, cols), np.float32)
    plan = hip.mean_iterate_plan(a.ptr, b.ptr, rows, cols, r, iters)
    assert sum(plan) == iters and max(plan) == gens and len(plan) % 2 == 0, plan
    inb = C.c_int()
    call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, rows, cols, r, 0.0, iters, C.byref(inb))
    assert inb.value == 0
    got = a.to_host().reshape(rows, cols)
    ref = R.mean_blur_iterate_f32(R.fill_uniform_f32(n, 900 + r).reshape(rows, cols), r, iters)
    np.testing.assert_array_equal(got, ref)


# Grids above 2^20 cells iterate four generations per launch (smaller ones
# eight): both window shapes against the oracle, with ragged tiles.
@pytest.mark.parametrize("shape,iters", [((1100, 1000), 9), ((1000, 1050), 6), ((1024, 1024), 11)])
def test_gol_iterate_both_window_shapes(shape, iters):
    rows, cols = shape
    n = rows * cols
    note(f"gol iterate {shape} x{iters}")
    a = DeviceArray((rows, cols), np.int32)
    b = DeviceArray((rows, cols), np.int32)
    call("msl_fill_bernoulli_i32", a.ptr, n, 77, R.BERNOULLI_03)
    inb = C.c_int()
    call("msl_gol_iterate_i32", a.ptr, b.ptr, rows, cols, iters, C.byref(inb))
    assert inb.value == (iters & 1)
    ref = R.fill_bernoulli_i32(n, 77).reshape(rows, cols)
    for _ in range(iters):
        ref = R.gol_i32(ref)
    np.testing.assert_array_equal((b if inb.value else a).to_host().reshape(rows, cols), ref)


def test_gol_1024_100_iterations_and_fold():
    rows = cols = 1024
    iters = 100
    note("gol 1024x1024 x100")
    n = rows * cols
    a = DeviceArray((rows, cols), np.int32)
    b = DeviceArray((rows, cols), np.int32)
    call("msl_fill_bernoulli_i32", a.ptr, n, 2024, R.BERNOULLI_03)
    inb = C.c_int()
    call("msl_gol_iterate_i32", a.ptr, b.ptr, rows, cols, iters, C.byref(inb))
    fin = b if inb.value else a
    ref = R.fill_bernoulli_i32(n, 2024).reshape(rows, cols)
    for _ in range(iters):
        ref = R.gol_i32(ref)
    np.testing.assert_array_equal(fin.to_host(), ref)
    s = C.c_int64()
    call("msl_fold_sum_i32_host", fin.ptr, n, C.byref(s))
    assert s.value == int(ref.sum())


# The iterate entry applies up to four generations per launch in LDS
# (partial tiles at the edges, launch plans split to keep the result-buffer
# parity); every count and ragged shape is bit-exact against the oracle.
@pytest.mark.parametrize("shape", [(3, 5), (17, 33), (100, 260), (33, 64), (1, 40), (40, 1), (70, 70)])
@pytest.mark.parametrize("iters", [1, 2, 3, 4, 5, 6, 9, 13])
def test_gol_iterate_multi_generation(shape, iters):
    rows, cols = shape
    note(f"gol iterate {shape} x{iters}")
    n = rows * cols
    a = DeviceArray((rows, cols), np.int32)
    b = DeviceArray((rows, cols), np.int32)
    ref = R.fill_bernoulli_i32(n, 77).reshape(rows, cols)
    for _ in range(iters):
        ref = R.gol_i32(ref)
    inb = C.c_int()
    for rep in range(2):  # the second call replays the cached graph
        call("msl_fill_bernoulli_i32", a.ptr, n, 77, R.BERNOULLI_03)
        call("msl_gol_iterate_i32", a.ptr, b.ptr, rows, cols, iters, C.byref(inb))
        assert inb.value == (iters & 1)
        np.testing.assert_array_equal((b if inb.value else a).to_host().reshape(rows, cols), ref,
                                      err_msg=f"replay {rep}")


# Small grids march 64-column strips; grids with enough 256-column strip
# segments (on a 256-CU device) march 256-column strips in the aligned,
# unaligned-row (cols % 4 != 0) and scalar (cols < 4) row modes.
WIDE_SHAPES_2D = [(1900, 2050), (16400, 3), (64, 65536), (2048, 2048)]


@pytest.mark.parametrize("r", [1, 3, 7, 10, 12, 16])
def test_mean_blur_wide_strip_modes(r):
    for (rows, cols) in WIDE_SHAPES_2D:
        for nv in (0.0, 3.5):
            note(f"mean wide r={r} {rows}x{cols} nv={nv}")
            n = rows * cols
            a = dev_uniform(n, 200 + r)
            out = DeviceArray((rows, cols), np.float32)
            call("msl_stencil2d_mean_f32", out.ptr, a.ptr, rows, cols, r, nv)
            ref = R.stencil2d_mean_f32(R.fill_uniform_f32(n, 200 + r).reshape(rows, cols), r, nv)
            np.testing.assert_array_equal(out.to_host(), ref, err_msg=f"r={r} shape={rows}x{cols} nv={nv}")


@pytest.mark.parametrize("shape", [(3, 5), (17, 33), (100, 260)] + WIDE_SHAPES_2D)
def test_gol_single_step_ragged(shape):
    rows, cols = shape
    note(f"gol {shape}")
    n = rows * cols
    a = DeviceArray((rows, cols), np.int32)
    b = DeviceArray((rows, cols), np.int32)
    call("msl_fill_bernoulli_i32", a.ptr, n, 31, R.BERNOULLI_03)
    call("msl_gol_i32", b.ptr, a.ptr, rows, cols)
    np.testing.assert_array_equal(b.to_host(), R.gol_i32(R.fill_bernoulli_i32(n, 31).reshape(rows, cols)))


# ---- 3-D heat -------------------------------------------------------------------
SHAPES_3D = [(1, 1, 1), (5, 7, 9), (17, 9, 65), (33, 70, 130), (64, 64, 64)]


@pytest.mark.parametrize("shape", SHAPES_3D)
def test_heat7_bit_exact(shape):
    nz, ny, nx = shape
    note(f"heat7 {shape}")
    n = nx * ny * nz
    a = dev_uniform(n, 40, R.SCALE_UNIT)
    out = DeviceArray(shape, np.float32)
    host = R.fill_uniform_f32(n, 40, R.SCALE_UNIT).reshape(shape)
    for nv in (0.0, 0.25):
        call("msl_heat7_f32", out.ptr, a.ptr, nx, ny, nz, nv)
        np.testing.assert_array_equal(out.to_host(), R.heat7_f32(host, nv))
    md = DeviceArray(1, np.float32)
    call("msl_fill_f32", md.ptr, 1, 0.0)
    call("msl_heat7_residual_f32", out.ptr, a.ptr, nx, ny, nz, 0.0, md.ptr)
    ref, m = R.heat7_residual_f32(host, 0.0)
    np.testing.assert_array_equal(out.to_host(), ref)
    assert float(md.to_host()[0]) == m


def test_heat7_512_single_step():
    shape = (512, 512, 512)
    note("heat7 512^3")
    n = 512 ** 3
    a = dev_uniform(n, 41, R.SCALE_UNIT)
    out = DeviceArray(shape, np.float32)
    call("msl_heat7_f32", out.ptr, a.ptr, 512, 512, 512, 0.0)
    np.testing.assert_array_equal(out.to_host(), R.heat7_f32(R.fill_uniform_f32(n, 41, R.SCALE_UNIT).reshape(shape)))


def test_heat7_residual_above_store_threshold():
    """A 138 MiB cube: the residual kernel with non-temporal stores."""
    shape = (336, 320, 336)
    nz, ny, nx = shape
    note(f"heat7 residual {shape}")
    n = nx * ny * nz
    a = dev_uniform(n, 43, R.SCALE_UNIT)
    out = DeviceArray(shape, np.float32)
    md = DeviceArray(1, np.float32)
    call("msl_fill_f32", md.ptr, 1, 0.0)
    call("msl_heat7_residual_f32", out.ptr, a.ptr, nx, ny, nz, 0.0, md.ptr)
    ref, m = R.heat7_residual_f32(R.fill_uniform_f32(n, 43, R.SCALE_UNIT).reshape(shape), 0.0)
    np.testing.assert_array_equal(out.to_host(), ref)
    assert float(md.to_host()[0]) == m


# The iterate entry runs two generations per launch where it can (a residual
# generation ends its launch); ragged cubes, counts and residual cadences are
# bit-exact against single oracle steps, residuals included.
@pytest.fixture
def heat_pairs(monkeypatch):
    """Force two generations per launch (the library's default above 64 MiB)."""
    monkeypatch.setenv("MSL_HEAT_GENS", "2")


@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 7, 9), (17, 9, 65), (9, 33, 130), (3, 70, 63), (20, 20, 124)])
@pytest.mark.parametrize("iters,every", [(1, 0), (2, 0), (3, 0), (5, 0), (2, 1), (4, 2), (6, 3), (7, 2), (9, 4)])
def test_heat7_iterate_two_generation_launches(shape, iters, every, heat_pairs):
    nz, ny, nx = shape
    note(f"heat7 iterate pairs {shape} x{iters} every {every}")
    n = nx * ny * nz
    for nv in (0.0, 0.25):
        a = dev_uniform(n, 44, R.SCALE_UNIT)
        b = DeviceArray(shape, np.float32)
        nres = iters // every if every else 0
        res = (C.c_float * max(nres, 1))()
        inb = C.c_int()
        ref = R.fill_uniform_f32(n, 44, R.SCALE_UNIT).reshape(shape)
        ref_res = []
        for it in range(iters):
            if every and (it + 1) % every == 0:
                ref, m = R.heat7_residual_f32(ref, nv)
                ref_res.append(m)
            else:
                ref = R.heat7_f32(ref, nv)
        call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, nv, iters, every, res if nres else None,
             C.byref(inb))
        assert inb.value == (iters & 1)
        np.testing.assert_array_equal((b if inb.value else a).to_host().reshape(shape), ref, err_msg=f"nv={nv}")
        assert list(res)[:nres] == ref_res


def test_heat7_iterate_large_cube_default_pairs(monkeypatch):
    """A 77 MiB cube takes the paired-generation launches by default."""
    monkeypatch.delenv("MSL_HEAT_GENS", raising=False)
    shape = (272, 272, 272)
    nz, ny, nx = shape
    iters, every = 5, 2
    note(f"heat7 iterate default {shape} x{iters}")
    n = nx * ny * nz
    assert 2 in hip.heat_iterate_plan(nx, ny, nz, iters, every)
    a = dev_uniform(n, 46, R.SCALE_UNIT)
    b = DeviceArray(shape, np.float32)
    res = (C.c_float * (iters // every))()
    inb = C.c_int()
    call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, iters, every, res, C.byref(inb))
    ref, ref_res = R.heat7_iterate_f32(R.fill_uniform_f32(n, 46, R.SCALE_UNIT).reshape(shape), iters, every)
    np.testing.assert_array_equal((b if inb.value else a).to_host().reshape(shape), ref)
    assert list(res) == ref_res


def test_heat7_iterate_pairs_above_store_threshold(monkeypatch):
    """A 138 MiB cube: two-generation launches with non-temporal stores, a
    fused residual ending each launch."""
    monkeypatch.delenv("MSL_HEAT_GENS", raising=False)
    shape = (336, 320, 336)
    nz, ny, nx = shape
    iters, every = 4, 2
    note(f"heat7 iterate pairs nt {shape} x{iters}")
    n = nx * ny * nz
    assert hip.heat_iterate_plan(nx, ny, nz, iters, every) == [2, 2]
    a = dev_uniform(n, 47, R.SCALE_UNIT)
    b = DeviceArray(shape, np.float32)
    res = (C.c_float * 2)()
    inb = C.c_int()
    call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, iters, every, res, C.byref(inb))
    assert inb.value == 0
    ref, ref_res = R.heat7_iterate_f32(R.fill_uniform_f32(n, 47, R.SCALE_UNIT).reshape(shape), iters, every)
    np.testing.assert_array_equal(a.to_host().reshape(shape), ref)
    assert list(res) == ref_res


# The 3-D configuration against the oracle: 512^3 x 200 heat iterations with
# a fused residual every 10 (two generations per launch between residuals,
# non-temporal stores), final state and all 20 residuals bit for bit; the
# drift against an FP64 recurrence on the same inputs is reported.
def test_heat7_512_x200_residuals_vs_oracle(monkeypatch):
    monkeypatch.delenv("MSL_HEAT_GENS", raising=False)
    shape = (512, 512, 512)
    nz, ny, nx = shape
    iters, every = 200, 10
    n = nx * ny * nz
    note("heat7 512^3 x200 residual every 10 vs oracle")
    plan = hip.heat_iterate_plan(nx, ny, nz, iters, every)
    assert sum(plan) == iters and plan.count(2) >= 80, plan
    a = dev_uniform(n, 48, R.SCALE_UNIT)
    b = DeviceArray(shape, np.float32)
    res = (C.c_float * (iters // every))()
    inb = C.c_int()
    call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, iters, every, res, C.byref(inb))
    assert inb.value == 0
    got = a.to_host().reshape(shape)
    x0 = R.fill_uniform_f32(n, 48, R.SCALE_UNIT).reshape(shape)
    ref, ref_res = R.heat7_iterate_f32(x0, iters, every)
    np.testing.assert_array_equal(got, ref)
    assert list(res) == ref_res
    del ref
    x64 = R.heat7_iterate_f64(x0, iters)
    d = float(np.abs(got.astype(np.float64) - x64).max())
    print(f"[drift] heat7 512^3 after {iters} steps: max|f32-f64| = {d:.3e}")
    assert d < DRIFT_ABS_TOL


def test_heat7_iterate_with_residuals():
    shape = (40, 48, 72)
    nz, ny, nx = shape
    iters, every = 30, 10
    note("heat7 iterate")
    n = nx * ny * nz
    a = dev_uniform(n, 42, R.SCALE_UNIT)
    b = DeviceArray(shape, np.float32)
    res = (C.c_float * (iters // every))()
    inb = C.c_int()
    ref = R.fill_uniform_f32(n, 42, R.SCALE_UNIT).reshape(shape)
    ref_res = []
    for it in range(iters):
        if (it + 1) % every == 0:
            ref, m = R.heat7_residual_f32(ref, 0.0)
            ref_res.append(m)
        else:
            ref = R.heat7_f32(ref, 0.0)
    for rep in range(2):
        call("msl_fill_uniform_f32", a.ptr, n, 42, float(R.SCALE_UNIT))
        call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, iters, every, res, C.byref(inb))
        got = (b if inb.value else a).to_host().reshape(shape)
        np.testing.assert_array_equal(got, ref, err_msg=f"replay {rep}")
        assert list(res) == ref_res


# Drift of the FP32 stencils against FP64 recurrences on the same inputs
# (reported; bounded by DRIFT_ABS_TOL).  The FP32 results themselves are
# bit-exact against the FP32 oracle.


def test_heat7_drift_vs_fp64():
    shape = (64, 64, 64)
    nz, ny, nx = shape
    iters = 200
    note("heat7 drift 64^3 x 200")
    n = nx * ny * nz
    a = dev_uniform(n, 43, R.SCALE_UNIT)
    b = DeviceArray(shape, np.float32)
    inb = C.c_int()
    call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, iters, 0, None, C.byref(inb))
    got = (b if inb.value else a).to_host().reshape(shape)
    x32 = R.fill_uniform_f32(n, 43, R.SCALE_UNIT).reshape(shape)
    x64 = x32.astype(np.float64)
    drift = []
    for it in range(iters):
        x32 = R.heat7_f32(x32, 0.0)
        x64 = R.heat7_f64(x64, 0.0)
        if (it + 1) % 50 == 0:
            drift.append(float(np.abs(x32.astype(np.float64) - x64).max()))
    np.testing.assert_array_equal(got, x32)
    rel = float((np.abs(got.astype(np.float64) - x64) / np.maximum(np.abs(x64), 1e-30)).max())
    print(f"[drift] heat7 64^3 max|f32-f64| every 50 steps {drift}, max rel after {iters}: {rel:.3e}")
    assert max(drift) < DRIFT_ABS_TOL


def test_mean_blur_drift_vs_fp64():
    rows, cols, r, iters = 256, 320, 5, 20
    note("mean blur drift 256x320 r5 x 20")
    a = dev_uniform(rows * cols, 44)
    b = DeviceArray((rows, cols), np.float32)
    inb = C.c_int()
    call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, rows, cols, r, 0.0, iters, C.byref(inb))
    got = (b if inb.value else a).to_host().reshape(rows, cols)
    x32 = R.fill_uniform_f32(rows * cols, 44).reshape(rows, cols)
    x64 = x32.astype(np.float64)
    for _ in range(iters):
        x32 = R.stencil2d_mean_f32(x32, r, 0.0)
        x64 = R.stencil2d_mean_f64(x64, r, 0.0)
    np.testing.assert_array_equal(got, x32)
    d = float(np.abs(got.astype(np.float64) - x64).max())
    print(f"[drift] mean blur r={r} after {iters} steps: max|f32-f64| = {d:.3e} (values < 255)")
    assert d < 255 * 1e-5


# ---- graph capture of arbitrary calls ------------------------------------------
def test_graph_capture_replays():
    note("graph capture")
    n = 100000
    a = dev_uniform(n, 50)
    out = DeviceArray(n, np.float32)
    tmp = DeviceArray(n, np.float32)
    g = C.c_void_p()
    call("msl_graph_begin")
    call("msl_map_affine_f32", tmp.ptr, a.ptr, n, 1.5, 3.0)
    call("msl_zip_add_f32", out.ptr, tmp.ptr, a.ptr, n)
    call("msl_graph_end", C.byref(g))
    ha = R.fill_uniform_f32(n, 50)
    expect = R.zip_add_f32(R.map_affine_f32(ha, 1.5, 3.0), ha)
    for rep in range(4):
        call("msl_fill_f32", out.ptr, n, -1.0)  # eager work between replays
        call("msl_graph_launch", g)
        np.testing.assert_array_equal(out.to_host(), expect, err_msg=f"replay {rep}")
    call("msl_graph_destroy", g)


def test_malloc_refused_during_capture():
    note("malloc during capture")
    call("msl_graph_begin")
    ptr = C.c_void_p()
    with pytest.raises(hip.MslError):
        call("msl_malloc", C.byref(ptr), 1024)
    g = C.c_void_p()
    call("msl_graph_end", C.byref(g))
    call("msl_graph_destroy", g)

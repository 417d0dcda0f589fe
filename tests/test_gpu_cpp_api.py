"""GPU parity of the C++ skeleton API (msl::DA / DM / DC class templates).

tests/cpp/skeletons_test.cpp runs user functors through the templates the
way the example programs do (functor classes, function-pointer stencils,
lambdas, the legacy mapStencilMM form, rotations) and writes each result;
here those results are compared with the oracle on the same seeded inputs.

Bar: bit-exact for every integer and stencil result (the user functors keep
the oracle's summation order; no FMA contraction; IEEE division); the float
fold of a user functor (tree order) within FOLD_REL_TOL of the FP64 sum.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import refcpu as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "skeletons_test")

# FP32 tree fold of n values in [0, 255): error <= ~log2(n) ulp of the sum.
FOLD_REL_TOL = 1e-5

GOL_SHAPE = (96, 130)  # rows, cols
BLUR_SHAPE = (70, 101)
HEAT_SHAPE = (17, 19, 23)  # nz, ny, nx


def _write(path, arr):
    a = np.ascontiguousarray(arr)
    dims = list(a.shape[::-1]) + [1] * (3 - a.ndim)
    with open(path, "wb") as f:
        f.write(np.asarray(dims, np.int32).tobytes())
        f.write(a.tobytes())


def _read(path, dtype):
    raw = open(path, "rb").read()
    dims = np.frombuffer(raw[:12], np.int32)
    data = np.frombuffer(raw[12:], dtype)
    x, y, z = (int(v) for v in dims)
    assert data.size == x * y * z
    if z > 1:
        return data.reshape(z, y, x)
    return data.reshape(y, x)


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (python -m muesli4_amd.build)")
    d = tmp_path_factory.mktemp("cppapi")
    gol = R.fill_bernoulli_i32(GOL_SHAPE[0] * GOL_SHAPE[1], 11).reshape(GOL_SHAPE)
    blur = R.fill_uniform_f32(BLUR_SHAPE[0] * BLUR_SHAPE[1], 12).reshape(BLUR_SHAPE)
    heat = R.fill_uniform_f32(int(np.prod(HEAT_SHAPE)), 13, R.SCALE_UNIT).reshape(HEAT_SHAPE)
    _write(d / "gol_in.bin", gol)
    _write(d / "blur_in.bin", blur)
    _write(d / "heat_in.bin", heat)
    print(f"[case] skeletons_test {d}", flush=True)
    res = subprocess.run([BIN, str(d)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    sys.stdout.write(res.stdout)
    return {"dir": d, "rc": res.returncode, "log": res.stdout, "gol": gol, "blur": blur, "heat": heat}


def test_closed_forms_pass(run):
    assert run["rc"] == 0, run["log"]
    assert "FAIL" not in run["log"]
    assert "0 failures" in run["log"]


@pytest.mark.parametrize("form", ["gol_mm", "gol_npl", "gol_fast"])
def test_game_of_life_forms(run, form):
    want = run["gol"]
    for _ in range(20):
        want = R.gol_i32(want)
    got = _read(run["dir"] / f"{form}.bin", np.int32)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("form", ["blur_npl", "blur_fast"])
def test_mean_blur_forms(run, form):
    want = R.stencil2d_mean_f32(run["blur"], 3, 0.5)
    got = _read(run["dir"] / f"{form}.bin", np.float32)
    np.testing.assert_array_equal(got, want)


def test_stencil_mm_border_functor(run):
    """Cross2: in(r+2,c) + in(r-2,c) + in(r,c+1) + in(r,c-2), radius-2 border
    band valued nvf(row, col) = row*7 + col*3 (float32 sums, left to right)."""
    x = run["blur"]
    rows, cols = x.shape
    b = 2
    pad = np.zeros((rows + 2 * b, cols + 2 * b), np.float32)
    rr, cc = np.meshgrid(np.arange(-b, rows + b), np.arange(-b, cols + b), indexing="ij")
    pad[:] = (rr * 7 + cc * 3).astype(np.float32)
    pad[b:b + rows, b:b + cols] = x

    def sh(dr, dc):
        return pad[b + dr:b + dr + rows, b + dc:b + dc + cols]

    want = np.float32(0) + sh(2, 0)
    want = want + sh(-2, 0)
    want = want + sh(0, 1)
    want = want + sh(0, -2)
    got = _read(run["dir"] / "cross_nvf.bin", np.float32)
    np.testing.assert_array_equal(got, want.astype(np.float32))


def test_rotations(run):
    """rotateRows(a): new row k = old row (k-a) mod rows (np.roll by a);
    rotateCols(a): new col j = old col (j+a) mod cols (np.roll by -a)."""
    x = run["blur"]
    cols = x.shape[1]
    step1 = np.roll(np.roll(x, 3, axis=0), 5, axis=1)
    np.testing.assert_array_equal(_read(run["dir"] / "rot_r3_cm5.bin", np.float32), step1)
    step2 = np.roll(np.roll(step1, -7, axis=0), -((cols + 2) % cols), axis=1)
    np.testing.assert_array_equal(_read(run["dir"] / "rot_rm7_c2.bin", np.float32), step2)


@pytest.mark.parametrize("form", ["heat_plc", "heat_fast"])
def test_heat_forms(run, form):
    want = run["heat"]
    for _ in range(5):
        want = R.heat7_f32(want, 0.0)
    got = _read(run["dir"] / f"{form}.bin", np.float32)
    np.testing.assert_array_equal(got, want)


def test_box_stencil_r2(run):
    want = R.stencil3d_boxsum_f32(run["heat"], 2, 0.25)
    got = _read(run["dir"] / "box2.bin", np.float32)
    np.testing.assert_array_equal(got, want)


def test_stencil_reads_beyond_radius(run):
    """Functors declared with radius 1 that read two cells away: the views
    answer from global memory (or nv off the grid), as the reference's
    views do for any offset (include/nplmatrix.h:123-139,
    include/plcube.h:105-135)."""
    f32 = np.float32
    b = run["blur"]
    rows, cols = b.shape
    p = np.full((rows + 4, cols + 4), f32(0.5), f32)
    p[2:-2, 2:-2] = b
    at = lambda dy, dx: p[2 + dy:2 + dy + rows, 2 + dx:2 + dx + cols]
    want = at(0, -2).copy()
    want = want + at(0, 2) * f32(2)
    want = want + at(-2, 0) * f32(3)
    want = want + at(2, 0) * f32(4)
    want = want - at(1, 1)
    np.testing.assert_array_equal(_read(run["dir"] / "reach_npl.bin", f32), want)

    h = run["heat"]
    nz, ny, nx = h.shape
    q = np.full((nz + 4, ny + 4, nx + 4), f32(0.75), f32)
    q[2:-2, 2:-2, 2:-2] = h
    at3 = lambda dz, dy, dx: q[2 + dz:2 + dz + nz, 2 + dy:2 + dy + ny, 2 + dx:2 + dx + nx]
    want3 = at3(0, 0, 2).copy()
    want3 = want3 + at3(0, -2, 0) * f32(2)
    want3 = want3 + at3(2, 0, 0) * f32(3)
    want3 = want3 + at3(-1, 1, -1)
    np.testing.assert_array_equal(_read(run["dir"] / "reach_plc.bin", f32), want3)


def test_user_fold_float(run):
    got = float(_read(run["dir"] / "fold_sum.bin", np.float32).ravel()[0])
    ref = math.fsum(run["blur"].astype(np.float64).ravel())
    assert abs(got - ref) <= FOLD_REL_TOL * abs(ref), (got, ref)


def test_extra_skeletons_and_random_numbers(run):
    """zipIndexInPlace3/MA/A/AAA, crossZipIndexInPlace, reduceRows,
    reducetwoColumns, setColumn, DA::*DMRows and the random-number API,
    checked in-process against host loops (exact integers; device draws
    equal host draws)."""
    names = ["DM zipIndexInPlace3", "DM zipIndexInPlaceMA", "DM zipIndexInPlaceA", "DM zipIndexInPlaceAAA",
             "DM crossZipIndexInPlace", "DM reduceRows cols=1", "DM reduceRows cols=65", "DM reduceRows cols=300",
             "DM reducetwoColumns", "DM setColumn", "DM setColumn argument checks", "DA mapIndexInPlaceDMRows",
             "DA zipIndexInPlaceDMRows", "random numbers: device == host", "random numbers: ranges"]
    for name in names:
        assert f"OK   {name}\n" in run["log"], name


def test_show_layouts(run):
    """show()/showLocal() print the reference's text layout
    (src/ds.cpp:493-534, src/dm.cpp:266-296, src/dc.cpp:284-319)."""
    want = ("sa: \n[0 10 20 30 40]\n\n"
            "[0 10 20]\n\n"
            "sal: [0 10 20 30 40 ]\n"
            "sm: \n[1 2 3\n 4 5 6]\n\n"
            "sc: \n[0 1\n 2 3\n-------------\n 4 5\n 6 7]\n\n")
    assert want in run["log"], run["log"][-800:]

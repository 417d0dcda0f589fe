"""ctypes wrapper of the oracle (liboracle.so, built from refcpu.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py as the checker; never by the product path.
Every function takes and returns numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_p = C.c_void_p
_sz = C.c_size_t
_i = C.c_int
_u64 = C.c_uint64
_f = C.c_float
_d = C.c_double

_SIG = {
    "orc_hash64": ([_u64, _u64], _u64),
    "orc_fill_uniform_f32": ([_p, _sz, _u64, _f], None),
    "orc_fill_bernoulli_i32": ([_p, _sz, _u64, _u64], None),
    "orc_fill_u32": ([_p, _sz, _u64], None),
    "orc_map_affine_f32": ([_p, _p, _sz, _f, _f], None),
    "orc_zip_add_f32": ([_p, _p, _p, _sz], None),
    "orc_mapindex_lin_f32": ([_p, _i, _i], None),
    "orc_map_affine_u32": ([_p, _p, _sz, C.c_uint32, C.c_uint32], None),
    "orc_fold_sum_f64": ([_p, _sz], _d),
    "orc_fold_sum_f32_seq": ([_p, _sz], _f),
    "orc_fold_sum_u32": ([_p, _sz], C.c_uint32),
    "orc_fold_sum_i32": ([_p, _sz], C.c_int64),
    "orc_fold_max_f32": ([_p, _sz], _f),
    "orc_max_abs_diff_f32": ([_p, _p, _sz], _f),
    "orc_stencil2d_mean_f32": ([_p, _p, _i, _i, _i, _f], None),
    "orc_stencil2d_mean_f64": ([_p, _p, _i, _i, _i, _d], None),
    "orc_gol_i32": ([_p, _p, _i, _i], None),
    "orc_heat7_f32": ([_p, _p, _i, _i, _i, _f], None),
    "orc_heat7_f64": ([_p, _p, _i, _i, _i, _d], None),
    "orc_heat7_residual_f32": ([_p, _p, _i, _i, _i, _f], _f),
    "orc_stencil3d_boxsum_f32": ([_p, _p, _i, _i, _i, _i, _f], None),
    "orc_num_threads": ([], _i),
}

BERNOULLI_03 = int(0.3 * 2.0 ** 64)  # threshold for p = 0.3 (double-rounded constant)
SCALE_255 = np.float32(255.0 / 16777216.0)  # exact: 255 * 2^-24
SCALE_UNIT = np.float32(1.0 / 16777216.0)  # exact: 2^-24


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError(f"{_PATH} missing: run `make -C oracle`")
        L = C.CDLL(_PATH)
        for name, (args, res) in _SIG.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data


def hash64(seed: int, idx: int) -> int:
    return int(lib().orc_hash64(seed, idx))


def fill_uniform_f32(n: int, seed: int, scale=SCALE_255) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().orc_fill_uniform_f32(_ptr(out), n, seed, float(scale))
    return out


def fill_bernoulli_i32(n: int, seed: int, threshold: int = BERNOULLI_03) -> np.ndarray:
    out = np.empty(n, np.int32)
    lib().orc_fill_bernoulli_i32(_ptr(out), n, seed, threshold)
    return out


def fill_u32(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    lib().orc_fill_u32(_ptr(out), n, seed)
    return out


def map_affine_f32(x: np.ndarray, a: float, b: float) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().orc_map_affine_f32(_ptr(out), _ptr(x), x.size, a, b)
    return out


def map_affine_u32(x: np.ndarray, a: int, b: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.uint32)
    out = np.empty_like(x)
    lib().orc_map_affine_u32(_ptr(out), _ptr(x), x.size, a, b)
    return out


def zip_add_f32(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.empty_like(a)
    lib().orc_zip_add_f32(_ptr(out), _ptr(a), _ptr(b), a.size)
    return out


def mapindex_lin_f32(rows: int, cols: int) -> np.ndarray:
    out = np.empty((rows, cols), np.float32)
    lib().orc_mapindex_lin_f32(_ptr(out), rows, cols)
    return out


def fold_sum_f64(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().orc_fold_sum_f64(_ptr(x), x.size))


def fold_sum_f32_seq(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().orc_fold_sum_f32_seq(_ptr(x), x.size))


def fold_sum_u32(x: np.ndarray) -> int:
    x = np.ascontiguousarray(x, np.uint32)
    return int(lib().orc_fold_sum_u32(_ptr(x), x.size))


def fold_sum_i32(x: np.ndarray) -> int:
    x = np.ascontiguousarray(x, np.int32)
    return int(lib().orc_fold_sum_i32(_ptr(x), x.size))


def fold_max_f32(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().orc_fold_max_f32(_ptr(x), x.size))


def max_abs_diff_f32(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return float(lib().orc_max_abs_diff_f32(_ptr(a), _ptr(b), a.size))


def _out(x: np.ndarray, out):
    """A fresh output like x, or `out` (same shape and dtype, not x itself)."""
    if out is None:
        return np.empty_like(x)
    assert out.shape == x.shape and out.dtype == x.dtype and out.flags.c_contiguous and out is not x
    return out


def stencil2d_mean_f32(x: np.ndarray, r: int, nv: float = 0.0, out=None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    rows, cols = x.shape
    out = _out(x, out)
    lib().orc_stencil2d_mean_f32(_ptr(out), _ptr(x), rows, cols, r, nv)
    return out


def stencil2d_mean_f64(x: np.ndarray, r: int, nv: float = 0.0, out=None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    rows, cols = x.shape
    out = _out(x, out)
    lib().orc_stencil2d_mean_f64(_ptr(out), _ptr(x), rows, cols, r, nv)
    return out


def gol_i32(x: np.ndarray, out=None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.int32)
    rows, cols = x.shape
    out = _out(x, out)
    lib().orc_gol_i32(_ptr(out), _ptr(x), rows, cols)
    return out


def heat7_f32(x: np.ndarray, nv: float = 0.0, out=None) -> np.ndarray:
    """x has shape (nz, ny, nx)."""
    x = np.ascontiguousarray(x, np.float32)
    nz, ny, nx = x.shape
    out = _out(x, out)
    lib().orc_heat7_f32(_ptr(out), _ptr(x), nx, ny, nz, nv)
    return out


def heat7_f64(x: np.ndarray, nv: float = 0.0, out=None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    nz, ny, nx = x.shape
    out = _out(x, out)
    lib().orc_heat7_f64(_ptr(out), _ptr(x), nx, ny, nz, nv)
    return out


def heat7_residual_f32(x: np.ndarray, nv: float = 0.0, out=None):
    x = np.ascontiguousarray(x, np.float32)
    nz, ny, nx = x.shape
    out = _out(x, out)
    m = lib().orc_heat7_residual_f32(_ptr(out), _ptr(x), nx, ny, nz, nv)
    return out, float(m)


def stencil3d_boxsum_f32(x: np.ndarray, r: int, nv: float = 0.0) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    nz, ny, nx = x.shape
    out = np.empty_like(x)
    lib().orc_stencil3d_boxsum_f32(_ptr(out), _ptr(x), nx, ny, nz, r, nv)
    return out


def num_threads() -> int:
    return int(lib().orc_num_threads())


def mean_blur_iterate_f32(x: np.ndarray, r: int, iters: int, nv: float = 0.0) -> np.ndarray:
    """`iters` oracle mean-blur steps, ping-ponging two buffers."""
    a = np.array(x, np.float32, copy=True, order="C")
    b = np.empty_like(a)
    for _ in range(iters):
        stencil2d_mean_f32(a, r, nv, out=b)
        a, b = b, a
    return a


def heat7_iterate_f32(x: np.ndarray, iters: int, every: int = 0, nv: float = 0.0):
    """`iters` oracle heat steps, a fused-residual step at every `every`-th
    generation (as msl_heat7_iterate_f32); returns (state, residuals)."""
    a = np.array(x, np.float32, copy=True, order="C")
    b = np.empty_like(a)
    res = []
    for it in range(iters):
        if every and (it + 1) % every == 0:
            res.append(heat7_residual_f32(a, nv, out=b)[1])
        else:
            heat7_f32(a, nv, out=b)
        a, b = b, a
    return a, res


def heat7_iterate_f64(x: np.ndarray, iters: int, nv: float = 0.0) -> np.ndarray:
    """`iters` FP64 heat steps (the drift reference)."""
    a = np.array(x, np.float64, copy=True, order="C")
    b = np.empty_like(a)
    for _ in range(iters):
        heat7_f64(a, nv, out=b)
        a, b = b, a
    return a

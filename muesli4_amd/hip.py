"""ctypes binding of libmsl_hip.so (the C ABI declared in include/msl_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MSL_HIP_LIB names another build of the library (e.g. to compare two
# builds in one process tree); by default the in-tree build is loaded.
_LIB_PATH = os.environ.get("MSL_HIP_LIB") or os.path.join(_HERE, "lib", "libmsl_hip.so")

MSL_OK = 0
ERROR_NAMES = {-1: "MSL_E_INVALID", -2: "MSL_E_HIP", -3: "MSL_E_NOMEM", -4: "MSL_E_STATE",
               -5: "MSL_E_TIMEOUT", -6: "MSL_E_FULL"}


class MslError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} -> {ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code
        self.func = func


p = C.c_void_p
i32 = C.c_int
u64 = C.c_uint64
sz = C.c_size_t
f32 = C.c_float
PF = C.POINTER(C.c_float)
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)
PU32 = C.POINTER(C.c_uint32)
PI64 = C.POINTER(C.c_int64)
PP = C.POINTER(C.c_void_p)



class BatchGrid(C.Structure):
    """msl_batch_grid: one grid of a batched mean blur."""

    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("rows", C.c_int), ("cols", C.c_int),
                ("r", C.c_int), ("nv", C.c_float)]


PBG = C.POINTER(BatchGrid)

MSL_JOB_BLUR2D = 1
MSL_JOB_HEAT3D = 2
MSL_JOB_MAPFOLD_U32 = 3


class Job(C.Structure):
    """msl_job: one serving request."""

    _fields_ = [("kind", C.c_int), ("steps", C.c_int), ("r", C.c_int), ("d0", C.c_int), ("d1", C.c_int),
                ("d2", C.c_int), ("n", C.c_int64), ("nv", C.c_float), ("mul", C.c_uint32), ("add", C.c_uint32),
                ("seed", C.c_uint64), ("keep_final", C.c_int)]


class JobResult(C.Structure):
    """msl_job_result."""

    _fields_ = [("id", C.c_int64), ("kind", C.c_int), ("sum", C.c_double), ("sum_u32", C.c_uint32),
                ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("bytes", C.c_size_t)]


PJOB = C.POINTER(Job)
PRES = C.POINTER(JobResult)

# name -> argtypes (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "msl_init": [i32],
    "msl_finalize": [],
    "msl_last_error": [],
    "msl_stream": [],
    "msl_synchronize": [],
    "msl_device_info": [PI, PI, C.POINTER(C.c_size_t), C.c_char_p, i32],
    "msl_malloc": [PP, sz],
    "msl_free": [p],
    "msl_memcpy_h2d": [p, p, sz],
    "msl_memcpy_d2h": [p, p, sz],
    "msl_memcpy_d2d": [p, p, sz],
    "msl_result_slot": [p, p],
    "msl_fill_f32": [p, sz, f32],
    "msl_fill_i32": [p, sz, C.c_int32],
    "msl_fill_uniform_f32": [p, sz, u64, f32],
    "msl_fill_bernoulli_i32": [p, sz, u64, u64],
    "msl_fill_u32": [p, sz, u64],
    "msl_map_affine_f32": [p, p, sz, f32, f32],
    "msl_map_affine_u32": [p, p, sz, C.c_uint32, C.c_uint32],
    "msl_zip_add_f32": [p, p, p, sz],
    "msl_mapindex_lin_f32": [p, i32, i32],
    "msl_fold_sum_f32": [p, sz, p],
    "msl_fold_sum_f32_host": [p, sz, PD, PF],
    "msl_fold_sum_f32_noalloc": [p, sz, p],
    "msl_fold_sum_u32_host": [p, sz, PU32],
    "msl_fold_sum_i32_host": [p, sz, PI64],
    "msl_fold_max_f32_host": [p, sz, PF],
    "msl_stencil2d_mean_f32": [p, p, i32, i32, i32, f32],
    "msl_stencil2d_mean_iterate_f32": [p, p, i32, i32, i32, f32, i32, PI],
    "msl_stencil2d_mean_iterate_plan": [p, p, i32, i32, i32, i32, PI, i32, PI],
    "msl_gol_i32": [p, p, i32, i32],
    "msl_gol_iterate_i32": [p, p, i32, i32, i32, PI],
    "msl_heat7_f32": [p, p, i32, i32, i32, f32],
    "msl_heat7_residual_f32": [p, p, i32, i32, i32, f32, p],
    "msl_heat7_iterate_f32": [p, p, i32, i32, i32, f32, i32, i32, PF, PI],
    "msl_heat7_iterate_plan": [i32, i32, i32, i32, i32, PI, i32, PI],
    "msl_selftest_div_odd_square": [i32, C.POINTER(C.c_uint64)],
    "msl_selftest_div_six": [C.POINTER(C.c_uint64)],
    "msl_graph_begin": [],
    "msl_graph_end": [PP],
    "msl_graph_launch": [p],
    "msl_graph_destroy": [p],
    "msl_batch_create": [PBG, i32, PP],
    "msl_batch_destroy": [p],
    "msl_batch_set_tile": [p, i32, i32, i32, i32],
    "msl_batch_get_tile": [p, i32, i32, PI, PI],
    "msl_batch_set_launch": [p, i32, i32],
    "msl_batch_autotune": [p, C.c_char_p, PI],
    "msl_batch_info": [p, PI, PI, PI, PI],
    "msl_batch_run": [p, i32, p, PI],
    "msl_batch_run_plan": [p, i32, i32, PI, PI, PI, PI],
    "msl_batch_run_loop": [p, i32, p, PI],
    "msl_server_create": [sz, i32, PP],
    "msl_server_destroy": [p],
    "msl_server_submit": [p, PJOB, C.c_double, PI64],
    "msl_server_step": [p, PI, PI],
    "msl_server_drain": [p],
    "msl_server_poll": [p, C.c_int64, PRES, p, sz],
    "msl_server_info": [p, PI64, C.POINTER(C.c_size_t), PD, PI64],
    "msl_server_run_trace": [p, PJOB, PD, i32, PRES],
    "msl_serve_plain_loop": [PJOB, PD, i32, sz, PRES],
    "msl_timer_start": [],
    "msl_timer_stop": [PF],
}
_RESTYPES = {"msl_last_error": C.c_char_p, "msl_stream": C.c_void_p}

_lib: Optional[C.CDLL] = None


def lib_path() -> str:
    return _LIB_PATH


def lib() -> C.CDLL:
    """Load libmsl_hip.so; raise loudly when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError(f"{_LIB_PATH} is missing: build it with `python -m muesli4_amd.build` "
                          "(there is no CPU fallback)")
    L = C.CDLL(_LIB_PATH, mode=C.RTLD_GLOBAL)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = L
    return L


def call(name: str, *args) -> int:
    rc = getattr(lib(), name)(*args)
    if rc != MSL_OK:
        msg = lib().msl_last_error()
        raise MslError(name, rc, msg.decode() if msg else "")
    return rc


def init(device: int = 0) -> None:
    call("msl_init", device)


def synchronize() -> None:
    call("msl_synchronize")


def mean_iterate_plan(d_a, d_b, rows: int, cols: int, r: int, iters: int) -> list:
    """Generations per launch that msl_stencil2d_mean_iterate_f32 would run."""
    cap = max(iters, 1)
    gens = (C.c_int * cap)()
    n = C.c_int()
    call("msl_stencil2d_mean_iterate_plan", d_a, d_b, rows, cols, r, iters, gens, cap, C.byref(n))
    return list(gens)[: n.value]


def heat_iterate_plan(nx: int, ny: int, nz: int, iters: int, every: int) -> list:
    """Generations per launch that msl_heat7_iterate_f32 would run."""
    cap = max(iters, 1)
    gens = (C.c_int * cap)()
    n = C.c_int()
    call("msl_heat7_iterate_plan", nx, ny, nz, iters, every, gens, cap, C.byref(n))
    return list(gens)[: n.value]


def device_info() -> dict:
    cus, ws, mem = C.c_int(), C.c_int(), C.c_size_t()
    name = C.create_string_buffer(128)
    call("msl_device_info", C.byref(cus), C.byref(ws), C.byref(mem), name, 128)
    return {"num_cus": cus.value, "warp_size": ws.value, "total_mem": mem.value, "arch": name.value.decode()}


class DeviceArray:
    """A device buffer owned by Python (msl_malloc / msl_free)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        ptr = C.c_void_p()
        call("msl_malloc", C.byref(ptr), self.nbytes)
        self.ptr = ptr.value

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        if d.nbytes:
            call("msl_memcpy_h2d", d.ptr, arr.ctypes.data, d.nbytes)
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("msl_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def free(self) -> None:
        if getattr(self, "ptr", None):
            call("msl_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and _lib is not None:
                _lib.msl_free(self.ptr)
                self.ptr = None
        except Exception:
            pass

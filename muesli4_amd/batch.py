"""Host-side handle of a batched mean blur (msl_batch_* in include/msl_hip.h).

The grids live in two device arenas (state a and state b), each grid's
start aligned to 64 elements; BatchArena fills arena a from a seed with the
library's counter-based generator, so the initial state of grid g is
`fill_uniform(total, seed)[off[g] : off[g] + rows*cols]`.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import hip
from .hip import BatchGrid, DeviceArray, call
from .workloads import arena_offsets

SCALE_255 = 255.0 / 16777216.0  # values in [0, 255)


class BatchArena:
    def __init__(self, rows: Sequence[int], cols: Sequence[int], radii: Sequence[int], nv=0.0, seed: int = 1,
                 scale: float = SCALE_255):
        self.rows = np.asarray(rows, np.int64)
        self.cols = np.asarray(cols, np.int64)
        self.radii = np.asarray(radii, np.int64)
        n = len(self.rows)
        self.nv = np.broadcast_to(np.asarray(nv, np.float32), (n,)).copy()
        self.offsets, self.total = arena_offsets(self.rows, self.cols)
        self.seed = seed
        self.scale = scale
        alloc = max(self.total, 1)
        self.a = DeviceArray(alloc, np.float32)
        self.b = DeviceArray(alloc, np.float32)
        self.sums = DeviceArray(max(n, 1), np.float64)
        self.reset()
        grids = (BatchGrid * max(n, 1))()
        for g in range(n):
            off = int(self.offsets[g]) * 4
            grids[g] = BatchGrid(self.a.ptr + off, self.b.ptr + off, int(self.rows[g]), int(self.cols[g]),
                                 int(self.radii[g]), float(self.nv[g]))
        self._grids = grids
        h = C.c_void_p()
        call("msl_batch_create", grids, n, C.byref(h))
        self.handle = h.value

    @property
    def n(self) -> int:
        return len(self.rows)

    def elements(self) -> int:
        return int((self.rows * self.cols).sum())

    def algorithmic_bytes(self, iters: int) -> int:
        """8 bytes per cell per step plus 4 bytes per cell for the fold."""
        a = self.rows * self.cols
        return int((8 * a * iters + 4 * a).sum())

    def reset(self) -> None:
        """Refill state a from the seed (b is left as is)."""
        call("msl_fill_uniform_f32", self.a.ptr, max(self.total, 1), self.seed, float(self.scale))

    def run(self, iters: int, fold: bool = True) -> int:
        res = C.c_int()
        call("msl_batch_run", self.handle, iters, self.sums.ptr if fold else None, C.byref(res))
        return res.value

    def run_loop(self, iters: int, fold: bool = True) -> int:
        res = C.c_int()
        call("msl_batch_run_loop", self.handle, iters, self.sums.ptr if fold else None, C.byref(res))
        return res.value

    def autotune(self, cache_path: Optional[str] = None) -> int:
        n = C.c_int()
        call("msl_batch_autotune", self.handle, cache_path.encode() if cache_path else None, C.byref(n))
        return n.value

    def set_tile(self, r: int, width_class: int, cols_per_lane: int, seg: int) -> None:
        call("msl_batch_set_tile", self.handle, r, width_class, cols_per_lane, seg)

    def tile(self, r: int, width_class: int):
        c, s = C.c_int(), C.c_int()
        call("msl_batch_get_tile", self.handle, r, width_class, C.byref(c), C.byref(s))
        return c.value, s.value

    def set_launch(self, blocks_per_cu: int, queue_mode) -> None:
        """queue_mode: 0 static, 1 per-XCD atomic queues, 2 one global queue
        (True/False map to 1/0)."""
        call("msl_batch_set_launch", self.handle, blocks_per_cu, int(queue_mode))

    def info(self) -> dict:
        v = [C.c_int() for _ in range(4)]
        call("msl_batch_info", self.handle, *[C.byref(x) for x in v])
        return {"items": v[0].value, "blocks": v[1].value, "blocks_per_cu": v[2].value, "dynamic": v[3].value}

    def run_plan(self, iters: int, fold: bool = True) -> dict:
        """Launch plan of run(iters): rounds, launches, two-generation grids, items."""
        v = [C.c_int() for _ in range(4)]
        call("msl_batch_run_plan", self.handle, iters, int(fold), *[C.byref(x) for x in v])
        return {"rounds": v[0].value, "launches": v[1].value, "pair_grids": v[2].value, "items": v[3].value}

    def state(self, which: str = "a") -> np.ndarray:
        """Whole arena of state a or b on the host."""
        return (self.a if which == "a" else self.b).to_host()

    def grid(self, arena: np.ndarray, g: int) -> np.ndarray:
        o = int(self.offsets[g])
        return arena[o:o + int(self.rows[g] * self.cols[g])].reshape(int(self.rows[g]), int(self.cols[g]))

    def fold_sums(self) -> np.ndarray:
        return self.sums.to_host()[: self.n].copy()

    def close(self) -> None:
        if getattr(self, "handle", None):
            call("msl_batch_destroy", self.handle)
            self.handle = None
        for d in (getattr(self, "a", None), getattr(self, "b", None), getattr(self, "sums", None)):
            if d is not None:
                d.free()

    def __del__(self):
        try:
            if getattr(self, "handle", None) and hip._lib is not None:
                hip._lib.msl_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

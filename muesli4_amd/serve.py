"""Host-side handle of the serving engine (msl_server_* in include/msl_hip.h)."""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import hip
from .hip import (MSL_JOB_BLUR2D, MSL_JOB_HEAT3D, MSL_JOB_MAPFOLD_U32, Job, JobResult, MslError, call)

MSL_NOT_READY = 1


def make_job(d: dict, keep_final: bool = False) -> Job:
    j = Job()
    j.kind = int(d["kind"])
    j.steps = int(d.get("steps", 0))
    j.r = int(d.get("r", 0))
    j.d0 = int(d.get("d0", 0))
    j.d1 = int(d.get("d1", 0))
    j.d2 = int(d.get("d2", 0))
    j.n = int(d.get("n", 0))
    j.nv = float(d.get("nv", 0.0))
    j.mul = int(d.get("mul", 0)) & 0xFFFFFFFF
    j.add = int(d.get("add", 0)) & 0xFFFFFFFF
    j.seed = int(d["seed"]) & 0xFFFFFFFFFFFFFFFF
    j.keep_final = 1 if keep_final else 0
    return j


def final_shape(d: dict):
    if d["kind"] == MSL_JOB_BLUR2D:
        return (d["d0"], d["d1"]), np.float32
    if d["kind"] == MSL_JOB_HEAT3D:
        return (d["d2"], d["d1"], d["d0"]), np.float32
    return (d["n"],), np.uint32


def job_bytes(d: dict) -> int:
    def al(x):
        return (x + 255) // 256 * 256
    if d["kind"] == MSL_JOB_BLUR2D:
        return 2 * al(d["d0"] * d["d1"] * 4)
    if d["kind"] == MSL_JOB_HEAT3D:
        return 2 * al(d["d0"] * d["d1"] * d["d2"] * 4)
    return al(d["n"] * 4)


def job_traffic_bytes(d: dict) -> int:
    """Algorithmic HBM bytes of a job: 8 B per cell per step plus the 4 B
    fold read and the initial 4 B write; map+fold 4 B write (init) + 8 B."""
    if d["kind"] == MSL_JOB_BLUR2D:
        c = d["d0"] * d["d1"]
        return c * (8 * d["steps"] + 8)
    if d["kind"] == MSL_JOB_HEAT3D:
        c = d["d0"] * d["d1"] * d["d2"]
        return c * (8 * d["steps"] + 8)
    return d["n"] * 12


def results_to_dicts(res: Sequence[JobResult]) -> List[dict]:
    return [dict(id=r.id, kind=r.kind, sum=r.sum, sum_u32=r.sum_u32, t_arrival=r.t_arrival, t_admit=r.t_admit,
                 t_done=r.t_done, bytes=r.bytes) for r in res]


def latency_stats(res: List[dict]) -> dict:
    lat = np.array([r["t_done"] - r["t_arrival"] for r in res])
    t_first = min(r["t_arrival"] for r in res)
    t_last = max(r["t_done"] for r in res)
    span = t_last - t_first
    return {"jobs": len(res), "makespan_s": round(span, 4), "jobs_per_s": round(len(res) / span, 1),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "mean_ms": round(1e3 * float(lat.mean()), 3)}


class Server:
    def __init__(self, budget_bytes: int = 16 << 30, max_active: int = 1024):
        h = C.c_void_p()
        call("msl_server_create", budget_bytes, max_active, C.byref(h))
        self.handle = h.value
        self.budget = budget_bytes

    def submit(self, d: dict, keep_final: bool = False, t_arrival: float = -1.0) -> int:
        j = make_job(d, keep_final)
        jid = C.c_int64()
        call("msl_server_submit", self.handle, C.byref(j), t_arrival, C.byref(jid))
        return jid.value

    def step(self):
        a, p = C.c_int(), C.c_int()
        call("msl_server_step", self.handle, C.byref(a), C.byref(p))
        return a.value, p.value

    def drain(self) -> None:
        call("msl_server_drain", self.handle)

    def poll(self, jid: int, d: Optional[dict] = None):
        """Result dict (and the final state when d is given and the job kept
        it), or None while the job is not done."""
        res = JobResult()
        final = None
        buf, nbytes = None, 0
        if d is not None:
            shape, dtype = final_shape(d)
            final = np.empty(shape, dtype)
            buf, nbytes = final.ctypes.data, final.nbytes
        rc = hip.lib().msl_server_poll(self.handle, jid, C.byref(res), buf, nbytes)
        if rc == MSL_NOT_READY:
            return None
        if rc != 0:
            msg = hip.lib().msl_last_error()
            raise MslError("msl_server_poll", rc, msg.decode() if msg else "")
        out = results_to_dicts([res])[0]
        if final is not None:
            out["final"] = final
        return out

    def info(self) -> dict:
        rounds, peak, mean_active, done = C.c_int64(), C.c_size_t(), C.c_double(), C.c_int64()
        call("msl_server_info", self.handle, C.byref(rounds), C.byref(peak), C.byref(mean_active), C.byref(done))
        return {"rounds": rounds.value, "peak_bytes": peak.value, "mean_active_jobs": round(mean_active.value, 2),
                "done_jobs": done.value}

    def run_trace(self, jobs: List[dict], arrivals) -> List[dict]:
        n = len(jobs)
        arr = (Job * max(n, 1))(*[make_job(d) for d in jobs])
        times = (C.c_double * max(n, 1))(*[float(t) for t in arrivals])
        res = (JobResult * max(n, 1))()
        call("msl_server_run_trace", self.handle, arr, times, n, res)
        return results_to_dicts(res[:n])

    def close(self) -> None:
        if getattr(self, "handle", None):
            call("msl_server_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            if getattr(self, "handle", None) and hip._lib is not None:
                hip._lib.msl_server_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def plain_loop(jobs: List[dict], arrivals, budget_bytes: int = 16 << 30) -> List[dict]:
    n = len(jobs)
    arr = (Job * max(n, 1))(*[make_job(d) for d in jobs])
    times = (C.c_double * max(n, 1))(*[float(t) for t in arrivals])
    res = (JobResult * max(n, 1))()
    call("msl_serve_plain_loop", arr, times, n, budget_bytes, res)
    return results_to_dicts(res[:n])

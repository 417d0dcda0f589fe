#!/usr/bin/env python3
"""Benchmark of the single-device skeleton path on one MI355X.

Headline workload (one "step"): DM<float> 8192 x 8192, values uniform in
[0,255) from a seed; mean-blur mapStencil at radius 1, 5 and 10 for 20
iterations each (ping-pong, graph-captured), plus map(x*1.5f+3.f),
zip(a+b), mapIndex(row*cols+col) and the compensated fold(sum).
``value`` is the mean-blur stencil throughput in algorithmic GB/s
(8 bytes per cell per iteration: one read, one write) with the grids
resident in HBM.  Sub-results cover the other skeletons and workloads.

Prints ONE JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s (1e9 B/s)
# wave64 FP32 VALU instructions per second: 1024 SIMDs x 2.4 GHz / 4 cycles
# each (packed FP32 instructions are not used)
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4
ROWS = COLS = 8192
RADII = (1, 5, 10)
BLUR_ITERS = 20
SEED = 20240611
BATCH_GRIDS = 4096
BATCH_ITERS = 10
BATCH_SEED = 2024
SERVE_JOBS = 400
SERVE_SEED = 2025
SERVE_BUDGET = 64 << 30


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _profile_json(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def radius_figures(hip, a, b, r, t_ms):
    """Figures of one radius of the headline: the event-timed 20-iteration
    call (t_ms), its launch plan, the HBM bytes the rocprofv3 counters saw
    for one such call and its vector-ALU issue floor (SQ_INSTS_VALU)."""
    n = ROWS * COLS
    plan = hip.mean_iterate_plan(a.ptr, b.ptr, ROWS, COLS, r, BLUR_ITERS)
    gens = max(plan)
    kname = ("mean3_multi_kernel<6, 16>" if r == 1 and gens > 1 else
             f"strip2_mean_kernel<{r}, 0, true>" if gens == 2 else
             f"strip_stencil_kernel<MeanPolicy<{r}>, 0, true, 4>")
    out = {"kernel": kname, "plan_generations_per_launch": plan, "plan": plan,
           "ms_per_iteration": round(t_ms / BLUR_ITERS, 5),
           "effective_GBs": round(8 * n * BLUR_ITERS / (t_ms * 1e-3) / 1e9, 1),
           "hbm_bytes_per_iteration": None}
    tr = _profile_json(f"traffic_iterate_r{r}_8192.json")
    hbm_frac = None
    if tr:
        per_call = tr["hbm_bytes_per_call"]
        gbs = per_call / (t_ms * 1e-3) / 1e9
        hbm_frac = gbs / HBM_PEAK_GBS
        out.update({"hbm_bytes_per_iteration": per_call // BLUR_ITERS,
                    "hbm_bytes_per_launch": per_call // len(plan),
                    "achieved_hbm_GBs": round(gbs, 1), "pct_hbm_peak_achieved": round(100 * hbm_frac, 1)})
    va = _profile_json(f"valu_iterate_r{r}_8192.json")
    if va:
        floor_ms = va["valu_wave_instructions_per_call"] / VALU_ISSUE_PER_S * 1e3
        out.update({"valu_lane_instructions_per_cell_generation": va["valu_lane_instructions_per_cell_generation"],
                    "valu_floor_ms_per_iteration": round(floor_ms / BLUR_ITERS, 5),
                    "valu_floor_ms_per_launch": round(floor_ms / len(plan), 5),
                    "valu_frac": round(floor_ms / t_ms, 4)})
    if hbm_frac is not None and "valu_frac" in out:
        out["bound"] = "valu" if out["valu_frac"] > hbm_frac else "hbm"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only-headline", action="store_true", help="skip the sub-workloads")
    ap.add_argument("--kernel", default=None,
                    help="run one kernel only, --reps times, for profiler counter runs: "
                         "map | zip | fold | blur_r1 | blur_r5 | blur_r10 | iter_r1 | iter_r5 | iter_r10 | "
                         "blur_r1_multi | blur_r5_pairs | heat7 | heat7_pairs | heat_iter | gol | gol_multi | batch")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    if args.kernel:
        run_single_kernel(args.kernel, args.reps)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part "
              "of this build", file=sys.stderr)
        sys.exit(2)

    from muesli4_amd import hip
    from muesli4_amd.hip import DeviceArray, call

    hip.init(0)
    info = hip.device_info()
    log(f"[bench] device {info}")
    n = ROWS * COLS
    a = DeviceArray((ROWS, COLS), np.float32)
    b = DeviceArray((ROWS, COLS), np.float32)
    c = DeviceArray((ROWS, COLS), np.float32)
    d_res = DeviceArray(1, np.float64)
    scale = float(np.float32(255.0 / 16777216.0))
    inb = C.c_int()
    ms = C.c_float()

    # ---- per-kernel timings: each launch timed alone with events on the
    # library stream, after a fold that reads 1 GiB and so evicts the grids
    # from the 256 MiB Infinity Cache (non-temporal stores, as in the fills,
    # do not allocate there and would leave the inputs cached): every launch
    # reads its inputs from HBM
    nflush = 1 << 28
    flush = DeviceArray(nflush, np.float32)
    call("msl_fill_f32", flush.ptr, nflush, 0.0)

    def timed_cold(fn, reps):
        for _ in range(3):  # first launches load the code object / grow the workspace
            fn()
        total = 0.0
        for _ in range(reps):
            call("msl_fold_sum_f32", flush.ptr, nflush, d_res.ptr)
            call("msl_timer_start")
            fn()
            call("msl_timer_stop", C.byref(ms))
            total += ms.value
        return total / reps

    def fresh():
        call("msl_fill_uniform_f32", a.ptr, n, SEED, scale)
        call("msl_fill_uniform_f32", b.ptr, n, SEED + 1, scale)

    fresh()
    hip.synchronize()
    reps = 20
    kernel_ms = {}
    kernel_ms["map"] = timed_cold(lambda: call("msl_map_affine_f32", c.ptr, a.ptr, n, 1.5, 3.0), reps)
    kernel_ms["zip"] = timed_cold(lambda: call("msl_zip_add_f32", c.ptr, a.ptr, b.ptr, n), reps)
    kernel_ms["mapindex"] = timed_cold(lambda: call("msl_mapindex_lin_f32", c.ptr, ROWS, COLS), reps)
    kernel_ms["fill"] = timed_cold(lambda: call("msl_fill_f32", c.ptr, n, 1.0), reps)
    kernel_ms["fold"] = timed_cold(lambda: call("msl_fold_sum_f32", a.ptr, n, d_res.ptr), reps)
    for r in RADII:
        kernel_ms[f"blur_r{r}"] = timed_cold(lambda: call("msl_stencil2d_mean_f32", c.ptr, a.ptr, ROWS, COLS, r, 0.0), reps)
    kernel_bytes = {"map": 8 * n, "zip": 12 * n, "mapindex": 4 * n, "fill": 4 * n, "fold": 4 * n,
                    **{f"blur_r{r}": 8 * n for r in RADII}}
    kernels = {k: {"ms": round(v, 5), "GBs": round(kernel_bytes[k] / (v * 1e-3) / 1e9, 1),
                   "pct_hbm_peak": round(100 * kernel_bytes[k] / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 1)}
               for k, v in kernel_ms.items()}
    log(f"[bench] kernels {kernels}")
    flush.free()

    # ---- fold error of the headline grid: the compensated FP32 fold (FP64
    # final combine) against an FP64 (pairwise) sum of the same floats
    fresh()
    s64, s32 = C.c_double(), C.c_float()
    call("msl_fold_sum_f32_host", a.ptr, n, C.byref(s64), C.byref(s32))
    host = a.to_host().ravel()
    ref64 = float(np.sum(host, dtype=np.float64))
    fold_check = {"fold_sum_f64_combine": s64.value, "fold_sum_f32": float(s32.value), "fp64_reference": ref64,
                  "rel_err": abs(s64.value - ref64) / abs(ref64), "tolerance": 1e-6}
    del host
    log(f"[bench] fold check {fold_check}")

    # ---- headline step ------------------------------------------------------
    def step():
        call("msl_map_affine_f32", c.ptr, a.ptr, n, 1.5, 3.0)
        call("msl_zip_add_f32", c.ptr, a.ptr, b.ptr, n)
        call("msl_mapindex_lin_f32", c.ptr, ROWS, COLS)
        call("msl_fold_sum_f32", a.ptr, n, d_res.ptr)
        for r in RADII:
            call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, ROWS, COLS, r, 0.0, BLUR_ITERS, C.byref(inb))

    for _ in range(args.warmup):
        fresh()
        step()
    hip.synchronize()
    step_ms = []
    blur_ms = {r: [] for r in RADII}
    for _ in range(args.steps):
        fresh()
        hip.synchronize()
        t0 = time.perf_counter()
        call("msl_map_affine_f32", c.ptr, a.ptr, n, 1.5, 3.0)
        call("msl_zip_add_f32", c.ptr, a.ptr, b.ptr, n)
        call("msl_mapindex_lin_f32", c.ptr, ROWS, COLS)
        call("msl_fold_sum_f32", a.ptr, n, d_res.ptr)
        for r in RADII:
            # events on the library stream around the 20 graph-replayed steps
            call("msl_timer_start")
            call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, ROWS, COLS, r, 0.0, BLUR_ITERS, C.byref(inb))
            call("msl_timer_stop", C.byref(ms))
            blur_ms[r].append(ms.value)
        hip.synchronize()
        t1 = time.perf_counter()
        step_ms.append((t1 - t0) * 1e3)
    med = {r: float(np.median(v)) for r, v in blur_ms.items()}
    blur_bytes = 8 * n * BLUR_ITERS * len(RADII)
    value = blur_bytes / (sum(med.values()) * 1e-3) / 1e9
    per_radius = {f"r{r}": radius_figures(hip, a, b, r, t) for r, t in med.items()}
    hbm_total = [per_radius[f"r{r}"]["hbm_bytes_per_iteration"] for r in RADII]
    hbm_achieved = None
    if all(x is not None for x in hbm_total):
        gbs = BLUR_ITERS * sum(hbm_total) / (sum(med.values()) * 1e-3) / 1e9
        hbm_achieved = {"GBs": round(gbs, 1), "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 1),
                        "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE bytes of one 20-iteration call per radius "
                                  "(profiles/traffic_iterate_r*_8192.json) over the event-timed iterations"}
    log(f"[bench] headline blur effective GB/s {value:.1f} achieved {hbm_achieved} per radius {per_radius} "
        f"step ms {np.median(step_ms):.3f}")

    # ---- roofline of the dominant kernel: the radius with the largest share
    # of the timed region; average launch duration = its 20-step time divided
    # by the launches of its plan.
    dom_r = max(med, key=lambda r: med[r])
    pr = per_radius[f"r{dom_r}"]
    gens = max(pr["plan"])
    launch_ms = med[dom_r] / len(pr["plan"]) if gens == 1 else pr["ms_per_iteration"] * gens
    alg_per_launch = 8 * n * gens
    achieved = alg_per_launch / (launch_ms * 1e-3) / 1e9
    traffic = pr.get("hbm_bytes_per_launch")
    hbm_frac = (traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else achieved / HBM_PEAK_GBS
    valu_frac = pr.get("valu_frac")
    bound = "valu" if (valu_frac is not None and valu_frac > hbm_frac) else "hbm"
    roofline = {"bound": bound,
                "kernel": f"{pr['kernel']} (8192x8192 ping-pong, 8 B/cell per generation)",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": alg_per_launch, "avg_launch_ms": round(launch_ms, 5),
                "hbm_achieved_GBs": round(hbm_frac * HBM_PEAK_GBS, 1), "hbm_frac": round(hbm_frac, 4),
                "valu_frac": valu_frac, "valu_floor_ms_per_launch": pr.get("valu_floor_ms_per_launch"),
                "note": "achieved/frac: algorithmic bytes per launch over the event-timed launch; "
                        "hbm_achieved/hbm_frac: rocprofv3 FETCH_SIZE+WRITE_SIZE bytes over the same time; "
                        "valu_frac: SQ_INSTS_VALU issue floor over the same time (bound = the larger)"}

    result = {
        "metric": "stencil GB/s (% HBM peak)",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.median(step_ms)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (SplitMix64-seeded uniform [0,255))",
        "config": {"workload": "DM<float> 8192x8192 mean-blur mapStencil r=1,5,10 x 20 iterations each "
                               "(+ map x*1.5f+3.f, zip a+b, mapIndex row*cols+col, compensated fold sum)",
                   "rows": ROWS, "cols": COLS, "radii": list(RADII), "iterations": BLUR_ITERS},
        "value_kind": "effective: 8 B per cell per generation (one read, one write), so launches that fuse "
                      "generations (six at r=1, two at r=5) count more bytes than they move; hbm_achieved "
                      "gives the bytes the rocprofv3 counters see",
        "pct_hbm_peak_effective": round(100 * value / HBM_PEAK_GBS, 1),
        "hbm_achieved": hbm_achieved,
        "roofline": roofline,
        "per_radius": per_radius,
        "kernels": kernels,
        "fold_check": fold_check,
        "kernels_note": "each launch timed alone (HIP events on the library stream) after a 1 GiB read that "
                        "evicts the 256 MiB Infinity Cache; blur kernels write a third grid",
        "device": info,
    }

    extra = {}  # GPU outputs that the CPU legs compare with (final states, sums)
    if not args.only_headline:
        result["gol"] = bench_gol(hip, call, DeviceArray)
        result["heat3d"] = bench_heat(hip, call, DeviceArray, extra)
        result["batched"] = bench_batch(hip, extra)
        result["serving"] = bench_serve(hip, extra)
        result["cpp_api"] = bench_cpp_api()

    result["cpu_baseline"] = None
    if not args.no_cpu:
        cpu_baselines(result, value, kernel_ms, extra)
    print(json.dumps(result), flush=True)


def run_single_kernel(name, reps):
    """Launch one kernel of the headline workloads `reps` times (after two
    warm-up launches) so that profiler runs see exactly that kernel.  Stencil
    kernels ping-pong between two grids as in the iteration loops."""
    from muesli4_amd import hip
    from muesli4_amd.hip import DeviceArray, call

    hip.init(0)
    if name == "batch":
        from muesli4_amd.batch import BatchArena
        from muesli4_amd.workloads import batch_shapes
        rows, cols, radii = batch_shapes(BATCH_GRIDS, BATCH_SEED)
        bt = BatchArena(rows, cols, radii, nv=0.0, seed=BATCH_SEED)
        bt.autotune(None)
        for _ in range(max(1, reps // BATCH_ITERS)):
            bt.run(BATCH_ITERS)
        hip.synchronize()
        log(f"[bench] ran batch x{max(1, reps // BATCH_ITERS)} runs of {BATCH_ITERS} steps")
        return
    if name in ("heat7", "heat7_pairs", "heat_iter"):
        nx = ny = nz = 512
        n = nx * ny * nz
        a = DeviceArray(n, np.float32)
        b = DeviceArray(n, np.float32)
        call("msl_fill_uniform_f32", a.ptr, n, SEED, float(np.float32(1.0 / 16777216.0)))
        bufs = [a, b]
        inb = C.c_int()
        for i in range(reps + 2):
            if name == "heat7":
                call("msl_heat7_f32", bufs[(i + 1) % 2].ptr, bufs[i % 2].ptr, nx, ny, nz, 0.0)
            elif name == "heat_iter":  # the 200-iteration run with a residual every 10 (the heat3d graph)
                res = (C.c_float * 20)()
                call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, 200, 10, res, C.byref(inb))
            else:  # two launches of two generations each per call (four keeps the buffer parity)
                call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, 4, 0, None, C.byref(inb))
    elif name in ("gol", "gol_multi"):
        rows = cols = 1024
        n = rows * cols
        a = DeviceArray(n, np.int32)
        b = DeviceArray(n, np.int32)
        call("msl_fill_bernoulli_i32", a.ptr, n, SEED, int(0.3 * 2.0 ** 64))
        bufs = [a, b]
        inb = C.c_int()
        for i in range(reps + 2):
            if name == "gol":
                call("msl_gol_i32", bufs[(i + 1) % 2].ptr, bufs[i % 2].ptr, rows, cols)
            else:  # one launch of four generations per call
                call("msl_gol_iterate_i32", a.ptr, b.ptr, rows, cols, 4, C.byref(inb))
    else:
        n = ROWS * COLS
        a = DeviceArray(n, np.float32)
        b = DeviceArray(n, np.float32)
        c = DeviceArray(n, np.float32)
        d_res = DeviceArray(1, np.float64)
        scale = float(np.float32(255.0 / 16777216.0))
        call("msl_fill_uniform_f32", a.ptr, n, SEED, scale)
        call("msl_fill_uniform_f32", b.ptr, n, SEED + 1, scale)
        bufs = [a, b]
        for i in range(reps + 2):
            if name == "map":
                call("msl_map_affine_f32", c.ptr, a.ptr, n, 1.5, 3.0)
            elif name == "zip":
                call("msl_zip_add_f32", c.ptr, a.ptr, b.ptr, n)
            elif name == "fold":
                call("msl_fold_sum_f32", a.ptr, n, d_res.ptr)
            elif name.startswith("iter_r"):  # one 20-iteration call (the headline's graph) per rep
                inb = C.c_int()
                r = int(name[len("iter_r"):])
                call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, ROWS, COLS, r, 0.0, BLUR_ITERS, C.byref(inb))
            elif name == "blur_r1_multi":  # one launch of six radius-1 generations per call
                inb = C.c_int()
                call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, ROWS, COLS, 1, 0.0, 6, C.byref(inb))
            elif name == "blur_r5_pairs":  # two launches of two radius-5 generations per call
                inb = C.c_int()
                call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, ROWS, COLS, 5, 0.0, 4, C.byref(inb))
            elif name.startswith("blur_r"):
                r = int(name[len("blur_r"):])
                call("msl_stencil2d_mean_f32", bufs[(i + 1) % 2].ptr, bufs[i % 2].ptr, ROWS, COLS, r, 0.0)
            else:
                raise SystemExit(f"unknown kernel {name}")
    hip.synchronize()
    log(f"[bench] ran {name} x{reps + 2}")


def bench_gol(hip, call, DeviceArray):
    rows = cols = 1024
    iters = 100
    n = rows * cols
    a = DeviceArray((rows, cols), np.int32)
    b = DeviceArray((rows, cols), np.int32)
    inb = C.c_int()
    ms = C.c_float()
    thr = int(0.3 * 2.0 ** 64)
    call("msl_fill_bernoulli_i32", a.ptr, n, SEED, thr)
    call("msl_gol_iterate_i32", a.ptr, b.ptr, rows, cols, iters, C.byref(inb))  # capture + warm
    times = []
    for _ in range(5):
        call("msl_fill_bernoulli_i32", a.ptr, n, SEED, thr)
        call("msl_timer_start")
        call("msl_gol_iterate_i32", a.ptr, b.ptr, rows, cols, iters, C.byref(inb))
        s = C.c_int64()
        fin = b if inb.value else a
        call("msl_fold_sum_i32_host", fin.ptr, n, C.byref(s))
        call("msl_timer_stop", C.byref(ms))
        times.append(ms.value)
    t = float(np.median(times)) * 1e-3
    out = {"workload": "GoL DM<int> 1024x1024 B3/S23, 100 iterations + fold(sum)", "ms": round(t * 1e3, 3),
           "cells_per_s": round(n * iters / t, 1), "effective_GBs": round(8 * n * iters / t / 1e9, 1),
           "live_cells": int(s.value), "note": "8 MiB working set is cache-resident; % HBM not meaningful"}
    log(f"[bench] gol {out}")
    return out


def bench_heat(hip, call, DeviceArray, extra):
    nx = ny = nz = 512
    iters, every = 200, 10
    n = nx * ny * nz
    a = DeviceArray((nz, ny, nx), np.float32)
    b = DeviceArray((nz, ny, nx), np.float32)
    res = (C.c_float * (iters // every))()
    inb = C.c_int()
    ms = C.c_float()
    unit = float(np.float32(1.0 / 16777216.0))
    call("msl_fill_uniform_f32", a.ptr, n, SEED, unit)
    call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, iters, every, res, C.byref(inb))
    times = []
    for _ in range(3):
        call("msl_fill_uniform_f32", a.ptr, n, SEED, unit)
        call("msl_timer_start")
        call("msl_heat7_iterate_f32", a.ptr, b.ptr, nx, ny, nz, 0.0, iters, every, res, C.byref(inb))
        call("msl_timer_stop", C.byref(ms))
        times.append(ms.value)
    t = float(np.median(times)) * 1e-3
    # algorithmic bytes: 8 B/cell/step (residual fused: read of the old value is the same read)
    gbs = 8 * n * iters / t / 1e9
    out = {"workload": "DC<float> 512^3 heat7 (sum of 6 neighbours / 6, nv=0), 200 iterations, "
                       "fused max|delta| residual every 10, hipGraph loop",
           "ms": round(t * 1e3, 3), "GBs": round(gbs, 1), "pct_hbm_peak_effective": round(100 * gbs / HBM_PEAK_GBS, 1),
           "plan_generations_per_launch": hip.heat_iterate_plan(nx, ny, nz, iters, every),
           "residuals": [float(x) for x in res]}
    tr = _profile_json("traffic_heat7_iterate200_512.json")
    if tr:
        hb = tr["hbm_bytes_per_call"] / t / 1e9
        out.update({"achieved_hbm_GBs": round(hb, 1), "pct_hbm_peak_achieved": round(100 * hb / HBM_PEAK_GBS, 1)})
    va = _profile_json("valu_heat7_iterate200_512.json")
    if va:
        out["valu_frac"] = round(va["valu_wave_instructions_per_call"] / VALU_ISSUE_PER_S / t, 4)
    extra["heat_final"] = (b if inb.value else a).to_host()
    log(f"[bench] heat {out['GBs']} GB/s")
    return out


def bench_batch(hip, extra):
    """4096 independent grids (long side log-uniform in [32,1024], aspect up
    to 4:1, r uniform in 1..8), 10 mean-blur steps each plus a per-grid
    fold(sum): launch rounds in one graph (two generations per item for
    r <= 8) vs the plain loop of one launch per grid per step and one fold
    per grid.  achieved_hbm_GBs: the HBM bytes rocprofv3 counted for one
    run (profiles/traffic_batch_4096grids.json) over the measured time."""
    from muesli4_amd.batch import BatchArena
    from muesli4_amd.workloads import batch_shapes

    rows, cols, radii = batch_shapes(BATCH_GRIDS, BATCH_SEED)
    bt = BatchArena(rows, cols, radii, nv=0.0, seed=BATCH_SEED)
    t0 = time.perf_counter()
    tuned = bt.autotune(None)
    tune_s = time.perf_counter() - t0
    ms = C.c_float()

    def timed(fn, reps):
        out = []
        for _ in range(reps):
            bt.reset()
            call = hip.call
            call("msl_timer_start")
            fn()
            call("msl_timer_stop", C.byref(ms))
            out.append(ms.value)
        return float(np.median(out))

    bt.run(BATCH_ITERS)  # capture the graph
    persistent_ms = timed(lambda: bt.run(BATCH_ITERS), 5)
    bt.run_loop(BATCH_ITERS)  # warm the single-grid kernels
    loop_ms = timed(lambda: bt.run_loop(BATCH_ITERS), 2)
    nbytes = bt.algorithmic_bytes(BATCH_ITERS)
    info = bt.info()
    plan = bt.run_plan(BATCH_ITERS)
    tiles = {f"r{r}": [bt.tile(r, wc) for wc in range(4)] for r in range(1, 9)}
    out = {"workload": f"{BATCH_GRIDS} DM<float> grids, sides log-uniform [32,1024], aspect 1-4, r in 1..8, "
                       f"{BATCH_ITERS} mean-blur steps + per-grid compensated fold(sum)",
           "grids": BATCH_GRIDS, "elements": bt.elements(), "algorithmic_bytes": nbytes,
           "persistent_ms": round(persistent_ms, 3), "loop_ms": round(loop_ms, 3),
           "speedup_vs_loop": round(loop_ms / persistent_ms, 2),
           "persistent_GBs": round(nbytes / (persistent_ms * 1e-3) / 1e9, 1),
           "achieved_hbm_GBs": (round(_profile_json("traffic_batch_4096grids.json")["hbm_bytes_per_run"]
                                      / (persistent_ms * 1e-3) / 1e9, 1)
                                if _profile_json("traffic_batch_4096grids.json") else None),
           "loop_GBs": round(nbytes / (loop_ms * 1e-3) / 1e9, 1),
           "pct_hbm_peak": round(100 * nbytes / (persistent_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 1),
           "launches_loop": int(BATCH_GRIDS * (BATCH_ITERS + 1)), "launches_persistent_graph": plan["launches"],
           "run_plan": plan,
           "autotune_s": round(tune_s, 2), "autotuned_classes": tuned, "launch": info,
           "tiles_cols_per_lane_seg_by_width_class": tiles}
    bt.reset()
    bt.run(BATCH_ITERS)
    extra["batch_sums"] = bt.fold_sums()
    log(f"[bench] batched {out}")
    bt.close()
    return out


def bench_cpp_api():
    """User functors through the C++ skeleton API (the reference examples'
    heat3D, GoL and mean-blur functors, map/zip/fold functors) beside the
    library kernels, measured by the compiled skeleton_bench program."""
    import subprocess
    exe = os.path.join(ROOT, "muesli4_amd", "bench_cpp", "build", "skeleton_bench")
    if not os.path.exists(exe):
        return {"error": f"{exe} not built"}
    res = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    if res.returncode != 0:
        return {"error": f"skeleton_bench exited with {res.returncode}: {res.stderr[-500:]}"}
    out = json.loads(res.stdout.strip().splitlines()[-1])
    log(f"[bench] cpp_api {out}")
    return out


def bench_serve(hip, extra):
    """Serving: a seeded mix of 2-D mean-blur jobs (sides log-uniform in
    [64,2048], r 1..10, 10..500 steps), 3-D heat jobs (side 32..256, 10..100
    steps) and u32 map+fold jobs (length log-uniform in [1e4,1e8]).
    Throughput: every job arrives at t=0 (saturation), continuous-batching
    engine vs the plain loop (one job at a time, one launch + sync per
    step).  Latency: Poisson arrivals at half of the plain loop's capacity,
    served by both, and at half of the engine's own capacity."""
    from muesli4_amd.serve import Server, job_traffic_bytes, latency_stats, plain_loop
    from muesli4_amd.workloads import serve_trace

    jobs, _ = serve_trace(SERVE_JOBS, SERVE_SEED, 0.0)
    zeros = np.zeros(len(jobs))
    srv = Server(SERVE_BUDGET, 1024)
    srv.run_trace(jobs[:20], zeros[:20])  # warm-up: code objects, pinned buffers
    res = srv.run_trace(jobs, zeros)
    info = srv.info()
    sat = latency_stats(res)
    base = plain_loop(jobs, zeros, SERVE_BUDGET)
    base_sat = latency_stats(base)
    mismatches = sum(1 for a, b in zip(res, base)
                     if (a["kind"] == 3 and a["sum_u32"] != b["sum_u32"])
                     or (a["kind"] != 3 and abs(a["sum"] - b["sum"]) > 1e-5 * abs(b["sum"])))
    traffic = sum(job_traffic_bytes(d) for d in jobs)
    rate_low = 0.5 * base_sat["jobs_per_s"]
    jobs_l, arr_l = serve_trace(SERVE_JOBS, SERVE_SEED, rate_low)
    lat_srv_low = latency_stats(srv.run_trace(jobs_l, arr_l))
    lat_base_low = latency_stats(plain_loop(jobs_l, arr_l, SERVE_BUDGET))
    rate_high = 0.5 * sat["jobs_per_s"]
    jobs_h, arr_h = serve_trace(SERVE_JOBS, SERVE_SEED, rate_high)
    lat_srv_high = latency_stats(srv.run_trace(jobs_h, arr_h))
    srv.close()
    out = {"workload": f"{SERVE_JOBS} jobs: 50% 2-D mean blur (sides log-U[64,2048], r 1..10, 10..500 steps), "
                       "25% 3-D heat7 (side 32..256, 10..100 steps), 25% u32 map(x*3+1)+fold "
                       "(length log-U[1e4,1e8]); seeded",
           "saturated": {"engine": sat, "plain_loop": base_sat,
                         "speedup_jobs_per_s": round(sat["jobs_per_s"] / base_sat["jobs_per_s"], 2),
                         "engine_GBs": round(traffic / sat["makespan_s"] / 1e9, 1),
                         "result_mismatches_vs_plain_loop": mismatches},
           "poisson_at_half_plain_capacity": {"rate_per_s": round(rate_low, 1), "engine": lat_srv_low,
                                              "plain_loop": lat_base_low},
           "poisson_at_half_engine_capacity": {"rate_per_s": round(rate_high, 1), "engine": lat_srv_high},
           "engine_info": info, "hbm_budget_bytes": SERVE_BUDGET}
    log(f"[bench] serving {out}")
    return out


# ---- CPU baseline legs -------------------------------------------------------
# The oracle (OpenMP C, test infrastructure) timed on this host's cores beside
# every GPU figure: rank 0, one GPU, bounded samples of the same workloads.
def _cpu_time(fn, reps=5, budget_s=5.0):
    """Median wall time of fn() over up to `reps` runs (one warm-up), stopping
    once `budget_s` seconds have been spent."""
    fn()
    ts = []
    t_start = time.perf_counter()
    while len(ts) < reps and (not ts or time.perf_counter() - t_start < budget_s):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), len(ts)


def cpu_baselines(result, gpu_value, kernel_ms, extra):
    from oracle import refcpu as R
    threads = R.num_threads()
    log(f"[bench] cpu legs on {threads} OpenMP threads")
    result["cpu_baseline"] = cpu_headline(R, threads, gpu_value)
    result["cpu_baseline_kernels"] = cpu_kernels(R, threads, kernel_ms)
    if "gol" in result:
        result["gol"]["cpu_baseline"] = cpu_gol(R, threads, result["gol"])
    if "heat3d" in result:
        result["heat3d"]["cpu_baseline"] = cpu_heat(R, threads, result["heat3d"], extra)
    if "batched" in result:
        result["batched"]["cpu_baseline"] = cpu_batch(R, threads, result["batched"], extra)
    if "serving" in result:
        result["serving"]["cpu_baseline"] = cpu_serve(R, threads, result["serving"], extra)
    if "cpp_api" in result and isinstance(result["cpp_api"], dict):
        kb = result["cpu_baseline_kernels"]
        pr = result["cpu_baseline"]["per_radius_GBs"]
        result["cpp_api"]["cpu_baseline"] = {
            "unit": "GB/s (algorithmic) or us/step", "cores": threads, "kind": "port",
            "map": kb["map"]["GBs"], "zip": kb["zip"]["GBs"], "fold": kb["fold"]["GBs"],
            "blur_r1_one_step": pr["r1"],
            "heat7_512_one_step": result.get("heat3d", {}).get("cpu_baseline", {}).get("GBs"),
            "gol_1024_us_per_step": result.get("gol", {}).get("cpu_baseline", {}).get("us_per_step"),
            "sample": "the same workloads timed in the other CPU legs (oracle, same sizes)"}


def cpu_headline(R, threads, gpu_value):
    """Same metric and mix as the GPU headline: full 8192x8192 grid, one
    mean-blur iteration at each of r=1, 5, 10 per round, rounds repeated until
    about 10 s have elapsed (at most 100)."""
    n = ROWS * COLS
    x = R.fill_uniform_f32(n, SEED).reshape(ROWS, COLS)
    y = np.empty_like(x)
    R.stencil2d_mean_f32(x[:16], 1, 0.0)  # warm the thread pool
    per_t = {r: [] for r in RADII}
    rounds = 0
    total_t = 0.0
    while rounds < 100 and (rounds == 0 or total_t < 10.0):
        for r in RADII:
            t0 = time.perf_counter()
            R.stencil2d_mean_f32(x, r, 0.0, out=y)
            dt = time.perf_counter() - t0
            per_t[r].append(dt)
            total_t += dt
        rounds += 1
    med = {r: float(np.median(v)) for r, v in per_t.items()}
    v = 8 * n * len(RADII) / sum(med.values()) / 1e9
    log(f"[bench] cpu headline {v:.2f} GB/s")
    return {"value": round(v, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"oracle mean-blur on the full 8192x8192 grid, one iteration at each of r=1,5,10 per round, "
                      f"{rounds} round(s), {total_t:.1f} s wall time on {threads} threads; value = 3 x 8 B/cell / sum of per-radius "
                      "median iteration times (the GPU headline's equal-iterations mix)",
            "per_radius_GBs": {f"r{r}": round(8 * n / t / 1e9, 3) for r, t in med.items()},
            "gpu_over_cpu": round(gpu_value / v, 1) if v > 0 else None}


def cpu_kernels(R, threads, kernel_ms):
    """map / zip / mapIndex / fold on the 8192x8192 grid, each the median of
    up to 5 runs."""
    n = ROWS * COLS
    x = R.fill_uniform_f32(n, SEED)
    y = R.fill_uniform_f32(n, SEED + 1)
    lib = R.lib()
    out = np.empty_like(x)
    legs = {
        "map": (lambda: lib.orc_map_affine_f32(out.ctypes.data, x.ctypes.data, n, 1.5, 3.0), 8 * n),
        "zip": (lambda: lib.orc_zip_add_f32(out.ctypes.data, x.ctypes.data, y.ctypes.data, n), 12 * n),
        "mapindex": (lambda: lib.orc_mapindex_lin_f32(out.ctypes.data, ROWS, COLS), 4 * n),
        "fold": (lambda: lib.orc_fold_sum_f64(x.ctypes.data, n), 4 * n),
    }
    res = {}
    for k, (fn, nbytes) in legs.items():
        t, reps = _cpu_time(fn)
        gbs = nbytes / t / 1e9
        gpu_gbs = nbytes / (kernel_ms[k] * 1e-3) / 1e9
        res[k] = {"ms": round(t * 1e3, 3), "GBs": round(gbs, 2), "runs": reps, "gpu_over_cpu": round(gpu_gbs / gbs, 1)}
    res["cores"] = threads
    res["kind"] = "port"
    res["sample"] = "oracle loops on the full 8192x8192 grid (fold: the FP64 oracle sum), median of up to 5 runs"
    log(f"[bench] cpu kernels {res}")
    return res


def cpu_gol(R, threads, gpu):
    rows = cols = 1024
    iters = 100
    n = rows * cols
    x0 = R.fill_bernoulli_i32(n, SEED).reshape(rows, cols)
    bufs = [x0.copy(), np.empty_like(x0)]

    def run():
        bufs[0][...] = x0
        for i in range(iters):
            R.gol_i32(bufs[i & 1], out=bufs[(i + 1) & 1])
        return R.fold_sum_i32(bufs[iters & 1])

    t, reps = _cpu_time(run, reps=3, budget_s=10.0)
    live = run()
    res = {"ms": round(t * 1e3, 2), "us_per_step": round(t * 1e6 / iters, 1), "cells_per_s": round(n * iters / t, 1),
           "cores": threads, "kind": "port", "runs": reps, "live_cells_match_gpu": int(live) == gpu.get("live_cells"),
           "gpu_over_cpu": round(t * 1e3 / gpu["ms"], 1),
           "sample": "oracle Game of Life, the full 1024x1024 grid x 100 iterations + fold(sum)"}
    log(f"[bench] cpu gol {res}")
    return res


def cpu_heat(R, threads, gpu, extra):
    """The full 512^3 x 200 heat run with a residual every 10 on the oracle
    (timed), checked bit for bit against the GPU's final state and residuals;
    then the FP64 recurrence on the same inputs gives the error the FP32 run
    carries after 200 steps."""
    nx = ny = nz = 512
    iters, every = 200, 10
    n = nx * ny * nz
    x0 = R.fill_uniform_f32(n, SEED, R.SCALE_UNIT).reshape(nz, ny, nx)
    t0 = time.perf_counter()
    ref, ref_res = R.heat7_iterate_f32(x0, iters, every)
    t = time.perf_counter() - t0
    gbs = 8 * n * iters / t / 1e9
    res = {"ms": round(t * 1e3, 1), "GBs": round(gbs, 2), "cores": threads, "kind": "port",
           "gpu_over_cpu": round(gpu["GBs"] / gbs, 1),
           "sample": "oracle heat7, the full 512^3 cube x 200 iterations with the fused residual every 10"}
    got = extra.get("heat_final")
    if got is not None:
        res["bit_exact_vs_gpu"] = bool(np.array_equal(got.reshape(ref.shape), ref)) and \
            list(ref_res) == list(gpu["residuals"])
        del ref
        t1 = time.perf_counter()
        x64 = R.heat7_iterate_f64(x0, iters)
        d = np.abs(got.reshape(x64.shape).astype(np.float64) - x64)
        res["fp64_error"] = {"max_abs": float(d.max()),
                             "max_rel": float((d / np.maximum(np.abs(x64), 1e-30)).max()),
                             "mean_abs": float(d.mean()), "fp64_ms": round((time.perf_counter() - t1) * 1e3, 1),
                             "what": "GPU FP32 state after 200 steps vs an FP64 recurrence from the same inputs"}
        del x64, d
    log(f"[bench] cpu heat {res}")
    return res


def cpu_batch(R, threads, gpu, extra):
    """All 4096 grids of the batch, 10 oracle mean-blur steps each plus the
    FP64 oracle fold of each final state; the per-grid sums are compared with
    the GPU's."""
    from muesli4_amd.workloads import arena_offsets, batch_shapes
    rows, cols, radii = batch_shapes(BATCH_GRIDS, BATCH_SEED)
    off, total = arena_offsets(rows, cols)
    arena = R.fill_uniform_f32(total, BATCH_SEED)
    sums = np.empty(BATCH_GRIDS)
    t0 = time.perf_counter()
    for g in range(BATCH_GRIDS):
        o = int(off[g])
        x = arena[o:o + int(rows[g] * cols[g])].reshape(int(rows[g]), int(cols[g]))
        sums[g] = R.fold_sum_f64(R.mean_blur_iterate_f32(x, int(radii[g]), BATCH_ITERS))
    t = time.perf_counter() - t0
    gbs = gpu["algorithmic_bytes"] / t / 1e9
    res = {"ms": round(t * 1e3, 1), "GBs": round(gbs, 2), "cores": threads, "kind": "port",
           "gpu_over_cpu": round(t * 1e3 / gpu["persistent_ms"], 1),
           "sample": "oracle, every grid of the batch x 10 steps + FP64 fold, grids one after another"}
    gs = extra.get("batch_sums")
    if gs is not None:
        res["max_rel_sum_diff_vs_gpu"] = float(np.max(np.abs(gs - sums) / np.maximum(np.abs(sums), 1e-30)))
    log(f"[bench] cpu batch {res}")
    return res


def cpu_serve(R, threads, gpu, extra):
    """The saturated serving mix on the oracle, jobs in trace order until
    about 10 s have elapsed: CPU jobs/s on that sample beside the engine."""
    from muesli4_amd.workloads import JOB_BLUR2D, JOB_HEAT3D, serve_trace
    jobs, _ = serve_trace(SERVE_JOBS, SERVE_SEED, 0.0)
    done = 0
    t0 = time.perf_counter()
    for d in jobs:
        if d["kind"] == JOB_BLUR2D:
            x = R.fill_uniform_f32(d["d0"] * d["d1"], d["seed"]).reshape(d["d0"], d["d1"])
            R.fold_sum_f64(R.mean_blur_iterate_f32(x, d["r"], d["steps"], d.get("nv", 0.0)))
        elif d["kind"] == JOB_HEAT3D:
            x = R.fill_uniform_f32(d["d0"] * d["d1"] * d["d2"], d["seed"], R.SCALE_UNIT)
            R.fold_sum_f64(R.heat7_iterate_f32(x.reshape(d["d2"], d["d1"], d["d0"]), d["steps"])[0])
        else:
            R.fold_sum_u32(R.map_affine_u32(R.fill_u32(d["n"], d["seed"]), d["mul"], d["add"]))
        done += 1
        if time.perf_counter() - t0 > 10.0 and done >= 5:
            break
    t = time.perf_counter() - t0
    rate = done / t
    eng = gpu["saturated"]["engine"]["jobs_per_s"]
    res = {"jobs_per_s": round(rate, 2), "jobs": done, "s": round(t, 2), "cores": threads, "kind": "port",
           "gpu_engine_over_cpu": round(eng / rate, 1),
           "sample": f"oracle, the first {done} jobs of the saturated {SERVE_JOBS}-job trace in order, "
                     "one after another (inputs filled, steps, fold)"}
    log(f"[bench] cpu serve {res}")
    return res


if __name__ == "__main__":
    main()

"""Seeded synthetic workloads of the benchmark configurations.

Every random draw is a SplitMix64 hash of (seed, counter), so the same seed
gives the same shapes on any machine.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser of seed + (idx+1)*gamma (uint64, wrapping)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(idx, np.uint64) + np.uint64(1)) * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, idx: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of the hash."""
    return (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def batch_shapes(n: int, seed: int, r_max: int = 8, side_min: int = 32, side_max: int = 1024):
    """Grid shapes of the batched mean-blur configuration.

    Long side L = round(side_min * (side_max/side_min)^U), log-uniform;
    aspect a ~ U[1, 4]; short side max(side_min, round(L / a)); a fair coin
    picks which side is the row count; radius uniform in {1..r_max}.
    Returns int arrays rows, cols, radii.
    """
    i = np.arange(n, dtype=np.uint64) * np.uint64(4)
    u_len = uniform01(seed, i)
    u_asp = uniform01(seed, i + np.uint64(1))
    u_ori = uniform01(seed, i + np.uint64(2))
    u_rad = uniform01(seed, i + np.uint64(3))
    long_side = np.rint(side_min * (side_max / side_min) ** u_len).astype(np.int64)
    aspect = 1.0 + 3.0 * u_asp
    short_side = np.maximum(side_min, np.rint(long_side / aspect).astype(np.int64))
    tall = u_ori < 0.5
    rows = np.where(tall, long_side, short_side)
    cols = np.where(tall, short_side, long_side)
    radii = 1 + np.minimum(r_max - 1, np.floor(u_rad * r_max).astype(np.int64))
    return rows.astype(np.int64), cols.astype(np.int64), radii.astype(np.int64)


def arena_offsets(rows: np.ndarray, cols: np.ndarray, align: int = 64):
    """Element offsets of grids packed into one buffer, each start aligned to
    `align` elements; returns (offsets, total elements)."""
    sizes = rows.astype(np.int64) * cols.astype(np.int64)
    padded = (sizes + align - 1) // align * align
    offsets = np.zeros(len(sizes), np.int64)
    if len(sizes) > 1:
        offsets[1:] = np.cumsum(padded)[:-1]
    total = int(padded.sum()) if len(sizes) else 0
    return offsets, total


JOB_BLUR2D, JOB_HEAT3D, JOB_MAPFOLD_U32 = 1, 2, 3


def serve_trace(n: int, seed: int, rate_per_s: float, mix=(0.5, 0.25, 0.25), blur_side=(64, 2048),
                blur_r=(1, 10), blur_steps=(10, 500), heat_side=(32, 256), heat_steps=(10, 100),
                map_len=(1e4, 1e8)):
    """Job trace of the serving configuration: Poisson arrivals at
    `rate_per_s` (rate <= 0: every job arrives at t = 0) of
      2-D mean-blur jobs: rows and cols log-uniform in blur_side, radius
        uniform in blur_r, steps uniform in blur_steps;
      3-D heat jobs: a cube with side uniform in heat_side, steps uniform in
        heat_steps;
      u32 map(x*3+1)+fold jobs: length log-uniform in map_len;
    in proportions `mix`.  Returns (list of job dicts, arrival times)."""
    i = np.arange(n, dtype=np.uint64) * np.uint64(8)
    u = [uniform01(seed, i + np.uint64(k)) for k in range(8)]
    if rate_per_s > 0:
        gaps = -np.log1p(-u[0]) / rate_per_s
        arrivals = np.cumsum(gaps) - gaps[0]
    else:
        arrivals = np.zeros(n)
    jobs = []
    c1, c2 = mix[0], mix[0] + mix[1]
    for k in range(n):
        sel = u[1][k]
        job_seed = int(splitmix64(seed ^ 0x5EED, np.asarray([k], np.uint64))[0])
        if sel < c1:
            lo, hi = blur_side
            rows = int(round(lo * (hi / lo) ** u[2][k]))
            cols = int(round(lo * (hi / lo) ** u[3][k]))
            r = blur_r[0] + min(blur_r[1] - blur_r[0], int(u[4][k] * (blur_r[1] - blur_r[0] + 1)))
            steps = blur_steps[0] + min(blur_steps[1] - blur_steps[0],
                                        int(u[5][k] * (blur_steps[1] - blur_steps[0] + 1)))
            jobs.append(dict(kind=JOB_BLUR2D, d0=rows, d1=cols, r=r, steps=steps, seed=job_seed))
        elif sel < c2:
            side = heat_side[0] + min(heat_side[1] - heat_side[0], int(u[2][k] * (heat_side[1] - heat_side[0] + 1)))
            steps = heat_steps[0] + min(heat_steps[1] - heat_steps[0],
                                        int(u[5][k] * (heat_steps[1] - heat_steps[0] + 1)))
            jobs.append(dict(kind=JOB_HEAT3D, d0=side, d1=side, d2=side, steps=steps, seed=job_seed))
        else:
            lo, hi = map_len
            length = int(round(lo * (hi / lo) ** u[2][k]))
            jobs.append(dict(kind=JOB_MAPFOLD_U32, n=length, mul=3, add=1, seed=job_seed))
    return jobs, arrivals

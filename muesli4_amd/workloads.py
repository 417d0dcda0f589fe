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

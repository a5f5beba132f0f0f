"""Deterministic scenes for the load-time BVH order tests.

Inputs come from an integer hash (splitmix64) of (seed, index, field) and are
binary fractions, so they do not depend on numpy's random streams and a
golden fixture only needs (kind, n, seed).
"""
from __future__ import annotations

import numpy as np

KINDS = ("uniform", "clusters", "dups", "flat", "line")


def _splitmix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def scene(n: int, seed: int, kind: str):
    """(means, rotations (w, x, y, z), scales) as float32 arrays."""
    i = np.arange(n, dtype=np.uint64)
    base = i * np.uint64(16) + np.uint64(seed) * np.uint64(0x100000001B3)

    def h(k):
        return _splitmix(base + np.uint64(k))

    def unit(k, bits=20):  # multiples of 2^-bits in [0, 1)
        return (h(k) >> np.uint64(64 - bits)).astype(np.float64) / float(1 << bits)

    # about two Gaussians per unit cube, side a power of two
    side = float(2.0 ** max(0, round(np.log2(max(n, 1) / 2.0) / 3.0)))
    means = (np.stack([unit(0), unit(1), unit(2)], 1) - 0.5) * side
    if kind == "clusters":
        cid = (h(3) % np.uint64(64)).astype(np.int64)
        cen = (_splitmix(np.arange(64 * 3, dtype=np.uint64) + np.uint64(seed)) >> np.uint64(44)).astype(np.float64)
        cen = cen.reshape(64, 3) / float(1 << 20) * 1024.0 - 512.0
        means = cen[cid] + means / side * 4.0
    elif kind == "dups":
        means = means[(np.arange(n) // 7) * 7]
    elif kind == "flat":
        means[:, 2] = 3.25
    elif kind == "line":
        means[:, 1] = 0.0
        means[:, 2] = 0.0
    e = (h(7) % np.uint64(8)).astype(np.float64)
    m = (h(8) % np.uint64(8)).astype(np.float64)
    s = np.ldexp(1.0 + m / 8.0, -(e.astype(np.int64) + 1))
    scales = np.stack([s, s * 0.5, s * 0.25], 1)
    if kind == "line":
        scales[:] = 0.0
    q = np.stack([np.ones(n), *[((h(9 + c) % np.uint64(17)).astype(np.float64) - 8.0) / 16.0 for c in range(3)]], 1)
    return (np.ascontiguousarray(means, np.float32), np.ascontiguousarray(q, np.float32),
            np.ascontiguousarray(scales, np.float32))


def perm_from_leaves(n: int, first: np.ndarray, count: np.ndarray) -> np.ndarray:
    """The reference's reorder as a closed form of the leaf ranges: the leaf
    walk visits right children first, so the ranges come in descending order
    and range [f, f + c) receives old indices n - f - c .. n - f - 1."""
    perm = np.empty(n, np.uint32)
    for f, c in zip(first.astype(np.int64), count.astype(np.int64)):
        perm[f:f + c] = np.arange(n - f - c, n - f, dtype=np.uint32)
    return perm

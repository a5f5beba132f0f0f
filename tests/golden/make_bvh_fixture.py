"""Writes tests/golden/bvh_tinybvh_leaves.npz: the leaves the reference's own
tinybvh 1.6.7 (packages/tinybvh/tiny_bvh.h, built into oracle/_ref by
oracle/Makefile where the reference tree is present) builds over the scenes
of tests/bvh_scenes.py, as leaf sizes in position order.

The triangles handed to tinybvh are the oracle's restatement of
gaussian_bvh.cu:150-155 (3-sigma boxes) and gaussian_bvh.cuh:20-26.  The leaf
walk is the reference's (right child first, primIdx positions), so its
output is a sequence of ascending runs in descending position order; the run
lengths, reversed, are the leaf sizes.

Run: python tests/golden/make_bvh_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import bvh_scenes  # noqa: E402
import orc  # noqa: E402

CASES = [("uniform", 3000, 1), ("clusters", 60000, 2), ("dups", 5000, 3), ("flat", 20000, 4), ("line", 2000, 5),
         ("uniform", 50000, 6)]


def leaf_sizes(leaf_order: np.ndarray) -> np.ndarray:
    brk = np.flatnonzero(np.diff(leaf_order.astype(np.int64)) != 1) + 1
    runs = np.diff(np.concatenate([[0], brk, [len(leaf_order)]]))
    return runs[::-1].astype(np.uint32)


def main():
    ref = orc.tinybvh_ref()
    assert ref is not None, "oracle/_ref/libtinybvh_ref.so is not built (needs the reference tree)"
    out = {}
    for kind, n, seed in CASES:
        m, q, s = bvh_scenes.scene(n, seed, kind)
        v = orc.bvh_fragments(m, q, s)
        lo = np.empty(n, np.uint32)
        assert ref.ref_tinybvh_leaf_order(v.ctypes.data, n, lo.ctypes.data) == 0
        sizes = leaf_sizes(lo)
        assert sizes.sum() == n
        out[f"{kind}_{n}_{seed}"] = sizes.astype(np.uint16 if sizes.max() < 65536 else np.uint32)
        print(kind, n, seed, "leaves", len(sizes), "max", sizes.max())
    np.savez_compressed(os.path.join(HERE, "bvh_tinybvh_leaves.npz"), **out)


if __name__ == "__main__":
    main()

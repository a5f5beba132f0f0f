"""Load-time BVH order on the GPU (gps_bvh_order) against the oracle
(orc_bvh.cpp, pinned to the reference's tinybvh in test_bvh_order.py) and
against the tinybvh golden leaves.  Integer output: bit-exact.
"""
import os

import numpy as np
import pytest

import bvh_scenes as B
import orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bvh_tinybvh_leaves.npz")


def _gpu_order(G, m, q, s):
    a = G.SceneArrays(m, q, s, np.zeros(len(m), np.float32), np.zeros((len(m), 3), np.float32), 0)
    return a.bvh_order()


@pytest.mark.parametrize("kind", B.KINDS)
@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 64, 1000, 49999, 50000, 300000])
def test_bvh_order_matches_oracle(gpu, kind, n):
    m, q, s = B.scene(n, 1000 + n, kind)
    print(f"bvh_order kind={kind} n={n}", flush=True)
    got = _gpu_order(gpu, m, q, s)
    want = orc.bvh_order(m, q, s)
    assert np.array_equal(got, want), (kind, n, int((got != want).sum()))


def test_bvh_order_matches_tinybvh_golden(gpu):
    g = np.load(GOLDEN)
    for key in g.files:
        kind, n, seed = key.rsplit("_", 2)
        n, seed = int(n), int(seed)
        m, q, s = B.scene(n, seed, kind)
        count = g[key].astype(np.int64)
        first = np.concatenate([[0], np.cumsum(count)[:-1]])
        print(f"bvh_order golden {key}", flush=True)
        assert np.array_equal(_gpu_order(gpu, m, q, s), B.perm_from_leaves(n, first, count)), key


def test_bvh_order_large_clusters(gpu):
    # 4M Gaussians: many levels of the level-synchronous build
    n = 4_000_000
    m, q, s = B.scene(n, 77, "clusters")
    print(f"bvh_order clusters n={n}", flush=True)
    got = _gpu_order(gpu, m, q, s)
    want = orc.bvh_order(m, q, s)
    assert np.array_equal(got, want), int((got != want).sum())


def test_bvh_order_synthetic_scene(gpu):
    # the renderer's own synthetic Gaussians (Gaussian-distributed positions,
    # log-normal scales, random rotations)
    p = gpu.synth_params(0x5EED0042, 200_000, kind="clusters", clusters=16, half_extent=20.0)
    a = gpu.synth_scene(p)
    print("bvh_order synth n=200000", flush=True)
    assert np.array_equal(a.bvh_order(), orc.bvh_order(a.means, a.rotations, a.scales))

"""Load-time BVH order (GaussianBVH::build_bvh_and_reorder,
src/core/rendering/occlusion/gaussian_bvh.cu:146-169): CPU checks of the
oracle.

Pinning of orc_bvh.cpp:
  * against the reference's own tinybvh 1.6.7 compiled from
    packages/tinybvh/tiny_bvh.h into oracle/_ref (skipped where that library
    was not built, i.e. without the reference tree);
  * against tests/golden/bvh_tinybvh_leaves.npz, leaves tinybvh built over
    the same scenes (written by tests/golden/make_bvh_fixture.py);
  * closed forms: one Gaussian or coincident boxes form one leaf (identity),
    two far-apart Gaussians split (reversal), and the literal leaf walk +
    cycle reorder equals the per-leaf mirror of the leaf ranges.
"""
import os

import numpy as np
import pytest

import bvh_scenes as B
import orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bvh_tinybvh_leaves.npz")

SIZES = [1, 2, 3, 33, 1000, 49999, 50000, 120000]


def _is_perm(p):
    return np.array_equal(np.sort(p), np.arange(len(p), dtype=p.dtype))


@pytest.mark.parametrize("kind", B.KINDS)
def test_oracle_matches_reference_tinybvh(kind):
    ref = orc.tinybvh_ref()
    if ref is None:
        pytest.skip("oracle/_ref/libtinybvh_ref.so not built (reference tree absent at build time)")
    for i, n in enumerate(SIZES):
        m, q, s = B.scene(n, 100 + i, kind)
        want = orc.tinybvh_ref_order(ref, orc.bvh_fragments(m, q, s))
        got = orc.bvh_order(m, q, s)
        assert np.array_equal(got, want), (kind, n)


def test_oracle_matches_golden_leaves():
    g = np.load(GOLDEN)
    assert len(g.files) >= 6
    for key in g.files:
        kind, n, seed = key.rsplit("_", 2)
        n, seed = int(n), int(seed)
        m, q, s = B.scene(n, seed, kind)
        first, count = orc.bvh_leaves(orc.bvh_fragments(m, q, s))
        assert np.array_equal(count, g[key].astype(np.uint32)), key
        assert np.array_equal(first, np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.uint32)), key
        assert np.array_equal(orc.bvh_order(m, q, s), B.perm_from_leaves(n, first, count)), key


def test_golden_has_multi_leaf_structure():
    # the fixture must pin more than "every Gaussian alone" (a plain reversal)
    g = np.load(GOLDEN)
    assert sum(int((g[k] > 1).sum()) for k in g.files) > 1000
    assert max(int(g[k].max()) for k in g.files) >= 4


def test_single_gaussian_is_identity():
    m, q, s = B.scene(1, 7, "uniform")
    assert orc.bvh_order(m, q, s).tolist() == [0]


def test_coincident_boxes_form_one_leaf():
    # identical boxes: every split candidate costs more than not splitting
    n = 40
    m = np.tile(np.float32([[1.0, 2.0, 3.0]]), (n, 1))
    q = np.tile(np.float32([[1.0, 0.0, 0.0, 0.0]]), (n, 1))
    s = np.full((n, 3), 0.5, np.float32)
    first, count = orc.bvh_leaves(orc.bvh_fragments(m, q, s))
    assert count.tolist() == [n]
    assert np.array_equal(orc.bvh_order(m, q, s), np.arange(n, dtype=np.uint32))


def test_two_far_gaussians_reverse():
    # two small boxes far apart: SAH cost 1 + (a0 + a1) / A < 2, two leaves,
    # and the right-first walk puts position 1 first
    m = np.float32([[0.0, 0.0, 0.0], [100.0, 50.0, 25.0]])
    q = np.float32([[1.0, 0.0, 0.0, 0.0]] * 2)
    s = np.full((2, 3), 0.1, np.float32)
    assert orc.bvh_order(m, q, s).tolist() == [1, 0]


@pytest.mark.parametrize("kind", ["uniform", "clusters", "dups", "flat"])
def test_leaf_walk_equals_leaf_mirror(kind):
    # the literal collect_leaf_order + reorder() of the oracle against the
    # closed form the GPU build writes directly
    for n, seed in [(5, 1), (777, 2), (60000, 3)]:
        m, q, s = B.scene(n, seed, kind)
        first, count = orc.bvh_leaves(orc.bvh_fragments(m, q, s))
        assert int(count.sum()) == n
        got = orc.bvh_order(m, q, s)
        assert _is_perm(got)
        assert np.array_equal(got, B.perm_from_leaves(n, first, count)), (kind, n)


def test_leaves_tile_positions_across_thread_threshold():
    # 49999 and 50000 Gaussians sit either side of tinybvh's threading
    # threshold; both builds' leaves tile [0, n) in order
    for n in (49999, 50000):
        m, q, s = B.scene(n, 9, "clusters")
        first, count = orc.bvh_leaves(orc.bvh_fragments(m, q, s))
        assert int(count.sum()) == n and first[0] == 0
        assert np.array_equal(first[1:], np.cumsum(count)[:-1])

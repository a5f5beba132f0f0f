"""The group occlusion test on the GPU (gps_pass_set_group_culling).

The preprocess skips a group of 64 Gaussians when the conservative group
test says every member fails the per-Gaussian Hi-Z test, so a pass with the
test on must give the same bits as a pass with it off:

* along the fly-through of the 100M workload's construction (4 translated
  copies, Morton order, blocks of 1024, previous-frame Hi-Z, TAA) at 4M
  Gaussians and 1920x1080: framebuffer words, colour, depth, TAA output,
  weights, valid mask and pass counters equal frame by frame, while the test
  skips a large share of the groups;
* at ss 3 (Hi-Z pixels by division, not shift), with the small-Gaussian cull
  and without block culling: the same equality;
* with min-reduced Hi-Z levels the test is off (no group skipped);
* the counting form of the preprocess runs every member of a skipped group
  through the full per-Gaussian steps and finds none that survives.

The soundness argument itself is checked on the CPU (test_group_bound.py);
parity with the oracle runs with the test on (its default) in every other
occlusion test.
"""
import numpy as np
import pytest

from helpers import assert_same, flush

pytestmark = pytest.mark.gpu

POSITIONS = (0, 1, 2, 3, 4, 5, 6, 150, 151, 299)


def _loops(w, scene):
    from gps_amd import workloads as W
    on, off = W.FrameLoop(w, scene), W.FrameLoop(w, scene)
    off.pass_.set_group_culling(False)
    for lp in (on, off):  # skipped groups are counted with the diagnostic counters on
        lp.pass_.set_counters(True)
    return on, off


def _compare(gpu, w, on, off, frames, tag):
    culled = tested_groups = 0
    for k, f in enumerate(frames):
        flush(f"{tag}: frame {k}")
        on.step(f)
        off.step(f)
        a, b = (k + 1) % 2, k % 2
        so, sf = on.pass_.stats(on.stream), off.pass_.stats(off.stream)
        for key in ("valid_gaussians", "num_points", "total_weight", "culled_blocks", "passes", "samples"):
            assert so[key] == sf[key], (tag, k, key, so[key], sf[key])
        assert sum(sf["culled_groups"]) == 0
        for dbg in (gpu.DBG_FRAMEBUFFER, gpu.DBG_WEIGHTS, gpu.DBG_MASK_CURRENT, gpu.DBG_MASK_PREVIOUS):
            assert_same(f"{tag} dbg {dbg} frame {k}", on.pass_.debug(dbg, on.stream), off.pass_.debug(dbg, off.stream))
        assert_same(f"{tag} color {k}", on.color[a].download(), off.color[a].download())
        assert_same(f"{tag} depth {k}", on.depth[b].download(), off.depth[b].download())
        if w.taa:
            assert_same(f"{tag} taa {k}", on.hist[a].download(), off.hist[a].download())
        culled += sum(so["culled_groups"][:so["passes"]])
        tested_groups += so["passes"] * ((w.n + 63) // 64)
    flush(f"{tag}: {culled} of {tested_groups} group tests skipped their group")
    return culled


def test_group_culling_changes_nothing_flythrough(gpu):
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 1_000_000)
    perm = W.order(w)
    scene = W.device_scene(w, perm)
    on, off = _loops(w, scene)
    frames = [W.frame(w, pos, k + 1) for k, pos in enumerate(POSITIONS)]
    culled = _compare(gpu, w, on, off, frames, "c3 4M")
    assert culled > 2_000


def test_group_culling_changes_nothing_ss3_small_noblocks(gpu):
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 100_000, width=640, height=360, ss=3, hierarchical=False, taa=False,
                 half_extent=20.0)
    perm = W.order(w)
    scene = W.device_scene(w, perm)
    on, off = _loops(w, scene)
    frames = []
    for k, pos in enumerate((0, 1, 2, 60, 61, 200)):
        f = W.frame(w, pos, k + 1)
        f.cull_small_gaussians = 1
        frames.append(f)
    culled = _compare(gpu, w, on, off, frames, "ss3")
    assert culled > 0


def test_group_culling_off_with_min_reduced_levels(gpu):
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 100_000, width=640, height=360)
    perm = W.order(w)
    scene = W.device_scene(w, perm)
    on, off = _loops(w, scene)
    frames = []
    for k, pos in enumerate((0, 1, 2, 150)):
        f = W.frame(w, pos, k + 1)
        f.occlusion_culling_min_reduce_count = 1
        frames.append(f)
    assert _compare(gpu, w, on, off, frames, "min-reduce") == 0


def test_group_culling_counting_form_finds_no_survivor(gpu):
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 500_000)
    perm = W.order(w)
    scene = W.device_scene(w, perm)
    on, off = _loops(w, scene)
    on.pass_.set_counters(False, count_rejects=True)
    off.pass_.set_counters(False, count_rejects=True)
    culled = 0
    for k, pos in enumerate((0, 1, 2, 3, 150, 151)):
        f = W.frame(w, pos, k + 1)
        flush(f"counting form: frame {k}")
        on.step(f)
        off.step(f)
        st = on.pass_.stats(on.stream)
        assert on.pass_.group_violations(on.stream) == [0, 0], k
        # the counting form tests every member: the histogram is the test-off one
        assert on.pass_.rejects(on.stream) == off.pass_.rejects(off.stream), k
        assert_same(f"counting fb {k}", on.pass_.debug(gpu.DBG_FRAMEBUFFER, on.stream),
                    off.pass_.debug(gpu.DBG_FRAMEBUFFER, off.stream))
        culled += sum(st["culled_groups"][:st["passes"]])
    assert culled > 0

"""The ordered point-parallel splat against the chunk-order splat and the oracle.

A dense pass (points of at least 8 samples, more than 32 samples per
framebuffer word; always, when forced) is splatted nearest first in four
phases; before each later phase, and before the first phase of
a second render pass, the maximum depth of every 16x16-pixel framebuffer tile
is taken, and a Gaussian whose bounds lie on tiles that are all nearer than
it is skipped.  Words only decrease and every sample of a Gaussian lands
inside its bounds, so the framebuffer must equal the chunk-order splat's and
the oracle's bit for bit: one and four sample layers, supersampling, two-pass
occlusion culling (the second pass tests against the first pass's tiles),
Gaussians too large for the test, and the automatic choice.
"""
import numpy as np
import pytest

import helpers as H
from helpers import assert_same, flush

pytestmark = pytest.mark.gpu


def slabs(seed: int, n_each: int, count: int, log_scale_mu: float = -2.0) -> H.G.SceneArrays:
    """`count` slabs of Gaussians 6 units apart along z: the nearer slabs hide
    most of the farther ones."""
    parts = [H.scene(seed + i, n_each, dict(H.C1, half_extent=5.0, log_scale_mu=log_scale_mu),
                     offset=(0.0, 0.0, 6.0 * i)) for i in range(count)]
    cat = lambda k: np.concatenate([getattr(p, k) for p in parts])
    arrays = H.G.SceneArrays(cat("means"), cat("rotations"), cat("scales"), cat("opacities"), cat("shs"), 0)
    return arrays.morton_sorted()


CASES = {
    "spp1_ss1": dict(w=96, h=64, spp=1, ss=1, occl=False, mu=-2.0),
    "spp4_ss2": dict(w=48, h=32, spp=4, ss=2, occl=False, mu=-2.0),
    "two_pass_occlusion": dict(w=96, h=64, spp=1, ss=2, occl=True, mu=-2.0),
    "big_gaussians": dict(w=96, h=64, spp=1, ss=1, occl=False, mu=0.3),
}


def render_modes(arrays, c, frames, modes):
    st = H.G.RenderSettings(c["w"], c["h"], c["spp"], c["ss"], gaussians_per_block=64)
    out = {}
    for mode in modes:
        r = H.GpuRunner(arrays, st)
        r.p.set_splat_mode(H.G.Pass.SPLAT_POINT_PARALLEL)
        r.p.set_ordered_splat(mode)
        r.p.set_counters(False, True)
        res = []
        for k, f in enumerate(frames):
            flush(f"ordered: mode {mode} frame {k}")
            rg = r.render(f, [H.G.DBG_FRAMEBUFFER])
            rg.stats["skipped"] = r.p.ordered_skipped(r.stream)
            res.append(rg)
        out[mode] = res
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_ordered_splat_identical(gpu, case):
    c = CASES[case]
    arrays = slabs(0x5EED0901 + 16 * len(case), 12_000, 4, c["mu"])
    n_frames = 3 if c["occl"] else 1
    frames = [gpu.frame_params((1.5 * i, 2.0, -22.0), (0.0, 0.0, 6.0), (0, 1, 0), 45.0, c["w"], c["h"], ss=c["ss"],
                               frame=i + 1, spp=c["spp"], occlusion=c["occl"], hierarchical=c["occl"])
              for i in range(n_frames)]
    G = gpu.Pass
    runs = render_modes(arrays, c, frames, (G.ORDERED_OFF, G.ORDERED_ALWAYS, G.ORDERED_AUTO))
    off, always, auto = runs[G.ORDERED_OFF], runs[G.ORDERED_ALWAYS], runs[G.ORDERED_AUTO]
    for k in range(n_frames):
        for name, run in (("always", always), ("auto", auto)):
            assert_same(f"{case} frame {k} {name} fb", run[k].debug[gpu.DBG_FRAMEBUFFER],
                        off[k].debug[gpu.DBG_FRAMEBUFFER])
            assert_same(f"{case} frame {k} {name} colour", run[k].color, off[k].color)
            assert_same(f"{case} frame {k} {name} depth", run[k].depth, off[k].depth)
        assert off[k].stats["skipped"] == [0, 0]
    skipped = [sum(r.stats["skipped"]) for r in always]
    points = [sum(r.stats["num_points"]) for r in always]
    flush(f"{case}: points {points} skipped {skipped}")
    if case == "big_gaussians":
        assert all(s < p for s, p in zip(skipped, points))
    else:
        assert sum(skipped) > 0.05 * sum(points), (skipped, points)
    if c["occl"]:
        # the second pass tests against the first pass's tiles before its first phase
        flush(f"second pass: points {[r.stats['num_points'][1] for r in always]} "
              f"skipped {[r.stats['skipped'][1] for r in always]}")


@pytest.mark.parametrize("case", ["spp1_ss1", "two_pass_occlusion"])
def test_ordered_splat_equals_oracle(gpu, case):
    """The ordered splat (forced) against the CPU oracle, frame by frame."""
    c = CASES[case]
    arrays = slabs(0x5EED0941 + len(case), 8_000, 4, c["mu"])
    n_frames = 3 if c["occl"] else 1
    frames = [gpu.frame_params((1.5 * i, 2.0, -22.0), (0.0, 0.0, 6.0), (0, 1, 0), 45.0, c["w"], c["h"], ss=c["ss"],
                               frame=i + 1, spp=c["spp"], occlusion=c["occl"], hierarchical=c["occl"])
              for i in range(n_frames)]
    st = gpu.RenderSettings(c["w"], c["h"], c["spp"], c["ss"], gaussians_per_block=64)
    g = H.GpuRunner(arrays, st)
    g.p.set_splat_mode(gpu.Pass.SPLAT_POINT_PARALLEL)
    g.p.set_ordered_splat(gpu.Pass.ORDERED_ALWAYS)
    g.p.set_counters(False, True)
    o = H.OracleRunner(arrays, st)
    total_skipped = 0
    for k, f in enumerate(frames):
        flush(f"ordered vs oracle {case}: frame {k}")
        rg = g.render(f, [gpu.DBG_FRAMEBUFFER])
        total_skipped += sum(g.p.ordered_skipped(g.stream))
        ro = o.render(f, [gpu.DBG_FRAMEBUFFER])
        assert_same(f"{case} frame {k} fb", rg.debug[gpu.DBG_FRAMEBUFFER], ro.debug[gpu.DBG_FRAMEBUFFER])
        assert_same(f"{case} frame {k} colour", rg.color, ro.color)
        assert_same(f"{case} frame {k} depth", rg.depth, ro.depth)
        for key in ("valid_gaussians", "num_points", "samples"):
            assert rg.stats[key] == ro.stats[key], key
    assert total_skipped > 0


def test_switches_reject_unknown_values(gpu):
    """An unknown ordered-splat mode and a preprocess split beyond the batch
    limit (256 views) are refused with an argument error."""
    arrays = slabs(0x5EED0971, 2_000, 1)
    scene = gpu.Scene(arrays, gpu.FORMAT_FP32)
    p = gpu.Pass(gpu.RenderSettings(64, 48, 1, 1), scene)
    for mode in (gpu.Pass.ORDERED_AUTO, gpu.Pass.ORDERED_OFF, gpu.Pass.ORDERED_ALWAYS):
        p.set_ordered_splat(mode)
    with pytest.raises(gpu.GpsError):
        p.set_ordered_splat(3)
    batch = gpu.Batch(scene, 2, 64, 48, max_spp=1, max_supersampling=1)
    batch.set_tuning(256, 4, False, 16)
    with pytest.raises(gpu.GpsError):
        batch.set_tuning(256, 4, False, 257)
    with pytest.raises(gpu.GpsError):
        batch.set_tuning(256, 3, False, 0)  # not a splat variant

"""Shared drivers for parity tests: run the same frame on the MI355X path and
on the CPU oracle and return every intermediate both expose."""
from __future__ import annotations

import sys
from dataclasses import dataclass, field

import numpy as np

import gps_amd as G

# Configurations of the benchmark (synthetic, seeded), scaled by `n`.
C1 = dict(kind="uniform", sh_degree=0, half_extent=10.0, log_scale_mu=-4.0, log_scale_sigma=0.5,
          opacity_sigma=1.0, dc_sigma=0.3)
C2 = dict(kind="clustered", sh_degree=3, half_extent=100.0, cluster_sigma=2.0, log_scale_mu=-3.0,
          log_scale_sigma=0.7, opacity_sigma=1.5, dc_sigma=0.3, rest_sigma=0.05)


def scene(seed: int, n: int, cfg: dict, clusters: int = 2000, offset=(0.0, 0.0, 0.0)) -> G.SceneArrays:
    kw = dict(cfg)
    if kw.get("kind") == "clustered":
        kw["clusters"] = clusters
    return G.synth_scene(G.synth_params(seed, n, offset=offset, **kw))


INTERMEDIATES = [G.DBG_WEIGHTS, G.DBG_MASK_CURRENT, G.DBG_VALID_IDS, G.DBG_VALID_OFFSETS, G.DBG_INDICES,
                 G.DBG_FRAMEBUFFER]


@dataclass
class Result:
    color: np.ndarray
    depth: np.ndarray
    stats: dict
    debug: dict = field(default_factory=dict)


def flush(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


class GpuRunner:
    def __init__(self, arrays: G.SceneArrays, settings: G.RenderSettings, fmt: int = G.FORMAT_FP32):
        self.scene = G.Scene(arrays, fmt)
        self.p = G.Pass(settings, self.scene)
        self.w, self.h = settings.width, settings.height
        self.color = G.DeviceArray((self.h, self.w, 3), np.float32, zero=True)
        self.depth = G.DeviceArray((self.h, self.w), np.float32, zero=True)
        self.stream = G.Stream()

    def render(self, frame: G.FrameParams, which=()) -> Result:
        self.p.render(frame, self.color, self.depth, self.stream)
        self.stream.synchronize()
        r = Result(self.color.download(), self.depth.download(), self.p.stats(self.stream))
        for k in which:
            r.debug[k] = self.p.debug(k, self.stream)
        return r


class OracleRunner:
    def __init__(self, arrays: G.SceneArrays, settings: G.RenderSettings, fmt: int = G.FORMAT_FP32,
                 threads: int = 8):
        import orc
        self.p = orc.OraclePass(settings, arrays, fmt, threads)

    def render(self, frame: G.FrameParams, which=()) -> Result:
        c, d = self.p.render(frame)
        r = Result(c, d, self.p.stats())
        # biased splat: framebuffer words are order-independent (so the GPU
        # result must equal this one bit for bit) when both counts are 0
        r.stats["biased_events"] = self.p.biased_events()
        for k in which:
            r.debug[k] = self.p.debug(k)
        return r


def records_valid(rec_u32: np.ndarray, ids: np.ndarray) -> np.ndarray:
    return rec_u32.reshape(-1, 16)[ids]


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    if mse == 0.0:
        return float("inf")
    return 10.0 * np.log10(1.0 / mse)


def assert_same(name: str, a: np.ndarray, b: np.ndarray) -> None:
    assert a.shape == b.shape, f"{name}: shape {a.shape} vs {b.shape}"
    if a.dtype.kind == "f":
        eq = (a.view(np.uint32 if a.dtype == np.float32 else np.uint64) ==
              b.view(np.uint32 if b.dtype == np.float32 else np.uint64))
    else:
        eq = a == b
    if not eq.all():
        bad = np.flatnonzero(~eq.reshape(-1))
        raise AssertionError(f"{name}: {bad.size} of {eq.size} differ; first at {bad[:8]}: "
                             f"{a.reshape(-1)[bad[:4]]} vs {b.reshape(-1)[bad[:4]]}")


def carried_state(p):
    """(Hi-Z chain, valid mask the next first pass leaves in place) of a GPU
    or oracle pass after a frame: the state a frame carries into the next.
    The ping-pong flips when a pass yields points, so the untouched mask is
    the last pass's own after a flip and the one before it otherwise."""
    st = p.stats()
    passes = st["passes"]
    flipped = passes > 0 and st["num_points"][passes - 1] > 0
    other = p.debug(G.DBG_MASK_CURRENT) if flipped else p.debug(G.DBG_MASK_PREVIOUS)
    return p.debug(G.DBG_HIZ), other

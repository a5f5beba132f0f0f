"""Synthetic workloads of the splatting hot path: seeded scene generators,
camera paths and the per-frame loop (render pass + resolve) of the reference's
render server (src/core/rendering/render_server.cu:86-160).

    c1  1M uniform-box SH0 Gaussians, 1280x720, ss1, 1 spp, fixed camera
    c2  10M clustered SH3 Gaussians (2000 clusters), 1920x1080, ss2, 4 spp,
        120-frame orbit, fp32 storage
    c3  100M = 4 copies of a 25M clustered SH3 scene (5000 clusters)
        translated by (250 i, 0, 0), fp16/fp8 storage, Morton load order,
        blocks of 1024, previous-frame Hi-Z + block culling, TAA over a
        300-frame fly-through (-300,0,0) -> (1050,0,0) looking +X,
        1920x1080, ss2, 1 spp

Every workload can be scaled (`scaled(w, n)`) so the CPU oracle can render the
same construction at test sizes.  Large scenes are generated and uploaded in
chunks (the Morton order is computed from a positions-only pass first, then
the Gaussians are generated directly in sorted order), so host memory stays
at a few GB for the 100M scene.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field, replace
from typing import Callable, Optional

import numpy as np

from . import (FORMAT_FP32, FORMAT_HALF_FP8, FORMAT_QUANTIZED, DeviceArray, Pass, RenderSettings, Scene, SceneArrays,
               Stream, merge_quantized_bounds, quantized_bounds,
               _lib, _ptr, check, frame_params, mat4_mul, resolve_progressive, resolve_taa, synth_params,
               taa_reprojection)
import ctypes as C


@dataclass(frozen=True)
class Workload:
    name: str
    desc: str
    n_copy: int                 # Gaussians per copy
    copies: int = 1
    copy_stride: float = 0.0    # x translation between copies
    kind: str = "uniform"
    sh_degree: int = 0
    clusters: int = 1
    half_extent: float = 10.0
    cluster_sigma: float = 2.0
    log_scale_mu: float = -4.0
    log_scale_sigma: float = 0.5
    opacity_sigma: float = 1.0
    dc_sigma: float = 0.3
    rest_sigma: float = 0.05
    seed: int = 1
    width: int = 1280
    height: int = 720
    ss: int = 1
    spp: int = 1
    fmt: int = FORMAT_FP32
    camera: str = "fixed"       # fixed | orbit | flythrough
    path_frames: int = 1
    eye: tuple = (0.0, 0.0, -30.0)
    radius: float = 350.0
    elevation: float = 20.0
    step_deg: float = 3.0
    fly_from: tuple = (-300.0, 0.0, 0.0)
    fly_to: tuple = (1050.0, 0.0, 0.0)
    occlusion: bool = False
    hierarchical: bool = False
    gaussians_per_block: int = 0
    taa: bool = False
    morton: bool = True

    @property
    def n(self) -> int:
        return self.n_copy * self.copies

    def settings(self) -> RenderSettings:
        return RenderSettings(self.width, self.height, self.spp, self.ss,
                              fully_disable_hierarchical_culling=not self.hierarchical,
                              gaussians_per_block=self.gaussians_per_block)


WORKLOADS = {
    "c1": Workload("c1", "1M uniform-box SH0 Gaussians, 1280x720, ss1 (K=1), 1 spp, fixed camera, no culling",
                   n_copy=1_000_000, kind="uniform", sh_degree=0, half_extent=10.0, log_scale_mu=-4.0,
                   log_scale_sigma=0.5, opacity_sigma=1.0, dc_sigma=0.3, rest_sigma=0.0, seed=0x5EED0001,
                   width=1280, height=720, ss=1, spp=1, camera="fixed", eye=(0.0, 0.0, -30.0)),
    "c2": Workload("c2", "10M clustered SH3 Gaussians (2000 clusters), 1920x1080, ss2 (K=4), 4 spp, 120-frame "
                         "orbit r=350 elev 20deg +3deg/frame, no culling, fp32 storage",
                   n_copy=10_000_000, kind="clustered", sh_degree=3, clusters=2000, half_extent=100.0,
                   cluster_sigma=2.0, log_scale_mu=-3.0, log_scale_sigma=0.7, opacity_sigma=1.5, dc_sigma=0.3,
                   rest_sigma=0.05, seed=0x5EED0002, width=1920, height=1080, ss=2, spp=4, camera="orbit",
                   path_frames=120, radius=350.0, elevation=20.0, step_deg=3.0),
    "c3": Workload("c3", "100M Gaussians = 4 copies of a 25M clustered SH3 scene (5000 clusters) at x+250i, "
                         "fp32 positions / fp16 rotation, scale, opacity / fp8-E4M3 SH with per-Gaussian scale, "
                         "Morton order, blocks of 1024, previous-frame Hi-Z + block culling, TAA, 300-frame "
                         "fly-through (-300,0,0)->(1050,0,0) looking +X, 1920x1080, ss2 (K=4), 1 spp",
                   n_copy=25_000_000, copies=4, copy_stride=250.0, kind="clustered", sh_degree=3, clusters=5000,
                   half_extent=100.0, cluster_sigma=2.0, log_scale_mu=-3.0, log_scale_sigma=0.7,
                   opacity_sigma=1.5, dc_sigma=0.3, rest_sigma=0.05, seed=0x5EED0003, width=1920, height=1080,
                   ss=2, spp=1, fmt=FORMAT_HALF_FP8, camera="flythrough", path_frames=300, occlusion=True,
                   hierarchical=True, gaussians_per_block=1024, taa=True),
}


def scaled(w: Workload, n_copy: int, **kw) -> Workload:
    """The same construction with `n_copy` Gaussians per copy (clusters scale
    along so the cluster density stays the same)."""
    clusters = max(1, round(w.clusters * n_copy / w.n_copy)) if w.kind == "clustered" else w.clusters
    return replace(w, n_copy=n_copy, clusters=clusters, **kw)


def params(w: Workload):
    return synth_params(w.seed, w.n_copy, kind=w.kind, sh_degree=w.sh_degree, clusters=w.clusters,
                        half_extent=w.half_extent, cluster_sigma=w.cluster_sigma, log_scale_mu=w.log_scale_mu,
                        log_scale_sigma=w.log_scale_sigma, opacity_sigma=w.opacity_sigma, dc_sigma=w.dc_sigma,
                        rest_sigma=w.rest_sigma)


# ----------------------------------------------------------------- scenes
def _generate(w: Workload, idx: np.ndarray, means_only: bool = False):
    """Gaussians `idx` (indices into the concatenation of the copies).  Copy i
    is the base scene with float32(copy_stride * i) added to x."""
    idx = np.ascontiguousarray(idx, np.uint64)
    n = idx.size
    p = params(w)
    g = idx % np.uint64(w.n_copy)
    nc = (w.sh_degree + 1) ** 2
    means = np.empty((n, 3), np.float32)
    if means_only:
        arrs = (None, None, None, None)
    else:
        arrs = (np.empty((n, 4), np.float32), np.empty((n, 3), np.float32), np.empty((n,), np.float32),
                np.empty((n, 3 * nc), np.float32))
    check(_lib.gps_synth_generate_indices(C.byref(p), _ptr(g), n, _ptr(means), *[_ptr(a) for a in arrs]),
          "synth")
    if w.copies > 1:
        copy = (idx // np.uint64(w.n_copy)).astype(np.float32)
        means[:, 0] += copy * np.float32(w.copy_stride)
    if means_only:
        return means
    return SceneArrays(means, *arrs, w.sh_degree)


def order(w: Workload, chunk: int = 1 << 23) -> np.ndarray:
    """Load order of the scene: perm[new] = old (Morton if enabled, the
    reference's --sort-morton-order over the whole scene)."""
    if not w.morton:
        return np.arange(w.n, dtype=np.uint64)
    base = np.empty((w.n_copy, 3), np.float32)
    for b in range(0, w.n_copy, chunk):
        base[b:b + chunk] = _generate(w, np.arange(b, min(w.n_copy, b + chunk), dtype=np.uint64), True)
    if w.copies > 1:
        means = np.empty((w.n, 3), np.float32)
        for i in range(w.copies):
            m = means[i * w.n_copy:(i + 1) * w.n_copy]
            m[:] = base
            m[:, 0] += np.float32(i) * np.float32(w.copy_stride)
        del base
    else:
        means = base
    perm = np.empty(w.n, np.uint32)
    check(_lib.gps_morton_order(_ptr(means), w.n, _ptr(perm)), "morton order")
    return perm.astype(np.uint64)


def host_scene(w: Workload, perm: Optional[np.ndarray] = None) -> SceneArrays:
    """The whole scene on the host in load order (test sizes)."""
    if perm is None:
        perm = order(w)
    return _generate(w, perm)


def host_sample(w: Workload, count: int, perm: Optional[np.ndarray] = None) -> SceneArrays:
    """Every (n/count)-th Gaussian of the load order: a spatially uniform
    sample of the scene that keeps its load order."""
    if perm is None:
        perm = order(w)
    step = max(1, w.n // count)
    return _generate(w, perm[::step][:count])


def device_scene(w: Workload, perm: Optional[np.ndarray] = None, chunk: int = 1 << 22,
                 log: Optional[Callable[[str], None]] = None) -> Scene:
    """Generate the scene in load order and stream it into HBM in chunks."""
    if perm is None:
        t0 = time.time()
        perm = order(w)
        if log:
            log(f"load order ({'morton' if w.morton else 'generation'}) of {w.n} Gaussians in "
                f"{time.time() - t0:.1f}s")
    t0 = time.time()
    scene = Scene.empty(w.n, w.sh_degree, w.fmt)
    if w.fmt == FORMAT_QUANTIZED:
        # The scene box over the whole scene comes first; per-chunk boxes merge
        # exactly (min / max, and the 0.001 widening commutes with them).
        box = None
        for b in range(0, w.n, chunk):
            cb = quantized_bounds(_generate(w, perm[b:b + chunk]))
            box = cb if box is None else merge_quantized_bounds(box, cb)
        scene.set_quantized_bounds(box)
        if log:
            log(f"scene box of {w.n} Gaussians in {time.time() - t0:.1f}s")
    for b in range(0, w.n, chunk):
        arr = _generate(w, perm[b:b + chunk])
        scene.upload_range(b, arr)
    if log:
        log(f"generated + uploaded {w.n} Gaussians ({scene.device_bytes / 1e9:.2f} GB resident) in "
            f"{time.time() - t0:.1f}s")
    return scene


# ----------------------------------------------------------------- cameras
def eye_at(w: Workload, i: int):
    """Camera of path position i: (eye, center)."""
    if w.camera == "orbit":
        az = math.radians(w.step_deg * (i % w.path_frames))
        el = math.radians(w.elevation)
        r = w.radius
        return (r * math.cos(el) * math.sin(az), r * math.sin(el), -r * math.cos(el) * math.cos(az)), (0.0, 0.0, 0.0)
    if w.camera == "flythrough":
        t = (i % w.path_frames) / max(1, w.path_frames - 1)
        e = tuple(a + (b - a) * t for a, b in zip(w.fly_from, w.fly_to))
        return e, (e[0] + 1.0, e[1], e[2])
    return tuple(w.eye), (0.0, 0.0, 0.0)


def frame(w: Workload, i: int, frame_number: Optional[int] = None):
    """FrameParams of path position i (RNG frame = frame_number, default i+1)."""
    eye, center = eye_at(w, i)
    return frame_params(eye, center, (0.0, 1.0, 0.0), 45.0, w.width, w.height, ss=w.ss,
                        frame=i + 1 if frame_number is None else frame_number, occlusion=w.occlusion,
                        hierarchical=w.hierarchical)


# ----------------------------------------------------------------- frame loop
class FrameLoop:
    """One render pass plus the resolve per frame, with the ping-pong buffers
    of render_server.cu:86-160: the pass writes color[a] and depth[b]
    (a = frame % 2); TAA blends color[a] with history[b] into history[a]
    using depth[a] as the current depth (the buffer the reference hands it)
    unless `taa_current_depth`; progressive accumulation sums into color[b]."""

    def __init__(self, w: Workload, scene: Scene, stream: Optional[Stream] = None, taa_current_depth: bool = False):
        self.w = w
        self.pass_ = Pass(w.settings(), scene)
        W, H = w.width, w.height
        self.color = [DeviceArray((H, W, 3), np.float32, zero=True) for _ in range(2)]
        self.depth = [DeviceArray((H, W), np.float32, zero=True) for _ in range(2)]
        self.hist = [DeviceArray((H, W, 3), np.float32, zero=True) for _ in range(2)] if w.taa else None
        self.rgb = DeviceArray((H, W, 3), np.uint8, zero=True)
        self.stream = stream or Stream()
        self.prev_vp = np.zeros(16, np.float32)
        self.frames_done = 0
        self.progressive = 0
        self.taa_current_depth = taa_current_depth
        self.last_view = None

    def step(self, f) -> None:
        self.frames_done += 1
        a = self.frames_done % 2
        b = 1 - a
        self.pass_.render(f, self.color[a], self.depth[b], self.stream)
        W, H = self.w.width, self.w.height
        if self.w.taa:
            M = taa_reprojection(self.prev_vp, np.array(f.view, np.float32), np.array(f.proj, np.float32))
            cur_depth = self.depth[b] if self.taa_current_depth else self.depth[a]
            resolve_taa(self.color[a], cur_depth, self.hist[b], self.depth[b], M, 0.9, W, H, self.frames_done,
                        self.hist[a], self.rgb, self.stream)
            self.prev_vp = mat4_mul(np.array(f.proj, np.float32), np.array(f.view, np.float32))
        else:
            view = tuple(f.view)
            moved = self.last_view is not None and view != self.last_view
            self.progressive = 1 if (moved or self.progressive == 0) else self.progressive + 1
            self.last_view = view
            resolve_progressive(self.color[a], self.color[b], W, H, self.progressive, self.rgb, self.stream)

    def output(self) -> np.ndarray:
        """The resolved float image of the last frame (TAA history or color)."""
        self.stream.synchronize()
        a = self.frames_done % 2
        return (self.hist[a] if self.w.taa else self.color[a]).download(self.stream)

    def close(self) -> None:
        self.pass_.close()


# ----------------------------------------------------------------- serving
@dataclass(frozen=True)
class ServeWorkload:
    """Render requests against resident scenes: Poisson arrivals at `rate`
    per second for `seconds`; each request draws a uniform scene, a camera
    on a sphere of that scene's radius around its centre (looking at the
    centre), a resolution and an spp, supersampling `ss`.  Occlusion
    culling is off for independent frames; block frustum culling is on."""
    scenes: tuple                       # Workload per resident scene
    radii: tuple
    centers: tuple
    rate: float = 200.0
    seconds: float = 60.0
    resolutions: tuple = ((854, 480), (1280, 720), (1920, 1080))
    spps: tuple = (1, 4)
    ss: int = 2
    seed: int = 0x5EED0005


SERVE = ServeWorkload(scenes=(WORKLOADS["c1"], WORKLOADS["c2"], WORKLOADS["c3"]), radii=(30.0, 350.0, 600.0),
                      centers=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (375.0, 0.0, 0.0)))


def serve_requests(sw: ServeWorkload, count: Optional[int] = None, rate: Optional[float] = None):
    """[(scene index, FrameParams, arrival seconds)] for the workload (or its
    first `count` requests), seeded and reproducible."""
    rng = np.random.default_rng(sw.seed)
    lam = sw.rate if rate is None else rate
    out = []
    t = 0.0
    i = 0
    while True:
        t += rng.exponential(1.0 / lam) if lam > 0 else 0.0
        if count is None and t > sw.seconds:
            break
        if count is not None and i >= count:
            break
        k = int(rng.integers(len(sw.scenes)))
        w, h = sw.resolutions[int(rng.integers(len(sw.resolutions)))]
        spp = int(sw.spps[int(rng.integers(len(sw.spps)))])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = np.array(sw.centers[k])
        eye = tuple(c + sw.radii[k] * d)
        f = frame_params(eye, tuple(c), (0.0, 1.0, 0.0), 45.0, w, h, ss=sw.ss, frame=i + 1, occlusion=False,
                         hierarchical=True, spp=spp)
        out.append((k, f, t))
        i += 1
    return out


def serve_stats(results, t_first=None) -> dict:
    """Latency percentiles and throughput of completed requests."""
    lat = np.array([r["done_s"] - r["arrival_s"] for r in results])
    done = np.array([r["done_s"] for r in results])
    arr = np.array([r["arrival_s"] for r in results])
    t0 = arr.min() if t_first is None else t_first
    span = max(done.max() - t0, 1e-9)
    return {"requests": len(results), "frames_per_s": len(results) / span, "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "max_ms": float(lat.max() * 1e3), "span_s": span}


# ----------------------------------------------------------------- batches
@dataclass(frozen=True)
class BatchWorkload:
    """Many camera views of one scene per batch: `views` cameras on a sphere
    of `radius` around the origin looking at it, resolution and spp drawn
    per view, supersampling `ss`, block frustum culling."""
    scene: Workload
    views: int = 64
    radius: float = 150.0
    resolutions: tuple = ((854, 480), (1280, 720), (1920, 1080))
    spps: tuple = (1, 2, 4)
    ss: int = 2
    seed: int = 0x5EED0004


BATCH = BatchWorkload(scene=replace(WORKLOADS["c2"], fmt=FORMAT_HALF_FP8))


def batch_views(bw: BatchWorkload, batch_index: int = 0, layout_seed: Optional[int] = None):
    """FrameParams of one batch.  The size / spp layout depends only on the
    workload seed (so a captured graph can be replayed); the camera poses and
    RNG frames change with `batch_index`."""
    lay = np.random.default_rng(bw.seed if layout_seed is None else layout_seed)
    sizes = [bw.resolutions[int(lay.integers(len(bw.resolutions)))] for _ in range(bw.views)]
    spps = [int(bw.spps[int(lay.integers(len(bw.spps)))]) for _ in range(bw.views)]
    rng = np.random.default_rng((bw.seed << 16) + batch_index)
    out = []
    for v in range(bw.views):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = tuple(bw.radius * d)
        w, h = sizes[v]
        out.append(frame_params(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, w, h, ss=bw.ss,
                                frame=batch_index * bw.views + v + 1, occlusion=False, hierarchical=True,
                                spp=spps[v]))
    return out

#!/usr/bin/env python3
"""Benchmark of the MI355X Gaussian point splatting hot path.

One step = one frame of the configured workload through the C ABI:
gps_pass_render (clear -> [cull] -> preprocess -> distribute -> splat ->
[combine -> Hi-Z -> pass 2] -> combine -> [Hi-Z]) plus the resolve
(progressive or TAA) into an RGB8 image, on one HIP stream, inputs resident in
HBM.  Prints ONE JSON line (rank 0).

value = points splatted per second (S = sum over passes of points * K * spp
RNG point draws per frame, summed over the timed frames / timed seconds).
The per-kernel roofline uses the HIP events the library records around every
stage on the render stream during the timed region.

The CPU baseline is the oracle (plain C++ transcription, test
infrastructure) timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gaussian-point-splatting_amd"))

METRIC = "Gpoints/s splatted + fps @1080p, 100M Gaussians; served frames/s at p99 latency"
HBM_PEAK_GBS = 8000.0

# Synthetic configurations (seeded; see DESIGN.md for the exact distributions).
CONFIGS = {
    "c1": dict(desc="1M uniform-box SH0 Gaussians, 1280x720, ss1 (K=1), 1 spp, no culling",
               n=1_000_000, kind="uniform", sh_degree=0, half_extent=10.0, log_scale_mu=-4.0, log_scale_sigma=0.5,
               opacity_sigma=1.0, dc_sigma=0.3, rest_sigma=0.0, clusters=1, seed=0x5EED0001, width=1280, height=720,
               ss=1, spp=1, camera="fixed", eye=(0.0, 0.0, -30.0), occlusion=False, hierarchical=False, taa=False,
               fmt=0),
    "c2": dict(desc="10M clustered SH3 Gaussians (2000 clusters), 1920x1080, ss2 (K=4), 4 spp, 120-frame orbit "
                    "r=350 elev 20deg +3deg/frame, no culling, fp32 storage",
               n=10_000_000, kind="clustered", sh_degree=3, half_extent=100.0, cluster_sigma=2.0,
               log_scale_mu=-3.0, log_scale_sigma=0.7, opacity_sigma=1.5, dc_sigma=0.3, rest_sigma=0.05,
               clusters=2000, seed=0x5EED0002, width=1920, height=1080, ss=2, spp=4, camera="orbit", radius=350.0,
               elevation=20.0, step_deg=3.0, occlusion=False, hierarchical=False, taa=False, fmt=0),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_scene(G, cfg, first=0, count=None):
    p = G.synth_params(cfg["seed"], cfg["n"], kind=cfg["kind"], sh_degree=cfg["sh_degree"],
                       clusters=cfg.get("clusters", 1), half_extent=cfg["half_extent"],
                       cluster_sigma=cfg.get("cluster_sigma", 2.0), log_scale_mu=cfg["log_scale_mu"],
                       log_scale_sigma=cfg["log_scale_sigma"], opacity_sigma=cfg["opacity_sigma"],
                       dc_sigma=cfg["dc_sigma"], rest_sigma=cfg["rest_sigma"])
    return G.synth_scene(p, first, count)


def frames_for(G, cfg, count, start=0):
    out = []
    for i in range(start, start + count):
        if cfg["camera"] == "orbit":
            az = math.radians(cfg["step_deg"] * (i % 120))
            el = math.radians(cfg["elevation"])
            r = cfg["radius"]
            eye = (r * math.cos(el) * math.sin(az), r * math.sin(el), -r * math.cos(el) * math.cos(az))
        else:
            eye = cfg["eye"]
        f = G.frame_params(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, cfg["width"], cfg["height"], ss=cfg["ss"],
                           frame=i + 1, occlusion=cfg["occlusion"], hierarchical=cfg["hierarchical"])
        out.append(f)
    return out


def stage_bytes(cfg, st, N):
    """Algorithmic (compulsory) HBM bytes per frame of each stage, from the
    run's own counters.  See DESIGN.md ("Roofline") for the derivation."""
    nc = (cfg["sh_degree"] + 1) ** 2
    sh_bytes = 3 * nc * 4 if cfg["fmt"] == 0 else 64
    geo_bytes = 44 if cfg["fmt"] == 0 else 12 + 16
    W, H, ss, spp = cfg["width"], cfg["height"], cfg["ss"], cfg["spp"]
    F = math.ceil(W * ss / 64) * math.ceil(H * ss / 64) * 4096 * spp
    passes = st["passes"]
    V = sum(st["valid_gaussians"][:passes])
    P = sum(st["num_points"][:passes])
    ntiles = math.ceil(N / 4096)
    return {
        "clear": F * 8,
        "preprocess": passes * N * (geo_bytes + 4 + 1 / 8) + V * (sh_bytes + 64),
        "distribute": passes * (N * 4 + ntiles * 16) + V * 8,
        "splat": V * 64 + P * 8,
        "combine": (1 + (1 if cfg["occlusion"] else 0)) * (F * 8 + W * H * 16),
        "hiz": (2 if cfg["occlusion"] else 0) * (W * H * 4 * 7 // 3),
        "cull": 0,
    }


def cpu_baseline(G, cfg, sample_n, threads, sort):
    """Oracle on a bounded sample: the first `sample_n` generated Gaussians of
    the scene (Morton-ordered like the GPU scene), frame 1 of the workload,
    same resolution / ss / spp."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # test infrastructure: the CPU baseline leg only
    arrays = make_scene(G, cfg, 0, sample_n)
    if sort:
        arrays = arrays.morton_sorted()
    st = G.RenderSettings(cfg["width"], cfg["height"], cfg["spp"], cfg["ss"],
                          fully_disable_hierarchical_culling=not cfg["hierarchical"])
    o = orc.OraclePass(st, arrays, cfg["fmt"], threads)
    f = frames_for(G, cfg, 1)[0]
    o.render(f)  # warm-up (page-in)
    t0 = time.perf_counter()
    o.render(f)
    dt = time.perf_counter() - t0
    s = o.stats()
    return {"value": s["samples"] / dt / 1e9, "unit": "Gpoints/s", "cores": threads, "kind": "port",
            "sample": f"first {sample_n} of the {cfg['n']} Gaussians, frame 1, {cfg['width']}x{cfg['height']} "
                      f"ss{cfg['ss']} spp{cfg['spp']}: {s['samples']} point draws in {dt:.2f} s",
            "frame_s": dt, "stage_ms": o.stage_ms()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0)
    ap.add_argument("--no-sort", action="store_true",
                    help="keep generation order instead of the load-time Morton order")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path. "
              "Run with --gpus 1.", file=sys.stderr)
        return 2
    import gps_amd as G
    if G.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    cfg = CONFIGS[args.config]
    t0 = time.time()
    arrays = make_scene(G, cfg)
    log(f"[bench] generated {cfg['n']} Gaussians in {time.time() - t0:.1f}s")
    if not args.no_sort:
        t0 = time.time()
        arrays = arrays.morton_sorted()
        log(f"[bench] Morton-ordered (reference --sort-morton-order) in {time.time() - t0:.1f}s")
    scene = G.Scene(arrays, cfg["fmt"])
    if args.no_cpu_baseline:
        del arrays
    settings = G.RenderSettings(cfg["width"], cfg["height"], cfg["spp"], cfg["ss"],
                                fully_disable_hierarchical_culling=not cfg["hierarchical"])
    ps = G.Pass(settings, scene)
    W, H = cfg["width"], cfg["height"]
    color = [G.DeviceArray((H, W, 3), np.float32, zero=True) for _ in range(2)]
    depth = [G.DeviceArray((H, W), np.float32, zero=True) for _ in range(2)]
    rgb = G.DeviceArray((H, W, 3), np.uint8)
    stream = G.Stream()
    K, Wm = args.steps, args.warmup
    frames = frames_for(G, cfg, Wm + K)
    log(f"[bench] scene {scene.device_bytes / 1e9:.2f} GB resident; warmup {Wm}")

    def step(i, f):
        a = (i + 1) % 2
        ps.render(f, color[a], depth[1 - a], stream)
        G.resolve_progressive(color[a], color[1 - a], W, H, 1, rgb, stream)

    for i in range(Wm):
        step(i, frames[i])
    stream.synchronize()
    ps.set_timing(K + 1)
    e0, e1 = G.Event(), G.Event()
    stream.synchronize()
    t_wall = time.perf_counter()
    e0.record(stream)
    for i in range(Wm, Wm + K):
        step(i, frames[i])
    e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t_wall
    gpu_ms = e0.elapsed_ms(e1)
    timing = ps.timing_summary()
    ps.set_timing(0)
    log(f"[bench] timed {K} frames: {gpu_ms:.2f} ms GPU, {wall * 1e3:.2f} ms wall")
    # Work counters: re-render the same (deterministic) frames untimed.
    samples = 0
    atomics = 0
    per_frame = []
    for i in range(Wm, Wm + K):
        step(i, frames[i])
        st = ps.stats(stream)
        samples += st["samples"]
        atomics += st["atomics"]
        per_frame.append(st)
    elapsed_s = max(gpu_ms, wall * 1e3) / 1e3
    ms_per_step = elapsed_s * 1e3 / K
    value = samples / elapsed_s / 1e9
    # Roofline of the dominant kernel.
    stages = timing["stages"]
    dom = max(stages, key=lambda k: stages[k]["ms"])
    bytes_frames = [stage_bytes(cfg, st, cfg["n"]) for st in per_frame]
    dom_bytes = sum(b[dom] for b in bytes_frames)
    dom_launches = max(1, stages[dom]["launches"])
    dom_avg_ms = stages[dom]["ms"] / dom_launches
    achieved = dom_bytes / dom_launches / (dom_avg_ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.config}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "algorithmic_bytes_per_launch": dom_bytes / dom_launches, "avg_launch_ms": dom_avg_ms}
    stage_report = {k: {"ms_per_frame": round(v["ms"] / K, 4), "launches": v["launches"],
                        "GBps": round(sum(b[k] for b in bytes_frames) / max(v["ms"], 1e-9) / 1e6, 1)}
                    for k, v in stages.items() if v["launches"]}
    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        sample_n = args.cpu_sample or min(cfg["n"], 2_000_000)
        log(f"[bench] cpu baseline on {sample_n} Gaussians, {threads} threads")
        cpu = cpu_baseline(G, cfg, sample_n, threads, not args.no_sort)
    st0 = per_frame[0]
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gpoints/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.config}: {cfg['desc']}", "gaussians": cfg["n"], "width": W, "height": H,
                   "ss": cfg["ss"], "spp": cfg["spp"], "format": "fp32" if cfg["fmt"] == 0 else "half_fp8",
                   "order": "generation" if args.no_sort else "morton (load-time, reference --sort-morton-order)"},
        "fps": round(1000.0 / ms_per_step, 2),
        "points_per_frame": round(samples / K),
        "atomics_per_frame": round(atomics / K),
        "scene_order": "generation" if args.no_sort else "morton",
        "gaussians_valid_first_frame": st0["valid_gaussians"][0],
        "points_first_frame": st0["num_points"][0],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stage_report,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gaussian-point-splatting_amd"))

METRIC = "Gpoints/s splatted + fps @1080p, 100M Gaussians; served frames/s at p99 latency"
HBM_PEAK_GBS = 8000.0



def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def stage_bytes(w, st):
    """Algorithmic (compulsory) HBM bytes per frame of each stage, from the
    frame's own counters: what each kernel must read or write at least once.
      clear       F words x 8 B
      cull        32 B per block node + 4 B per 32 nodes, per pass
      preprocess  per pass: cull bit + mask bit per Gaussian; weight (4 B),
                  position and geometry for Gaussians of unculled blocks;
                  SH + a 64-B record per valid Gaussian
      distribute  weights (4 B) of Gaussians of unculled blocks, 16 B per
                  tile per pass, 8 B (index, offset) per valid Gaussian
      splat       the 64-B record per valid Gaussian + one 8-B FB word per
                  point draw
      combine     FB read + colour (12 B) and depth (4 B) per pixel
      hiz         depth read + mip chain written (4/3 of the base level)"""
    N = w.n
    nc = (w.sh_degree + 1) ** 2
    sh_bytes = 3 * nc * 4 if w.fmt == 0 else 64
    pos_geo = 44 if w.fmt == 0 else 12 + 16
    W, H, ss, spp = w.width, w.height, w.ss, w.spp
    F = math.ceil(W * ss / 64) * math.ceil(H * ss / 64) * 4096 * spp
    passes = st["passes"]
    gpb = max(1, st["gpb"])
    nblocks = st["num_blocks"]
    V = sum(st["valid_gaussians"][:passes])
    P = sum(st["num_points"][:passes])
    unculled = sum(max(0, N - st["culled_blocks"][i] * gpb) for i in range(passes)) if w.hierarchical \
        else passes * N
    ntiles = math.ceil(N / 4096)
    combines = passes
    return {
        "clear": F * 8,
        "cull": passes * (nblocks * 32 + math.ceil(nblocks / 32) * 4),
        "preprocess": passes * N * 2 / 8 + unculled * (4 + pos_geo) + V * (sh_bytes + 64),
        "distribute": unculled * 4 + passes * ntiles * 16 + V * 8,
        "splat": V * 64 + P * 8,
        "combine": combines * (F * 8 + W * H * 16),
        "hiz": combines * (W * H * 4 * 7 // 3) if w.occlusion else 0,
    }


def cpu_baseline(G, W, w, sample_n, threads, perm):
    """Oracle on a bounded sample: every (n/sample_n)-th Gaussian of the load
    order (a spatially uniform subset), first timed camera of the workload,
    same resolution / ss / spp / culling."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # test infrastructure: the CPU baseline leg only
    arrays = W.host_sample(w, sample_n, perm)
    o = orc.OraclePass(w.settings(), arrays, w.fmt, threads)
    f = W.frame(w, 0)
    o.render(f)  # warm-up (page-in)
    t0 = time.perf_counter()
    o.render(f)
    dt = time.perf_counter() - t0
    s = o.stats()
    return {"value": s["samples"] / dt / 1e9, "unit": "Gpoints/s", "cores": threads, "kind": "port",
            "sample": f"every {w.n // sample_n}-th of the {w.n} Gaussians ({arrays.n}), camera 0, "
                      f"{w.width}x{w.height} ss{w.ss} spp{w.spp}: {s['samples']} point draws in {dt:.2f} s",
            "frame_s": dt}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed frames (default: the whole camera path)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c3", choices=("c1", "c2", "c3"))
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
    from gps_amd import workloads as W
    if G.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    w = W.WORKLOADS[args.config]
    if args.no_sort:
        w = W.replace(w, morton=False)
    K = args.steps or max(60, w.path_frames)
    Wm = args.warmup
    t0 = time.time()
    perm = W.order(w)
    log(f"[bench] {args.config}: load order of {w.n} Gaussians in {time.time() - t0:.1f}s")
    scene = W.device_scene(w, perm, log=lambda m: log("[bench] " + m))
    if args.no_cpu_baseline:
        perm = None

    def run(loop, positions, stats=None):
        for i in positions:
            loop.step(W.frame(w, i % w.path_frames, i + 1))
            if stats is not None:
                stats.append(loop.pass_.stats(loop.stream))

    stream = G.Stream()
    loop = W.FrameLoop(w, scene, stream)
    log(f"[bench] warmup {Wm} frames, then {K} timed frames (camera positions {Wm}..{Wm + K - 1} mod "
        f"{w.path_frames})")
    run(loop, range(Wm))
    stream.synchronize()
    loop.pass_.set_timing(K + 1)
    e0, e1 = G.Event(), G.Event()
    stream.synchronize()
    t_wall = time.perf_counter()
    e0.record(stream)
    run(loop, range(Wm, Wm + K))
    e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t_wall
    gpu_ms = e0.elapsed_ms(e1)
    timing = loop.pass_.timing_summary()
    loop.close()
    log(f"[bench] timed {K} frames: {gpu_ms:.2f} ms GPU, {wall * 1e3:.2f} ms wall")
    # Work counters: replay the same deterministic frame sequence on a fresh
    # pass (same initial state), reading the counters after every frame.
    loop = W.FrameLoop(w, scene, stream)
    per_frame = []
    run(loop, range(Wm))
    run(loop, range(Wm, Wm + K), per_frame)
    loop.close()
    samples = sum(st["samples"] for st in per_frame)
    atomics = sum(st["atomics"] for st in per_frame)
    elapsed_s = max(gpu_ms, wall * 1e3) / 1e3
    ms_per_step = elapsed_s * 1e3 / K
    value = samples / elapsed_s / 1e9
    # Roofline of the dominant kernel.
    stages = timing["stages"]
    dom = max(stages, key=lambda k: stages[k]["ms"])
    bytes_frames = [stage_bytes(w, st) for st in per_frame]
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
                "algorithmic_bytes_per_launch": round(dom_bytes / dom_launches), "avg_launch_ms": dom_avg_ms}
    stage_report = {k: {"ms_per_frame": round(v["ms"] / K, 4), "launches": v["launches"],
                        "GBps": round(sum(b[k] for b in bytes_frames) / max(v["ms"], 1e-9) / 1e6, 1)}
                    for k, v in stages.items() if v["launches"]}
    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        sample_n = args.cpu_sample or min(w.n, 2_000_000)
        log(f"[bench] cpu baseline on {sample_n} Gaussians, {threads} threads")
        cpu = cpu_baseline(G, W, w, sample_n, threads, perm)
    passes = [st["passes"] for st in per_frame]
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gpoints/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.config}: {w.desc}", "gaussians": w.n, "width": w.width, "height": w.height,
                   "ss": w.ss, "spp": w.spp, "format": "fp32" if w.fmt == 0 else "half_fp8",
                   "order": "generation" if args.no_sort else "morton (load-time, reference --sort-morton-order)",
                   "culling": "occlusion+blocks" if w.hierarchical else ("occlusion" if w.occlusion else "none"),
                   "resolve": "taa" if w.taa else "progressive"},
        "fps": round(1000.0 / ms_per_step, 2),
        "points_per_frame": round(samples / K),
        "atomics_per_frame": round(atomics / K),
        "two_pass_frames": sum(1 for p in passes if p == 2),
        "valid_gaussians_per_frame": round(sum(sum(st["valid_gaussians"][:st["passes"]]) for st in per_frame) / K),
        "culled_blocks_per_frame": round(sum(sum(st["culled_blocks"][:st["passes"]]) for st in per_frame) / K),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stage_report,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

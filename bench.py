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
    pos_geo = {0: 44, 1: 12 + 16, 2: 16 + 1}[w.fmt]  # fp32 planes / half-fp8 / quantized words + opacity byte
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


def atomic_rate(G, words: int) -> float:
    """Random 64-bit atomicMin per second over a `words`-word buffer
    (gps_selftest_atomics, best of 3): the memory-side update rate the
    point-parallel splat's framebuffer atomics run against."""
    buf = G.DeviceArray((words,), np.uint64, zero=True)
    st = G.Stream()
    e0, e1 = G.Event(), G.Event()
    blocks, iters = 8192, 32
    G.check(G.lib.gps_selftest_atomics(buf.ptr, words, blocks, iters, st.handle), "atomics")
    st.synchronize()
    best = 0.0
    for _ in range(3):
        e0.record(st)
        G.check(G.lib.gps_selftest_atomics(buf.ptr, words, blocks, iters, st.handle), "atomics")
        e1.record(st)
        st.synchronize()
        best = max(best, blocks * 256 * iters / (e0.elapsed_ms(e1) / 1e3))
    return best


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


def serving(G, W, args, resident):
    """Serving workload: resident 1M / 10M / 100M scenes, Poisson requests at
    200/s of mixed resolution and spp for `--serve-seconds`, admitted by the
    native scheduler into `--serve-slots` HIP-stream slots.  Also the same
    request mix closed-loop on 1 slot (one request at a time, the plain form)
    and on all slots, for the throughput ratio."""
    sw = W.replace(W.SERVE, seconds=float(args.serve_seconds))
    scenes = []
    for w in sw.scenes:
        if w.name in resident:
            scenes.append(resident[w.name])
        else:
            log(f"[serve] building {w.name} ({w.n} Gaussians)")
            scenes.append(W.device_scene(w, log=lambda m: log("[serve] " + m)))
    mw = max(r[0] for r in sw.resolutions)
    mh = max(r[1] for r in sw.resolutions)

    def server(slots):
        srv = G.Server(slots=slots, max_width=mw, max_height=mh, max_spp=max(sw.spps), max_supersampling=sw.ss)
        ids = [srv.add_scene(sc) for sc in scenes]
        srv.prepare()
        srv.start()
        return srv, ids

    def run(slots, reqs, lead):
        srv, ids = server(slots)
        base = srv.clock() + lead
        for k, f, t in reqs:
            srv.submit(ids[k], f, base + t)
        srv.drain(900_000)
        res = srv.poll(len(reqs) + 8)
        st = srv.stats()
        srv.close()
        bad = [r for r in res if r["status"] != 0]
        if bad or len(res) != len(reqs):
            raise RuntimeError(f"serving: {len(bad)} failed, {len(res)} of {len(reqs)} done")
        return W.serve_stats(res, base), st

    closed = W.serve_requests(sw, count=args.serve_closed, rate=0.0)
    log(f"[serve] closed loop, {len(closed)} requests, 1 slot (plain form)")
    plain, _ = run(1, closed, 0.2)
    log(f"[serve] plain: {plain['frames_per_s']:.1f} frames/s")
    log(f"[serve] closed loop, {args.serve_slots} slots")
    batched, _ = run(args.serve_slots, closed, 0.2)
    log(f"[serve] batched: {batched['frames_per_s']:.1f} frames/s")
    reqs = W.serve_requests(sw)
    log(f"[serve] open loop: {len(reqs)} requests at {sw.rate:.0f}/s over {sw.seconds:.0f}s, {args.serve_slots} slots")
    live, st = run(args.serve_slots, reqs, 1.0)
    log(f"[serve] p50 {live['p50_ms']:.2f} ms p99 {live['p99_ms']:.2f} ms, {live['frames_per_s']:.1f} frames/s")
    # The same mix at 80 % of the closed-loop capacity, where queues stay bounded.
    rate2 = round(0.8 * batched["frames_per_s"])
    sustained = None
    if args.serve_sustained_seconds > 0 and rate2 > 0:
        sw2 = W.replace(sw, rate=float(rate2), seconds=float(args.serve_sustained_seconds))
        reqs2 = W.serve_requests(sw2)
        log(f"[serve] open loop at 80 % of capacity: {len(reqs2)} requests at {rate2}/s")
        s2, _ = run(args.serve_slots, reqs2, 1.0)
        log(f"[serve] p50 {s2['p50_ms']:.2f} ms p99 {s2['p99_ms']:.2f} ms, {s2['frames_per_s']:.1f} frames/s")
        sustained = {"rate_per_s": rate2, "seconds": sw2.seconds, "requests": s2["requests"],
                     "frames_per_s": round(s2["frames_per_s"], 2), "p50_ms": round(s2["p50_ms"], 3),
                     "p99_ms": round(s2["p99_ms"], 3)}
    return {"workload": "c5: resident 1M fp32 / 10M fp32 / 100M compact scenes; Poisson arrivals "
                        f"{sw.rate:.0f}/s for {sw.seconds:.0f} s; 854x480 / 1280x720 / 1920x1080, spp 1 / 4, ss 2, "
                        "random sphere cameras; frustum block culling, no occlusion culling",
            "slots": args.serve_slots, "requests": live["requests"], "frames_per_s": round(live["frames_per_s"], 2),
            "p50_ms": round(live["p50_ms"], 3), "p99_ms": round(live["p99_ms"], 3), "max_ms": round(live["max_ms"], 3),
            "closed_loop_requests": len(closed), "plain_frames_per_s": round(plain["frames_per_s"], 2),
            "batched_frames_per_s": round(batched["frames_per_s"], 2),
            "batched_over_plain": round(batched["frames_per_s"] / plain["frames_per_s"], 3),
            "closed_loop_p99_ms": round(batched["p99_ms"], 3), "sustained": sustained,
            "max_batch": int(st["max_batch"]),
            "hbm_used_gb": round(st["hbm_used_bytes"] / 1e9, 2)}


def batched(G, W, args):
    """Batched views: the 10M SH3 scene in compact form, 64 views per batch
    on a sphere of radius 150, resolution and spp drawn per view, ss 2.
    Plain form: the views rendered one after another through one pass.
    Batched form: one gps_batch_render per batch (autotuned splat variant,
    replayed from a HIP graph).  Both render the same frames."""
    bw = W.replace(W.BATCH, views=args.batch_views)
    log(f"[batch] building {bw.scene.name} compact scene ({bw.scene.n} Gaussians)")
    scene = W.device_scene(bw.scene, log=lambda m: log("[batch] " + m))
    nb = args.batch_count
    frames = [W.batch_views(bw, b) for b in range(nb + 1)]
    stream = G.Stream()
    cols = [G.DeviceArray((f.height, f.width, 3), np.float32) for f in frames[0]]
    deps = [G.DeviceArray((f.height, f.width), np.float32) for f in frames[0]]
    mw = max(r[0] for r in bw.resolutions)
    mh = max(r[1] for r in bw.resolutions)
    single = G.Pass(G.RenderSettings(mw, mh, max(bw.spps), bw.ss, flexible_frame_size=True), scene)
    for f, c, d in zip(frames[0], cols, deps):
        single.render(f, c, d, stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for b in range(1, nb + 1):
        for f, c, d in zip(frames[b], cols, deps):
            single.render(f, c, d, stream)
    stream.synchronize()
    plain_s = time.perf_counter() - t0
    single.close()
    log(f"[batch] plain: {nb * bw.views} frames in {plain_s * 1e3:.1f} ms")
    batch = G.Batch(scene, bw.views, mw, mh, max(bw.spps), bw.ss)
    tuned = batch.autotune(frames[0], cols, deps, stream, repeats=2)
    log(f"[batch] autotune: {tuned}")
    batch.set_tuning(tuned["splat_threads"], tuned["splat_points_per_thread"], True)
    batch.render(frames[0], cols, deps, stream)  # capture
    stream.synchronize()
    t0 = time.perf_counter()
    for b in range(1, nb + 1):
        batch.render(frames[b], cols, deps, stream)
    stream.synchronize()
    batch_s = time.perf_counter() - t0
    samples = sum(batch.view_stats(v, stream)["samples"] for v in range(bw.views))
    log(f"[batch] batched: {nb * bw.views} frames in {batch_s * 1e3:.1f} ms")
    out = {"workload": f"c4: {bw.scene.n} clustered SH3 Gaussians, compact format, {bw.views} views per batch on a "
                       f"sphere of radius {bw.radius:.0f}, resolutions 854x480 / 1280x720 / 1920x1080 and spp 1 / 2 / 4 "
                       f"per view, ss {bw.ss}, block frustum culling",
           "views_per_batch": bw.views, "batches": nb,
           "plain_frames_per_s": round(nb * bw.views / plain_s, 2),
           "batched_frames_per_s": round(nb * bw.views / batch_s, 2),
           "batched_over_plain": round(plain_s / batch_s, 3),
           "batched_ms_per_batch": round(batch_s * 1e3 / nb, 3),
           "points_per_batch": samples, "batched_gpoints_per_s": round(samples * nb / batch_s / 1e9, 3),
           "tuning": tuned, "graph": True, "device_gb": round(batch.device_bytes / 1e9, 2)}
    batch.close()
    scene.close()
    return out


def bvh_order_leg(G, W, args) -> dict:
    """The load-time BVH order (gps_bvh_order, the reference's
    GaussianBVH::build_bvh_and_reorder) of the configured workload's scene in
    generation order, i.e. as an un-presorted PLY would arrive; host arrays
    in, perm out, timed end to end.  cpu_baseline: the reference's own
    tinybvh (oracle/_ref, its threaded build) on a bounded prefix of the same
    scene; the oracle's serial restatement where _ref is absent."""
    w = W.WORKLOADS[args.config]
    t0 = time.time()
    a = W.host_scene(w, np.arange(w.n, dtype=np.uint64))
    log(f"[bench] bvh order: generated {w.n} Gaussians in {time.time() - t0:.1f}s")
    a.bvh_order()  # first call: device memory and code objects
    runs = []
    for _ in range(args.bvh_order_runs):
        t = time.perf_counter()
        perm = a.bvh_order()
        runs.append(time.perf_counter() - t)
    sec = min(runs)
    assert np.array_equal(np.sort(perm), np.arange(w.n, dtype=np.uint32))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # CPU baseline only
    ns = min(w.n, args.bvh_ref_sample)
    m, q, sc = a.means[:ns], a.rotations[:ns], a.scales[:ns]
    ref = orc.tinybvh_ref()
    t = time.perf_counter()
    if ref is not None:
        orc.tinybvh_ref_order(ref, orc.bvh_fragments(m, q, sc))
        kind, cores = "reference", 32  # tinybvh runs the subtrees below depth 5 on 32 threads
    else:
        orc.bvh_order(m, q, sc)
        kind, cores = "port", 1
    cpu_sec = time.perf_counter() - t
    return {
        "metric": "load-time BVH order (reference build_bvh_and_reorder), Gaussians/s", "value": round(w.n / sec),
        "unit": "Gaussians/s", "n_gpus": 1, "steps": len(runs), "warmup": 1, "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.config}: {w.desc}, generation order", "gaussians": w.n},
        "runs_ms": [round(r * 1e3, 2) for r in runs],
        "cpu_baseline": {"value": round(ns / cpu_sec), "unit": "Gaussians/s", "cores": cores, "kind": kind,
                         "sample": f"first {ns} Gaussians of the same scene, {cpu_sec:.2f} s"},
    }


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
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="serving measurement (0 = skip)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-sustained-seconds", type=float, default=30.0,
                    help="open-loop seconds at 80 %% of the closed-loop capacity (0 = skip)")
    ap.add_argument("--serve-closed", type=int, default=300, help="requests of the closed-loop comparison")
    ap.add_argument("--batch-views", type=int, default=64, help="views per batch (0 = skip the batched form)")
    ap.add_argument("--format", choices=("workload", "fp32", "half_fp8", "quantized"), default="workload",
                    help="scene storage format (default: the workload's own)")
    ap.add_argument("--first-frame", type=int, default=-1,
                    help="camera-path position of the first timed frame (default: right after the warm-up)")
    ap.add_argument("--no-splat-forms", action="store_true",
                    help="skip re-timing the frames with each splat form forced")
    ap.add_argument("--batch-count", type=int, default=5, help="timed batches")
    ap.add_argument("--bvh-order", action="store_true",
                    help="measure only the load-time BVH order of the workload's scene (one JSON line)")
    ap.add_argument("--bvh-order-runs", type=int, default=3)
    ap.add_argument("--bvh-ref-sample", type=int, default=8_000_000,
                    help="Gaussians of the reference tinybvh CPU timing")
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
    if args.bvh_order:
        print(json.dumps(bvh_order_leg(G, W, args)), flush=True)
        return 0
    w = W.WORKLOADS[args.config]
    if args.format != "workload":
        w = W.replace(w, fmt={"fp32": G.FORMAT_FP32, "half_fp8": G.FORMAT_HALF_FP8,
                              "quantized": G.FORMAT_QUANTIZED}[args.format])
    if args.no_sort:
        w = W.replace(w, morton=False)
    K = args.steps or max(60, w.path_frames)
    Wm = args.warmup
    F0 = max(0, args.first_frame - Wm) if args.first_frame >= 0 else 0  # warm-up positions F0 .. F0 + Wm - 1
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
    log(f"[bench] warmup {Wm} frames, then {K} timed frames (camera positions {F0 + Wm}..{F0 + Wm + K - 1} mod "
        f"{w.path_frames})")
    run(loop, range(F0, F0 + Wm))
    stream.synchronize()
    loop.pass_.set_timing(K + 1)
    e0, e1 = G.Event(), G.Event()
    stream.synchronize()
    t_wall = time.perf_counter()
    e0.record(stream)
    run(loop, range(F0 + Wm, F0 + Wm + K))
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
    loop.pass_.set_counters(True)
    per_frame = []
    run(loop, range(F0, F0 + Wm))
    run(loop, range(F0 + Wm, F0 + Wm + K), per_frame)
    loop.close()
    # The same frames with each splat form forced (binned vs point-parallel
    # with one global atomic per sample): the chosen and the plain form.
    splat_forms = {}
    forms = () if args.no_splat_forms else (("point_parallel", G.Pass.SPLAT_POINT_PARALLEL),
                                            ("binned", G.Pass.SPLAT_BINNED))
    for name, mode in forms:
        loop = W.FrameLoop(w, scene, stream)
        loop.pass_.set_splat_mode(mode)
        run(loop, range(F0, F0 + Wm))
        stream.synchronize()
        loop.pass_.set_timing(K + 1)
        t0 = time.perf_counter()
        run(loop, range(F0 + Wm, F0 + Wm + K))
        stream.synchronize()
        sp = loop.pass_.timing_summary()["stages"]["splat"]["ms"] / K
        splat_forms[name] = {"ms_per_frame": round((time.perf_counter() - t0) * 1e3 / K, 4),
                             "splat_ms_per_frame": round(sp, 4)}
        loop.close()
    log(f"[bench] splat forms: {splat_forms}")
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
    if dom == "splat" and atomics:
        # The splat's own bound: memory-side 64-bit atomics at random
        # framebuffer words, priced against the measured random-atomicMin rate.
        peak_atomics = atomic_rate(G, w.width * w.ss * w.height * w.ss * w.spp)
        done = atomics / (stages["splat"]["ms"] / 1e3)
        roofline["atomics"] = {"achieved_per_s": round(done), "peak_per_s": round(peak_atomics),
                               "frac": round(done / peak_atomics, 4), "per_launch": round(atomics / dom_launches),
                               "peak": "gps_selftest_atomics: random 64-bit atomicMin over the framebuffer's words"}
    stage_report = {k: {"ms_per_frame": round(v["ms"] / K, 4), "launches": v["launches"],
                        "GBps": round(sum(b[k] for b in bytes_frames) / max(v["ms"], 1e-9) / 1e6, 1)}
                    for k, v in stages.items() if v["launches"]}
    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        sample_n = args.cpu_sample or min(w.n, 2_000_000)
        log(f"[bench] cpu baseline on {sample_n} Gaussians, {threads} threads")
        cpu = cpu_baseline(G, W, w, sample_n, threads, perm)
    batch_report = None
    if args.batch_views > 0:
        batch_report = batched(G, W, args)
    serve = None
    if args.serve_seconds > 0:
        serve = serving(G, W, args, {w.name: scene})
    passes = [st["passes"] for st in per_frame]
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gpoints/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.config}: {w.desc}", "gaussians": w.n, "width": w.width, "height": w.height,
                   "ss": w.ss, "spp": w.spp, "format": ("fp32", "half_fp8", "quantized")[w.fmt],
                   "order": "generation" if args.no_sort else "morton (load-time, reference --sort-morton-order)",
                   "culling": "occlusion+blocks" if w.hierarchical else ("occlusion" if w.occlusion else "none"),
                   "resolve": "taa" if w.taa else "progressive"},
        "fps": round(1000.0 / ms_per_step, 2),
        "points_per_frame": round(samples / K),
        "atomics_per_frame": round(atomics / K),
        "two_pass_frames": sum(1 for p in passes if p == 2),
        "valid_gaussians_per_frame": round(sum(sum(st["valid_gaussians"][:st["passes"]]) for st in per_frame) / K),
        "culled_blocks_per_frame": round(sum(sum(st["culled_blocks"][:st["passes"]]) for st in per_frame) / K),
        "splat_forms": splat_forms,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batch_report,
        "serving": serve,
        "stages": stage_report,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

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

Parity (unless --no-parity): the workload at its full size against the oracle
(plain C++ transcription, test infrastructure, used only as the checker):
the first path positions in order from a fresh pass, later ones with the
frame state carried from the GPU; bit-exactness and PSNR go into the line.
The CPU baseline is the oracle's time for the first timed frame of that
check (the full scene, one frame); with --no-parity, the oracle on every
50-th Gaussian of the scene.
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
      cull        32 B per block node + 4 B per 32 nodes, per pass; with
                  occlusion, the group test: 48 B per group of unculled
                  blocks + its mask bits
      preprocess  per pass: cull bit + mask bit per Gaussian; weight (4 B)
                  for Gaussians of unculled blocks, position and geometry for
                  those not in groups the group test skipped; SH + a 64-B
                  record per valid Gaussian
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
    skipped = min(unculled, 64 * sum(st.get("culled_groups", (0, 0))[:passes]))
    group_test = (unculled / 64 * 48 + passes * math.ceil(N / 64) / 8) if w.occlusion else 0
    ntiles = math.ceil(N / 4096)
    combines = passes
    return {
        "clear": F * 8,
        "cull": passes * (nblocks * 32 + math.ceil(nblocks / 32) * 4) + group_test,
        "preprocess": passes * N * 2 / 8 + unculled * 4 + (unculled - skipped) * pos_geo + V * (sh_bytes + 64),
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


def _psnr(a, b):
    """PSNR in dB against a peak of 1.0 ("inf" for identical images)."""
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return "inf" if mse == 0.0 else round(10.0 * math.log10(1.0 / mse), 2)


def _host_scene_for_oracle(G, W, w, perm, orc, threads):
    """The oracle's copy of the resident scene, at its full size, generated
    in the same load order and encoded in chunks into the storage format the
    GPU holds (the compact format is never expanded to fp32 on the host)."""
    n = w.n
    if w.fmt != G.FORMAT_HALF_FP8:
        return orc.OraclePass(w.settings(), W.host_scene(w, perm), w.fmt, threads)
    means = np.empty((n, 3), np.float32)
    enc = {"rotations": np.empty((n, 4), np.uint16), "scales": np.empty((n, 3), np.uint16),
           "opacities": np.empty((n,), np.uint16), "sh_dc": np.empty((n, 3), np.uint16),
           "sh_rest": np.empty((n, 45), np.uint8), "sh_scale": np.empty((n,), np.float32)}
    ch = 1 << 22
    for b in range(0, n, ch):
        e = min(n, b + ch)
        a = W._generate(w, perm[b:e])
        means[b:e] = a.means
        ce = G.encode_half_fp8(a)
        for k in enc:
            enc[k][b:e] = ce[k]
        del a, ce
    return orc.OraclePass.from_half_fp8(w.settings(), means, enc, w.sh_degree, threads)


def parity_leg(G, W, w, scene, perm, threads, in_order, carried):
    """Parity at the workload's full size: the GPU path against the oracle
    on the same scene, frame by frame.  Positions in `in_order` are rendered
    by both from a fresh pass in path order; for a position in `carried` the
    GPU renders the path up to it and hands the oracle the state a frame
    carries into the next (Hi-Z chain, the valid mask the next first pass
    leaves in place) and the TAA history, then both render that position.
    Compared: framebuffer words, colour and depth images (bit for bit and as
    PSNR), the TAA output and its RGB8 image.  The oracle's render time of
    each frame is kept for the CPU baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # the checker, test infrastructure
    t_start = time.time()
    o = _host_scene_for_oracle(G, W, w, perm, orc, threads)
    log(f"[parity] oracle scene of {w.n} Gaussians ready in {time.time() - t_start:.1f}s")
    Wd, Hd = w.width, w.height
    stream = G.Stream()
    loop = W.FrameLoop(w, scene, stream)
    p = loop.pass_
    o_color = [np.zeros((Hd, Wd, 3), np.float32) for _ in range(2)]
    o_depth = [np.zeros((Hd, Wd), np.float32) for _ in range(2)]
    o_hist = [np.zeros((Hd, Wd, 3), np.float32) for _ in range(2)]
    check = set(in_order) | set(carried)
    frames = []
    oracle_in_order_done = 0
    for pos in range(max(check) + 1):
        f = W.frame(w, pos % w.path_frames, pos + 1)
        if pos not in check:
            loop.step(f)
            continue
        k = loop.frames_done + 1
        a, b = k % 2, 1 - k % 2
        prev_vp = loop.prev_vp.copy()
        if pos in carried:
            stream.synchronize()
            st = p.stats(stream)
            passes = st["passes"]
            flipped = passes > 0 and st["num_points"][passes - 1] > 0
            other = p.debug(G.DBG_MASK_CURRENT if flipped else G.DBG_MASK_PREVIOUS, stream)
            o.set_state(p.debug(G.DBG_HIZ, stream), other)
            if w.taa:
                o_depth[a] = loop.depth[a].download(stream)
                o_hist[b] = loop.hist[b].download(stream)
            state = f"Hi-Z + mask + TAA history carried from the GPU after position {pos - 1}"
        else:
            assert oracle_in_order_done == pos, "in-order positions must start at 0 and be consecutive"
            oracle_in_order_done += 1
            state = "in path order from a fresh pass"
        loop.step(f)
        stream.synchronize()
        fb = p.debug(G.DBG_FRAMEBUFFER, stream)
        g_color = loop.color[a].download(stream)
        g_depth = loop.depth[b].download(stream)
        st = p.stats(stream)
        t0 = time.perf_counter()
        oc, od = o.render(f)
        dt = time.perf_counter() - t0
        so = o.stats()
        ofb = o.debug(G.DBG_FRAMEBUFFER)
        o_color[a], o_depth[b] = oc, od
        rec = {"position": pos, "state": state, "samples": int(so["samples"]), "oracle_s": round(dt, 3),
               "fb_words": int(fb.size), "fb_words_differ": int(np.count_nonzero(fb != ofb)),
               "color_bit_exact": bool(np.array_equal(g_color.view(np.uint32), oc.view(np.uint32))),
               "color_psnr_db": _psnr(g_color, oc),
               "depth_bit_exact": bool(np.array_equal(g_depth.view(np.uint32), od.view(np.uint32))),
               "counters_equal": all(st[key] == so[key] for key in ("valid_gaussians", "num_points",
                                                                     "culled_blocks", "passes"))}
        if w.taa:
            M = G.taa_reprojection(prev_vp, np.array(f.view, np.float32), np.array(f.proj, np.float32))
            o_hist[a], o_rgb = orc.resolve_taa(o_color[a], o_depth[a], o_hist[b], o_depth[b], M, 0.9, Wd, Hd, k)
            g_hist = loop.hist[a].download(stream)
            rec["taa_bit_exact"] = bool(np.array_equal(g_hist.view(np.uint32), o_hist[a].view(np.uint32)))
            rec["taa_psnr_db"] = _psnr(g_hist, o_hist[a])
            rec["rgb8_equal"] = bool(np.array_equal(loop.rgb.download(stream), o_rgb))
        log(f"[parity] position {pos}: {rec}")
        frames.append(rec)
    loop.close()
    o.close()
    exact = all(r["fb_words_differ"] == 0 and r["color_bit_exact"] and r["depth_bit_exact"] and r["counters_equal"]
                and r.get("taa_bit_exact", True) and r.get("rgb8_equal", True) for r in frames)
    psnrs = [r[k] for r in frames for k in ("color_psnr_db", "taa_psnr_db") if k in r and r[k] != "inf"]
    return {"workload": f"{w.name} at full size ({w.n} Gaussians, {w.width}x{w.height}, ss {w.ss}, spp {w.spp})",
            "oracle_threads": threads, "bit_exact": exact, "min_psnr_db": min(psnrs) if psnrs else "inf",
            "frames": frames, "seconds": round(time.time() - t_start, 1)}


def _read_csv(path):
    import csv
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


COPY_CAL_BYTES = 1 << 30  # the calibration copy of --profile-window: 1 GiB read + 1 GiB written


def _kname(r):
    """Kernel name without arguments, return type and namespace."""
    return r["Kernel_Name"].split("(")[0].replace("void ", "").strip().replace("gps::", "")


def _window(trace_csv, gap_s):
    """Dispatches after the last idle gap of more than gap_s seconds."""
    rows = _read_csv(trace_csv)
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    cut = 0
    end = int(rows[0]["End_Timestamp"]) if rows else 0
    for i in range(1, len(rows)):
        if int(rows[i]["Start_Timestamp"]) - end > gap_s * 1e9:
            cut = i
        end = max(end, int(rows[i]["End_Timestamp"]))
    return rows, rows[cut:]


def summarize_trace(traces, counters=(), gap_s=0.5, traffic_out=None, window_name=""):
    """Per-kernel statistics of the timed frames of --profile-window runs
    under rocprofv3: the dispatches after the last idle gap of more than
    `gap_s` seconds (the run sleeps between the warm-up and the timed
    frames).  The first trace gives the kernel times.  Each further trace
    comes with the counter CSV of the same run (--counters, in order), whose
    per-dispatch means over the same window are listed; FETCH_SIZE and
    WRITE_SIZE (KiB) are also scaled by what they read for the calibration
    copy of exactly 1 GiB each way that closes the window (selftest_copy_kernel)
    and, with `traffic_out`, written as HBM bytes per dispatch per stage."""
    rows, win = _window(traces[0], gap_s)
    per = {}
    for r in win:
        d = per.setdefault(_kname(r), {"calls": 0, "ns": 0})
        d["calls"] += 1
        d["ns"] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    total = sum(d["ns"] for k, d in per.items() if k != "selftest_copy_kernel") or 1
    t0, t1 = int(win[0]["Start_Timestamp"]), max(int(r["End_Timestamp"]) for r in win)
    lines = [f"window: {len(win)} of {len(rows)} dispatches (after an idle gap > {gap_s} s), "
             f"{(t1 - t0) / 1e6:.3f} ms first start to last end",
             f"{'kernel':50s} {'calls':>6s} {'avg us':>10s} {'total ms':>10s} {'share':>6s}"]
    for k, d in sorted(per.items(), key=lambda kv: -kv[1]["ns"]):
        share = f"{100.0 * d['ns'] / total:5.1f}%" if k != "selftest_copy_kernel" else "  cal."
        lines.append(f"{k[:50]:50s} {d['calls']:6d} {d['ns'] / d['calls'] / 1e3:10.1f} {d['ns'] / 1e6:10.3f} {share}")
    means = {}  # (kernel, counter) -> mean per dispatch
    for tr, cc in zip(traces[1:], counters):
        _, w2 = _window(tr, gap_s)
        ids = {}
        for r in w2:
            ids.setdefault(_kname(r), set()).add(r["Dispatch_Id"])
        crows = _read_csv(cc)
        sums = {}
        for r in crows:
            k = _kname(r)
            if r["Dispatch_Id"] not in ids.get(k, ()):
                continue
            acc = sums.setdefault((k, r["Counter_Name"]), [0.0, set()])
            acc[0] += float(r["Counter_Value"])
            acc[1].add(r["Dispatch_Id"])
        for key, (v, dset) in sums.items():
            means[key] = (v / len(dset), len(dset))
    if means:
        lines.append("")
        lines.append(f"{'kernel':50s} {'counter':24s} {'per dispatch':>16s} {'dispatches':>10s}")
        for (k, c), (v, n) in sorted(means.items(), key=lambda kv: (-per.get(kv[0][0], {"ns": 0})["ns"], kv[0][1])):
            lines.append(f"{k[:50]:50s} {c:24s} {v:16.1f} {n:10d}")
    # VALU issue utilisation: a wave64 VALU instruction occupies a 16-lane
    # SIMD for 4 cycles; GRBM_GUI_ACTIVE counts the busy cycles of every XCD
    # (MI355X: 8 XCDs, 256 CUs, 4 SIMDs each).
    valu = {}
    for (k, c), (v, n) in means.items():
        if c == "SQ_INSTS_VALU" and (k, "GRBM_GUI_ACTIVE") in means:
            cycles = means[(k, "GRBM_GUI_ACTIVE")][0] / 8.0
            if cycles > 0:
                valu[k] = v * 4.0 / (256 * 4 * cycles)
    if valu:
        lines.append("")
        lines.append("VALU issue utilisation (SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x GRBM_GUI_ACTIVE / 8 XCDs)): " +
                     ", ".join(f"{k} {u:.2f}" for k, u in sorted(valu.items(), key=lambda kv: -per.get(kv[0], {"ns": 0})["ns"])
                               if per.get(k, {"ns": 0})["ns"] > 0))
    cal = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        m = means.get(("selftest_copy_kernel", c))
        if m and m[0] > 0:
            cal[c] = COPY_CAL_BYTES / (m[0] * 1024.0)
    if cal:
        lines.append("")
        lines.append("calibration (1 GiB copy each way): " +
                     ", ".join(f"{c} x {f:.3f} per KiB counted = bytes" for c, f in cal.items()))
        stage_of = {"preprocess_kernel": "preprocess", "splat_kernel": "splat", "distribute_kernel": "distribute",
                    "tile_scan_kernel": "distribute", "combine_staged_kernel": "combine", "combine_kernel": "combine",
                    "hiz_build_kernel": "hiz", "hiz_top_kernel": "hiz", "cull_kernel": "cull", "clear_kernel": "clear",
                    "group_cull_live_kernel": "cull", "group_cull_kernel": "cull",
                    "taa_kernel": "taa", "progressive_kernel": "resolve"}
        traffic = {}
        for (k, c), (v, n) in means.items():
            base = k.split("<")[0]
            if c not in cal or base not in stage_of:
                continue
            calls = per.get(k, {"calls": n})["calls"]
            st = traffic.setdefault(stage_of[base], {"fetch_bytes": 0.0, "write_bytes": 0.0, "kernels": []})
            # a stage's bytes per launch: its kernels' bytes per dispatch times their dispatches per stage launch
            st["fetch_bytes" if c == "FETCH_SIZE" else "write_bytes"] += v * 1024.0 * cal[c] * calls
            if k not in st["kernels"]:
                st["kernels"].append(k)
            st["_calls"] = max(st.get("_calls", 0), calls)
        out = {"_meta": {"window": window_name, "calibration": {c: round(f, 4) for c, f in cal.items()},
                         "valu_busy": {k: round(u, 3) for k, u in valu.items()},
                         "unit": "HBM bytes per stage launch = sum over the stage's kernels of "
                                 "(FETCH_SIZE KiB x 1024 x fetch factor + WRITE_SIZE KiB x 1024 x write factor) "
                                 "per dispatch, x their dispatches per stage launch"}}
        for stg, d in traffic.items():
            calls = d.pop("_calls")
            launches = max(per[k]["calls"] for k in d["kernels"] if k in per) if any(k in per for k in d["kernels"]) \
                else calls
            # kernels that run once per stage launch (the common case): per-dispatch bytes
            f = d["fetch_bytes"] / launches
            wb = d["write_bytes"] / launches
            out[stg] = round(f + wb)
            out.setdefault("_detail", {})[stg] = {"fetch_bytes": round(f), "write_bytes": round(wb),
                                                  "kernels": d["kernels"]}
        lines.append("traffic per stage launch (bytes): " +
                     ", ".join(f"{k} {v}" for k, v in out.items() if not k.startswith("_")))
        if traffic_out:
            with open(traffic_out, "w") as fh:
                json.dump(out, fh, indent=1)
    return "\n".join(lines)


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


def request_types(G, W, args) -> dict:
    """GPU time per serving request type: for every resident scene and spp of
    the serving mix, 12 of its requests (mixed resolutions) rendered back to
    back on one pass with the server's settings (the largest request size,
    flexible frame size, automatic block size), with the splat form each
    request got and the points the ordered splat skipped on the first one."""
    sw = W.SERVE
    reqs = W.serve_requests(sw, count=400, rate=0.0)
    mw = max(r[0] for r in sw.resolutions)
    mh = max(r[1] for r in sw.resolutions)
    color = G.DeviceArray((mh, mw, 3), np.float32)
    depth = G.DeviceArray((mh, mw), np.float32)
    stream = G.Stream()
    rows = []
    only = set(args.request_scenes.split(",")) if args.request_scenes else None
    for k, w in enumerate(sw.scenes):
        if only is not None and w.name not in only:
            continue
        log(f"[types] {w.name}: {w.n} Gaussians")
        scene = W.device_scene(w, log=lambda m: log("[types] " + m))
        p = G.Pass(G.RenderSettings(mw, mh, max(sw.spps), sw.ss, flexible_frame_size=True), scene)
        for spp in sw.spps:
            fs = [f for kk, f, _ in reqs if kk == k and f.samples_per_pixel == spp][:12]
            p.set_counters(False, False)
            p.render(fs[0], color, depth, stream)
            stream.synchronize()
            p.set_timing(len(fs) + 1)
            e0, e1 = G.Event(), G.Event()
            e0.record(stream)
            for f in fs:
                p.render(f, color, depth, stream)
            e1.record(stream)
            stream.synchronize()
            stages = p.timing_summary()["stages"]
            p.set_counters(False, True)
            p.render(fs[0], color, depth, stream)
            st = p.stats(stream)
            rows.append({"scene": w.name, "spp": spp, "requests": len(fs),
                         "ms_per_request": round(e0.elapsed_ms(e1) / len(fs), 3),
                         "stages_ms": {n: round(v["ms"] / len(fs), 3) for n, v in stages.items() if v["launches"]},
                         "first_request": {"size": [fs[0].width, fs[0].height], "points": st["num_points"][0],
                                           "splat_binned": st["splat_binned"], "splat_fallback": st["splat_fallback"],
                                           "ordered_skipped": sum(p.ordered_skipped(stream))}})
            log(f"[types] {w.name} spp {spp}: {rows[-1]['ms_per_request']} ms per request")
        p.close()
        del scene
    out = {"request_types": rows, "workload": "c5 request mix: 1M fp32 / 10M fp32 / 100M compact scenes, "
           "854x480 / 1280x720 / 1920x1080, ss 2, random sphere cameras, no occlusion culling"}
    if only is None:
        mean = sum(r["ms_per_request"] for r in rows) / len(rows)
        out["mean_ms_per_request"] = round(mean, 3)
        out["requests_per_s_back_to_back"] = round(1000.0 / mean, 1)
    return out


def _e4m3_table():
    """FP8 E4M3 (OCP, finite-only) value of every code, from its bit fields."""
    codes = np.arange(256)
    sign = np.where(codes & 0x80, -1.0, 1.0)
    e = (codes >> 3) & 0xF
    m = codes & 0x7
    val = np.where(e == 0, m / 8.0 * 2.0 ** -6, (1.0 + m / 8.0) * 2.0 ** (e.astype(np.float64) - 7))
    val = np.where((e == 15) & (m == 7), np.nan, val)
    return sign * val


def attribute_errors(G, W, w, perm, chunk=1 << 21):
    """Largest error of the compact storage (fp16 rotation / scale / opacity,
    fp8-E4M3 SH coefficients with a per-Gaussian fp32 scale) against the
    original values, all in float64, over the Gaussians `perm` lists
    (relative rotation error over components in fp16's normal range)."""
    lut = _e4m3_table()
    worst = {"rotation_abs": 0.0, "rotation_rel": 0.0, "scale_rel": 0.0, "opacity_abs": 0.0, "sh_dc_abs": 0.0,
             "sh_rest_abs": 0.0, "sh_rest_rel_to_max": 0.0}
    nc = (w.sh_degree + 1) ** 2
    for b in range(0, len(perm), chunk):
        a = W._generate(w, perm[b:b + chunk])
        e = G.encode_half_fp8(a)
        q = a.rotations.astype(np.float64)
        dq = e["rotations"].view(np.float16).astype(np.float64) - q
        worst["rotation_abs"] = max(worst["rotation_abs"], float(np.abs(dq).max()))
        nz = np.abs(q) >= 2.0 ** -14  # fp16 normal range
        worst["rotation_rel"] = max(worst["rotation_rel"], float((np.abs(dq[nz]) / np.abs(q[nz])).max()))
        sc = a.scales.astype(np.float64)
        worst["scale_rel"] = max(worst["scale_rel"],
                                 float((np.abs(e["scales"].view(np.float16).astype(np.float64) - sc) / sc).max()))
        worst["opacity_abs"] = max(worst["opacity_abs"], float(np.abs(
            e["opacities"].view(np.float16).astype(np.float64) - a.opacities.astype(np.float64)).max()))
        sh = a.shs.reshape(-1, nc * 3).astype(np.float64)
        worst["sh_dc_abs"] = max(worst["sh_dc_abs"], float(np.abs(
            e["sh_dc"].view(np.float16).astype(np.float64) - sh[:, :3]).max()))
        if nc > 1:
            rest = sh[:, 3:]
            dec = lut[e["sh_rest"][:, :rest.shape[1]]] * e["sh_scale"].astype(np.float64)[:, None]
            err = np.abs(dec - rest)
            mx = np.abs(rest).max(axis=1)
            worst["sh_rest_abs"] = max(worst["sh_rest_abs"], float(err.max()))
            ok = mx > 0
            worst["sh_rest_rel_to_max"] = max(worst["sh_rest_rel_to_max"],
                                              float((err[ok].max(axis=1) / mx[ok]).max()))
        del a, e
    return {k: float(f"{v:.4g}") for k, v in worst.items()}


def precision_leg(G, W, args):
    """The plain (fp32) and the chosen (compact: fp16 rotation / scale /
    opacity, fp8-E4M3 SH with a per-Gaussian scale; positions, projection
    and depth fp32) storage of the 10M SH3 scene on the same GPU: the same
    camera and RNG frames rendered `--precision-frames` times and
    accumulated progressively in each form; PSNR of the compact form's
    converged image against the fp32 one, beside the Monte-Carlo noise floor
    (two independent fp32 accumulations of the same length); frames/s of
    each form; the largest per-attribute storage error against float64."""
    w32 = W.WORKLOADS["c2"]
    w8 = W.replace(w32, fmt=G.FORMAT_HALF_FP8)
    M = args.precision_frames
    perm = W.order(w32)
    stream = G.Stream()
    out = {"workload": f"{w32.n} clustered SH3 Gaussians, {w32.width}x{w32.height} ss {w32.ss} spp {w32.spp}, orbit "
                       f"position 0, {M} frames accumulated (progressive), RNG frames 1..{M} in both forms",
           "frames": M}
    imgs = {}
    for name, w in (("fp32", w32), ("half_fp8", w8)):
        scene = W.device_scene(w, perm, log=lambda m: log("[precision] " + m))
        seqs = [range(1, M + 1)] + ([range(M + 1, 2 * M + 1)] if name == "fp32" else [])
        for k, frames in enumerate(seqs):
            loop = W.FrameLoop(w, scene, stream)
            loop.step(W.frame(w, 0, frames[0]))  # warm-up frame, restarted below
            loop.close()
            loop = W.FrameLoop(w, scene, stream)
            stream.synchronize()
            t0 = time.perf_counter()
            for fr in frames:
                loop.step(W.frame(w, 0, fr))
            stream.synchronize()
            dt = time.perf_counter() - t0
            imgs[(name, k)] = loop.output().astype(np.float64) / M
            loop.close()
            if k == 0:
                out[name] = {"fps": round(M / dt, 2), "device_gb": round(scene.device_bytes / 1e9, 3)}
        scene.close()
    out["psnr_db_half_fp8_vs_fp32"] = _psnr(imgs[("half_fp8", 0)], imgs[("fp32", 0)])
    out["psnr_db_noise_floor_fp32_vs_fp32"] = _psnr(imgs[("fp32", 1)], imgs[("fp32", 0)])
    out["half_fp8_over_fp32_fps"] = round(out["half_fp8"]["fps"] / out["fp32"]["fps"], 3)
    out["max_attribute_error_vs_fp64"] = attribute_errors(G, W, w32, perm[::5])
    out["attribute_sample"] = f"every 5th Gaussian of the scene ({len(perm[::5])})"
    log(f"[precision] {out}")
    return out


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
    batch.set_tuning(tuned["splat_threads"], tuned["splat_points_per_thread"], True,
                     tuned["preprocess_views_per_workgroup"])
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
    ap.add_argument("--no-group-cull", action="store_true",
                    help="time the path without the group occlusion test (same outputs)")
    ap.add_argument("--request-types", action="store_true",
                    help="measure only the GPU time per serving request type (one JSON line)")
    ap.add_argument("--request-scenes", default="",
                    help="with --request-types: comma-separated scene names (c1,c2,c3; default all)")
    ap.add_argument("--no-ordered", action="store_true",
                    help="time the path with the ordered splat off (chunk order; same outputs)")
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
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-views comparison and print its report (for a profiler: its "
                         "kernels per form)")
    ap.add_argument("--bvh-order", action="store_true",
                    help="measure only the load-time BVH order of the workload's scene (one JSON line)")
    ap.add_argument("--bvh-order-runs", type=int, default=3)
    ap.add_argument("--bvh-ref-sample", type=int, default=8_000_000,
                    help="Gaussians of the reference tinybvh CPU timing")
    ap.add_argument("--profile-window", action="store_true",
                    help="render the warm-up, sleep 1 s, render the timed frames, print the per-stage times and "
                         "exit: the run to put under rocprofv3 (the sleep marks the timed frames in its trace)")
    ap.add_argument("--summarize-trace", metavar="KERNEL_TRACE_CSV", action="append",
                    help="print per-kernel statistics of the timed frames of --profile-window traces and exit: "
                         "the first gives the times, each further one pairs with a --counters file")
    ap.add_argument("--counters", metavar="COUNTER_CSV", action="append", default=[],
                    help="with --summarize-trace: the counter CSV of the same rocprofv3 run as the trace it follows")
    ap.add_argument("--traffic-out", metavar="JSON",
                    help="with --summarize-trace: write the calibrated HBM bytes per stage launch here")
    ap.add_argument("--precision-frames", type=int, default=256,
                    help="frames accumulated per storage form in the precision comparison (0 = skip)")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the full-size parity check against the oracle (the CPU baseline then times the "
                         "oracle on a sample of the scene)")
    args = ap.parse_args()
    if args.summarize_trace:
        if len(args.counters) != len(args.summarize_trace) - 1:
            print("--summarize-trace: one trace for the times, then one per --counters file", file=sys.stderr)
            return 2
        print(summarize_trace(args.summarize_trace, args.counters, traffic_out=args.traffic_out,
                              window_name=f"{args.config} positions "
                                          f"{max(0, args.first_frame) if args.first_frame >= 0 else args.warmup}.."
                                          f"{(max(0, args.first_frame) if args.first_frame >= 0 else args.warmup) + (args.steps or 0) - 1}"))
        return 0
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
    if args.batch_only:
        print(json.dumps({"batched": batched(G, W, args)}), flush=True)
        return 0
    if args.request_types:
        print(json.dumps(request_types(G, W, args)), flush=True)
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
    loop.pass_.set_group_culling(not args.no_group_cull)
    loop.pass_.set_ordered_splat(G.Pass.ORDERED_OFF if args.no_ordered else G.Pass.ORDERED_AUTO)
    log(f"[bench] warmup {Wm} frames, then {K} timed frames (camera positions {F0 + Wm}..{F0 + Wm + K - 1} mod "
        f"{w.path_frames})")
    run(loop, range(F0, F0 + Wm))
    stream.synchronize()
    if args.profile_window:
        time.sleep(1.0)  # an idle gap that separates the timed frames in a profiler's trace
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
    if args.profile_window:
        # the counter calibration copy (1 GiB each way), inside the window
        buf_a = G.DeviceArray((COPY_CAL_BYTES,), np.uint8, zero=True)
        buf_b = G.DeviceArray((COPY_CAL_BYTES,), np.uint8)
        G.check(G.lib.gps_selftest_copy(buf_b.ptr, buf_a.ptr, COPY_CAL_BYTES, stream.handle), "copy")
        stream.synchronize()
        st_ = timing["stages"]
        print(json.dumps({"profile_window": f"{args.config} positions {F0 + Wm}..{F0 + Wm + K - 1}",
                          "ms_per_frame": round(max(gpu_ms, wall * 1e3) / K, 4),
                          "stages_us_per_launch": {k: round(v["ms"] * 1e3 / v["launches"], 1)
                                                   for k, v in st_.items() if v["launches"]},
                          "launches": {k: v["launches"] for k, v in st_.items() if v["launches"]}}), flush=True)
        return 0
    # Work counters: replay the same deterministic frame sequence on a fresh
    # pass (same initial state), reading the counters after every frame.
    loop = W.FrameLoop(w, scene, stream)
    loop.pass_.set_group_culling(not args.no_group_cull)
    loop.pass_.set_ordered_splat(G.Pass.ORDERED_OFF if args.no_ordered else G.Pass.ORDERED_AUTO)
    loop.pass_.set_counters(True, count_rejects=True)
    per_frame = []
    rejects = [dict.fromkeys(G.REJECT_REASONS, 0) for _ in range(2)]
    group_violations = 0
    ordered_skipped = 0
    run(loop, range(F0, F0 + Wm))
    for i in range(F0 + Wm, F0 + Wm + K):
        loop.step(W.frame(w, i % w.path_frames, i + 1))
        per_frame.append(loop.pass_.stats(loop.stream))
        for acc, r in zip(rejects, loop.pass_.rejects(loop.stream)):
            for key, v in r.items():
                acc[key] += v
        group_violations += sum(loop.pass_.group_violations(loop.stream))
        ordered_skipped += sum(loop.pass_.ordered_skipped(loop.stream))
    loop.close()
    reject_report = [{key: round(v / K) for key, v in acc.items()} for acc in rejects]
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
    # HBM traffic of the dominant kernel, from rocprofv3 counters of this same
    # window (bench.py --profile-window under rocprofv3; bench.py
    # --summarize-trace ... --traffic-out writes the file); null for a window
    # that was not profiled.
    traffic = None
    window = f"{args.config} positions {F0 + Wm}..{F0 + Wm + K - 1}"
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.config}_positions_{F0 + Wm}_{F0 + Wm + K - 1}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        if tj.get("_meta", {}).get("window") == window:
            traffic = tj.get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "window": window,
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
    threads = min(16, os.cpu_count() or 1)
    parity = None
    first_timed = F0 + Wm
    if not args.no_parity:
        in_order = tuple(range(min(2, w.path_frames)))
        carried = tuple(sorted({first_timed, w.path_frames // 2, w.path_frames - 1} - set(in_order))) \
            if w.path_frames > 2 else ()
        if perm is None:
            perm = W.order(w)
        log(f"[parity] {args.config} at full size: positions {in_order} in order, {carried} with carried state, "
            f"oracle on {threads} threads")
        parity = parity_leg(G, W, w, scene, perm, threads, in_order, carried)
    cpu = None
    timed_frames = [r for r in (parity or {}).get("frames", []) if r["position"] == first_timed]
    if timed_frames and not args.no_cpu_baseline:
        r = timed_frames[0]
        cpu = {"value": round(r["samples"] / r["oracle_s"] / 1e9, 5), "unit": "Gpoints/s", "cores": threads,
               "kind": "port",
               "sample": f"one frame of the timed window (camera position {first_timed}) on the full {w.n}-Gaussian "
                         f"scene, {w.width}x{w.height} ss{w.ss} spp{w.spp}, frame state carried from the GPU: "
                         f"{r['samples']} point draws in {r['oracle_s']:.2f} s (render pass only, no resolve)",
               "frame_s": r["oracle_s"]}
    elif not args.no_cpu_baseline:
        sample_n = args.cpu_sample or min(w.n, 2_000_000)
        log(f"[bench] cpu baseline on {sample_n} Gaussians, {threads} threads")
        cpu = cpu_baseline(G, W, w, sample_n, threads, perm)
    precision = None
    if args.precision_frames > 0:
        precision = precision_leg(G, W, args)
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
        "ordered_skipped_points_per_frame": round(ordered_skipped / K),
        "skipped_groups_per_frame": round(sum(sum(st.get("culled_groups", (0, 0))[:st["passes"]]) for st in per_frame) / K),
        "preprocess_rejects_per_frame": {"first_pass": reject_report[0], "second_pass": reject_report[1],
                                         "skipped_group_survivors": group_violations},
        "splat_forms": splat_forms,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "precision": precision,
        "batched": batch_report,
        "serving": serve,
        "stages": stage_report,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

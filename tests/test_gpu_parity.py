"""GPU parity: the MI355X kernels against the CPU oracle, through the C ABI.

Integer work (weights, valid masks, compacted offsets, per-point Gaussian
indices, framebuffer words) must be bit-exact; because the scalar math is the
shared deterministic header, projection results and final colour/depth images
are bit-exact too (tolerance 0 ulp, PSNR = inf).  Sizes are chosen so the
oracle finishes in seconds.
"""
import numpy as np
import pytest

import helpers as H
from helpers import assert_same, flush

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------------ scalar math
@pytest.mark.parametrize("kind,lo,hi", [(0, -110.0, 90.0), (1, 1e-38, 1e4), (2, -7.0, 7.0), (3, -7.0, 7.0),
                                        (4, 0.0, 1e6), (5, -1e3, 1e3), (6, 0.0, 1.0), (7, 0.0, 1.7),
                                        (8, -1e10, 1e10), (9, 0.0, 1.0), (10, 0.0, 1.0)])
def test_device_math_bitexact(gpu, kind, lo, hi):
    import orc
    flush(f"device math kind={kind}")
    rng = np.random.default_rng(kind)
    x = rng.uniform(lo, hi, 200_000).astype(np.float32)
    if kind == 1:
        x = np.exp(rng.uniform(np.log(1e-38), np.log(1e30), 200_000)).astype(np.float32)
    specials = np.array([0.0, -0.0, 1.0, 0.5, 2.0 ** -126, 2.0 ** -149, 1e-37, 0.99999994], np.float32)
    x[:specials.size] = np.clip(specials, lo, hi)
    d_in = gpu.DeviceArray(x.shape, np.float32)
    d_out = gpu.DeviceArray(x.shape, np.float32)
    d_in.upload(x)
    gpu.check(gpu.lib.gps_selftest_math(kind, d_in.ptr, d_out.ptr, x.size, None), "selftest")
    dev = d_out.download()
    host = orc.detmath(kind, x)
    assert_same(f"math kind {kind}", dev, host)


def test_device_log_quotient_exhaustive(gpu):
    """The log's quotient f / (2 + f) on the device (reciprocal and one fma
    correction) equals the IEEE-rounded division for every input it gets:
    f = m - 1, m a float in (sqrt(2)/2, sqrt(2)] from a mantissa (halved above
    sqrt(2)), i.e. all 2^23 mantissas."""
    k = np.arange(1 << 23, dtype=np.uint32)
    m = ((k | np.uint32(0x3F800000)).view(np.float32)).copy()
    big = m > np.float32(1.41421354)  # 0x1.6a09e6p+0
    m[big] = m[big] * np.float32(0.5)
    f = (m - np.float32(1.0)).astype(np.float32)
    host = (f / (np.float32(2.0) + f)).astype(np.float32)
    d_in = gpu.DeviceArray(f.shape, np.float32)
    d_out = gpu.DeviceArray(f.shape, np.float32)
    d_in.upload(f)
    flush("log quotient: 2^23 inputs")
    gpu.check(gpu.lib.gps_selftest_math(12, d_in.ptr, d_out.ptr, f.size, None), "selftest")
    assert_same("log quotient", d_out.download(), host)


def test_device_sqrt_range_exhaustive(gpu):
    """The sampling transform's square root (hardware root and the +-1-ulp
    fma correction, no tiny-input scaling) equals det_sqrt's IEEE-rounded
    root for every float in [1, 4) -- every mantissa at both exponent
    parities -- and for 0, a sample of floats from 2^-96 to 2^128, the 4096
    largest floats, infinity, and NaN for negative and NaN inputs."""
    k = np.arange(1 << 24, dtype=np.uint32)
    x = ((k + np.uint32(0x3F800000)).view(np.float32)).copy()  # [1, 4)
    rng = np.random.default_rng(13)
    sample = np.exp(rng.uniform(np.log(2.0 ** -96), np.log(2.0 ** 127.99), 1 << 20)).astype(np.float32)
    big = (np.float32(3.4028235e38).view(np.uint32) - np.arange(4096, dtype=np.uint32)).view(np.float32)
    special = np.array([0.0, 2.0 ** -96, 170.5, np.inf, -1.0, -np.inf, np.nan], np.float32)
    x = np.concatenate([x, sample, big, special])
    d_in = gpu.DeviceArray(x.shape, np.float32)
    d_fast = gpu.DeviceArray(x.shape, np.float32)
    d_ref = gpu.DeviceArray(x.shape, np.float32)
    d_in.upload(x)
    flush("sqrt: 2^24 + 2^20 inputs")
    gpu.check(gpu.lib.gps_selftest_math(13, d_in.ptr, d_fast.ptr, x.size, None), "selftest")
    gpu.check(gpu.lib.gps_selftest_math(4, d_in.ptr, d_ref.ptr, x.size, None), "selftest")
    ref = d_ref.download()
    fast = d_fast.download()
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(fast), nan), "NaN where det_sqrt is not, or the reverse"
    assert nan.sum() == 3  # -1, -inf, NaN
    assert_same("sqrt vs device det_sqrt", fast[~nan], ref[~nan])
    ok = np.isfinite(x) & (x >= 0)
    assert_same("device det_sqrt vs numpy", ref[ok], np.sqrt(x[ok]))


def test_device_rng_bitexact(gpu):
    import orc
    n, draws, frame = 4096, 6, 4 * 7 + 1
    d = gpu.DeviceArray((n * (2 + 2 * draws),), np.uint32)
    gpu.check(gpu.lib.gps_selftest_rng(n, frame, draws, d.ptr, None), "rng")
    dev = d.download().reshape(n, -1)
    host = np.empty_like(dev)
    for i in range(n):
        orc.lib.orc_rng_probe(i, frame, draws, host[i].ctypes.data_as(orc.C.c_void_p))
    assert_same("rng", dev, host)


def test_device_poisson_bitexact(gpu):
    import orc
    n = 20000
    lam = np.exp(np.random.default_rng(3).uniform(np.log(1e-6), np.log(1e6), n)).astype(np.float32)
    d_l = gpu.DeviceArray((n,), np.float32)
    d_o = gpu.DeviceArray((n,), np.uint32)
    d_l.upload(lam)
    gpu.check(gpu.lib.gps_selftest_poisson(n, d_l.ptr, d_o.ptr, None), "poisson")
    dev = d_o.download()
    seeds = np.empty(2, np.uint32)
    host = np.empty(n, np.uint32)
    for i in range(n):
        orc.lib.orc_rng_probe(i, 12345, 0, seeds.ctypes.data_as(orc.C.c_void_p))
        host[i] = orc.lib.orc_poisson(int(seeds[0]), int(seeds[1]), float(lam[i]))
    assert_same("poisson", dev, host)


# ------------------------------------------------------------------ one frame
def expected_records(o_proj: np.ndarray, opac: np.ndarray, near: float, far: float) -> np.ndarray:
    """Splat records rebuilt in float32 numpy from the oracle's FP32
    projection (pixel mean, dilated cov2d, colour, view z, bounds)."""
    import orc
    f32 = np.float32
    a, b, c = o_proj[:, 4].astype(f32), o_proj[:, 5].astype(f32), o_proj[:, 6].astype(f32)
    c0 = np.sqrt(a)
    c1 = b / c0
    c2 = np.sqrt(c - c1 * c1)
    det = a * c - b * b
    inv_det = f32(1.0) / det
    op = np.minimum(opac.astype(f32), f32(1.0))
    rec = np.zeros((o_proj.shape[0], 16), np.uint32)
    cols = [o_proj[:, 2].astype(f32), o_proj[:, 3].astype(f32), c0, c1, c2, c * inv_det, (-b) * inv_det,
            a * inv_det, op, f32(1.0) / np.maximum(op, f32(1e-37)), orc.detmath(6, op)]
    for k, v in enumerate(cols):
        rec[:, k] = np.asarray(v, f32).view(np.uint32)
    rec[:, 11] = o_proj[:, 12].astype(np.uint32) | (o_proj[:, 13].astype(np.uint32) << 16)
    rec[:, 12] = o_proj[:, 14].astype(np.uint32) | (o_proj[:, 15].astype(np.uint32) << 16)
    vz = o_proj[:, 1].astype(f32)
    depth = np.array([orc.lib.orc_depth_bits(float(v), near, far) for v in vz], np.uint64)
    col = o_proj[:, 7:10].astype(f32)
    with np.errstate(invalid="ignore"):
        q = (col * f32(255.0)).astype(np.uint64) & np.uint64(0xFFF)
    packed = (depth << np.uint64(36)) | (q[:, 0] << np.uint64(24)) | (q[:, 1] << np.uint64(12)) | q[:, 2]
    rec[:, 13] = (packed & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    rec[:, 14] = (packed >> np.uint64(32)).astype(np.uint32)
    return rec


def run_pair(arrays, settings, frames, fmt=0, which=H.INTERMEDIATES, check_records=True):
    g = H.GpuRunner(arrays, settings, fmt)
    o = H.OracleRunner(arrays, settings, fmt)
    out = []
    for fi, f in enumerate(frames):
        flush(f"frame {fi}: render gpu")
        rg = g.render(f, which)
        flush(f"frame {fi}: render oracle")
        ro = o.render(f, which)
        if not f.use_unbiased_2d_splatting:
            assert ro.stats["biased_events"] == {"zero_colour": 0, "wrapped": 0}, ro.stats["biased_events"]
        for k in which:
            assert_same(f"frame {fi} debug[{k}]", rg.debug[k], ro.debug[k])
        assert_same(f"frame {fi} color", rg.color, ro.color)
        assert_same(f"frame {fi} depth", rg.depth, ro.depth)
        for key in ("valid_gaussians", "total_weight", "num_points", "samples", "passes", "budget_truncated"):
            assert rg.stats[key] == ro.stats[key], (key, rg.stats[key], ro.stats[key])
        if check_records and not f.enable_occlusion_culling and fmt == 0:
            ids = rg.debug.get(H.G.DBG_VALID_IDS)
            if ids is not None and ids.size:
                rec = g.p.debug(H.G.DBG_RECORDS).reshape(-1, 16)[ids]
                proj = o.p.project(f, fp64=False)[ids]
                exp = expected_records(proj, arrays.opacities[ids], f.near_plane, f.far_plane)
                assert_same(f"frame {fi} records", rec[:, :15], exp[:, :15])
        out.append((rg, ro))
    return out


def c1_frame(w=640, h=360, **kw):
    return H.G.frame_params((0.0, 0.0, -30.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, w, h, **kw)


def test_frame_c1_small(gpu):
    arrays = H.scene(0x5EED0001, 60_000, H.C1)
    st = gpu.RenderSettings(640, 360, 1, 1, fully_disable_hierarchical_culling=True)
    (rg, ro), = run_pair(arrays, st, [c1_frame(ss=1)])
    assert rg.stats["num_points"][0] > 10_000
    assert H.psnr(rg.color, ro.color) == float("inf")


@pytest.mark.parametrize("case", ["one_word", "grid64"])
def test_frame_splat_word_table_contention(gpu, case):
    """Point-parallel splat under heavy merging in its LDS word table: every
    Gaussian on one sub-pixel spot (all lanes race for one slot, the
    per-thread repeat skip), or on a coarse grid of 64 spots at two depths
    (slot collisions, the neighbour probe and the framebuffer fallback)."""
    flush(f"word table {case}")
    arrays = H.scene(0x5EED0300 + len(case), 30_000, H.C1)
    n = arrays.n
    if case == "one_word":
        arrays.means[:] = 0.0
        arrays.scales[:] = 0.002
    else:
        k = np.arange(n)
        arrays.means[:, 0] = ((k % 8) - 3.5) * 2.0
        arrays.means[:, 1] = (((k // 8) % 8) - 3.5) * 2.0
        arrays.means[:, 2] = np.where(k % 2 == 0, 0.0, 4.0)
        arrays.scales[:] = 0.05
    arrays.opacities[:] = 0.95
    st = gpu.RenderSettings(320, 200, 1, 2, fully_disable_hierarchical_culling=True)
    (rg, ro), = run_pair(arrays, st, [c1_frame(320, 200, ss=2)])
    assert rg.stats["num_points"][0] > 1000 and rg.stats["splat_binned"] == 0


def test_frame_c2_small_sh3_ss2_spp4(gpu):
    arrays = H.scene(0x5EED0002, 40_000, H.C2, clusters=200)
    st = gpu.RenderSettings(480, 270, 4, 2, fully_disable_hierarchical_culling=True)
    f = H.G.frame_params((0.0, 120.0, -330.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 480, 270, ss=2)
    (rg, ro), = run_pair(arrays, st, [f])
    assert rg.stats["K"] == 4 and rg.stats["num_points"][0] > 0


@pytest.mark.parametrize("case", ["empty", "one", "ragged37", "behind", "budget", "biased", "one_point",
                                  "cull_small", "ss3_spp3", "background", "odd_res"])
def test_frame_edge_cases(gpu, case):
    flush(f"edge case {case}")
    n = {"empty": 0, "one": 1, "ragged37": 37}.get(case, 20_000)
    arrays = H.scene(0x5EED0100 + len(case), n, H.C1) if n else H.G.SceneArrays(
        np.zeros((0, 3), np.float32), np.zeros((0, 4), np.float32), np.zeros((0, 3), np.float32),
        np.zeros((0,), np.float32), np.zeros((0, 3), np.float32), 0)
    if case == "one":
        arrays.means[:] = 0.0
        arrays.scales[:] = 0.5
        arrays.opacities[:] = 0.9
    w, h, ss, spp = 320, 200, 1, 1
    if case == "ss3_spp3":
        ss, spp = 3, 3
    if case == "odd_res":
        w, h, ss = 237, 131, 2
    st = gpu.RenderSettings(w, h, spp, ss, fully_disable_hierarchical_culling=True,
                            max_number_of_points=1000 if case == "budget" else 250_000_000)
    kw = dict(ss=ss)
    if case == "biased":
        kw["unbiased"] = False
    if case == "one_point":
        kw["reduce_point_count"] = False
    if case == "cull_small":
        kw["cull_small"] = True
    if case == "background":
        kw["background"] = (0.25, 0.5, 1.0)
    eye, ctr = ((0.0, 0.0, -30.0), (0.0, 0.0, 0.0)) if case != "behind" else ((0.0, 0.0, -30.0), (0.0, 0.0, -60.0))
    f = H.G.frame_params(eye, ctr, (0.0, 1.0, 0.0), 45.0, w, h, **kw)
    (rg, ro), = run_pair(arrays, st, [f])
    if case == "budget":
        assert rg.stats["budget_truncated"] == 1 and rg.stats["num_points"][0] == 1000
    if case in ("empty", "behind"):
        assert rg.stats["num_points"][0] == 0
    if case == "biased":
        # equal-depth samples added their colours: some word's colour exceeds
        # every single Gaussian's (the min-splat would never do that)
        fb = rg.debug[H.G.DBG_FRAMEBUFFER]
        assert rg.stats["splat_binned"] == 0
        assert np.count_nonzero(fb != ro.debug[H.G.DBG_FRAMEBUFFER]) == 0


@pytest.mark.parametrize("ss,spp,occl", [(1, 1, False), (2, 4, False), (2, 1, True)])
def test_frame_biased_splat(gpu, ss, spp, occl):
    """Biased splat (use_unbiased_2d_splatting = 0): the reference's
    pre-check + atomicAddColor per sample.  Large Gaussians with many samples
    per pixel so equal-depth adds are common; the oracle reports no
    zero-colour sample and no colour carry out of bit 35, so every word is
    independent of sample order and the framebuffers must match bit for bit.
    Also with 4 samples per point per layer (where the unbiased splat would
    take the binned form) and with two-pass occlusion + block culling."""
    flush(f"biased ss={ss} spp={spp} occl={occl}")
    arrays = H.scene(0x5EED0B00 + 10 * ss + spp, 8_000, dict(H.C1, log_scale_mu=-1.5, log_scale_sigma=0.3))
    arrays.opacities[:] = np.clip(arrays.opacities, 0.3, 0.95)
    st = gpu.RenderSettings(320, 200, spp, ss, fully_disable_hierarchical_culling=not occl, gaussians_per_block=64)
    frames = moving_frames(2, 320, 200, ss, unbiased=False, occlusion=occl, hierarchical=occl)
    which = H.INTERMEDIATES + ([H.G.DBG_HIZ] if occl else [])
    res = run_pair(arrays, st, frames, which=which)
    for rg, ro in res:
        assert rg.stats["splat_binned"] == 0
        fb = rg.debug[H.G.DBG_FRAMEBUFFER]
        bg = np.uint64(((1 << 28) - 1) << 36)
        hit = fb[fb != bg]
        assert hit.size > 1000
        # some word holds more than one sample's colour: its 12-bit red
        # channel exceeds 255 (one sample's colour is at most ~1.5 * 255 here)
        assert int(((hit >> np.uint64(24)) & np.uint64(0xFFF)).max()) > 600


def moving_frames(n_frames, w, h, ss, radius=40.0, **kw):
    frames = []
    for i in range(n_frames):
        a = 0.15 * i
        eye = (radius * np.sin(a), 3.0, -radius * np.cos(a))
        f = H.G.frame_params(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, w, h, ss=ss, frame=i + 1, **kw)
        frames.append(f)
    return frames


def test_frames_occlusion_and_block_culling(gpu):
    # Two overlapping slabs so the Hi-Z culls something; Morton-free index order.
    a = H.scene(0x5EED0201, 30_000, dict(H.C1, half_extent=6.0, log_scale_mu=-2.5))
    b = H.scene(0x5EED0202, 30_000, dict(H.C1, half_extent=6.0, log_scale_mu=-2.5), offset=(0.0, 0.0, 20.0))
    arrays = H.G.SceneArrays(np.concatenate([a.means, b.means]), np.concatenate([a.rotations, b.rotations]),
                             np.concatenate([a.scales, b.scales]), np.concatenate([a.opacities, b.opacities]),
                             np.concatenate([a.shs, b.shs]), 0)
    st = gpu.RenderSettings(320, 200, 1, 2, gaussians_per_block=64)
    frames = moving_frames(4, 320, 200, 2, occlusion=True, hierarchical=True)
    which = H.INTERMEDIATES + [H.G.DBG_MASK_PREVIOUS, H.G.DBG_HIZ, H.G.DBG_CULL_MASK]
    res = run_pair(arrays, st, frames, which=which)
    culled = sum(r[0].stats["culled_blocks"][0] for r in res)
    second = sum(r[0].stats["num_points"][1] for r in res)
    flush(f"culled blocks {culled}, second-pass points {second}")


def test_frames_half_fp8(gpu):
    arrays = H.scene(0x5EED0301, 30_000, H.C2, clusters=100)
    st = gpu.RenderSettings(320, 200, 2, 2, gaussians_per_block=128)
    frames = moving_frames(3, 320, 200, 2, radius=300.0, occlusion=True, hierarchical=True)
    run_pair(arrays, st, frames, fmt=H.G.FORMAT_HALF_FP8,
             which=H.INTERMEDIATES + [H.G.DBG_HIZ, H.G.DBG_CULL_MASK])


@pytest.mark.parametrize("case", ["sh3_two_pass_culling", "sh0_plain"])
def test_frames_quantized(gpu, case):
    """The reference's compressed scene (its default build): on-the-fly decode
    of position / rotation / scale / opacity bytes, the Sloan-basis colour of
    the SH bytes, and the whole frame bit for bit against the oracle, which
    decodes the same bytes with its own restatement."""
    if case == "sh0_plain":
        arrays = H.scene(0x5EED0312, 30_000, H.C1)
        st = gpu.RenderSettings(320, 200, 1, 1, fully_disable_hierarchical_culling=True)
        frames = [c1_frame(320, 200)]
        which = H.INTERMEDIATES
    else:
        arrays = H.scene(0x5EED0311, 30_000, H.C2, clusters=100)
        st = gpu.RenderSettings(320, 200, 2, 2, gaussians_per_block=128)
        frames = moving_frames(3, 320, 200, 2, radius=300.0, occlusion=True, hierarchical=True)
        which = H.INTERMEDIATES + [H.G.DBG_HIZ, H.G.DBG_CULL_MASK]
    res = run_pair(arrays, st, frames, fmt=H.G.FORMAT_QUANTIZED, which=which)
    assert res[0][0].stats["num_points"][0] > 1000


def test_c3_construction_streamed_culling_taa(gpu):
    """The 100M-Gaussian workload's construction at test size: 4 translated
    copies of a clustered SH3 scene, Morton load order over the whole scene,
    compact fp16/fp8 storage uploaded in several ranges, blocks + Hi-Z
    culling and TAA along the fly-through (in order, then two camera cuts).
    Every frame's framebuffer, colour, depth and TAA output bit-exact."""
    import orc
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 30_000, width=320, height=180, gaussians_per_block=64)
    perm = W.order(w)
    arrays = W.host_scene(w, perm)
    scene = W.device_scene(w, perm, chunk=25_000)  # 5 upload ranges, last one ragged
    loop = W.FrameLoop(w, scene)
    o = orc.OraclePass(w.settings(), arrays, w.fmt, 8)
    h_color = [np.zeros((w.height, w.width, 3), np.float32) for _ in range(2)]
    h_depth = [np.zeros((w.height, w.width), np.float32) for _ in range(2)]
    h_hist = [np.zeros((w.height, w.width, 3), np.float32) for _ in range(2)]
    prev_vp = np.zeros(16, np.float32)
    culled = second = 0
    for k, pos in enumerate([0, 1, 2, 3, 150, 151, 299]):
        f = W.frame(w, pos, k + 1)
        flush(f"c3 tiny frame {k} (camera {pos})")
        loop.step(f)
        fg = loop.pass_.debug(gpu.DBG_FRAMEBUFFER, loop.stream)
        sg = loop.pass_.stats(loop.stream)
        c, d = o.render(f)
        a, b = (k + 1) % 2, k % 2
        h_color[a], h_depth[b] = c, d
        M = gpu.taa_reprojection(prev_vp, np.array(f.view, np.float32), np.array(f.proj, np.float32))
        h_hist[a], rgb = orc.resolve_taa(h_color[a], h_depth[a], h_hist[b], h_depth[b], M, 0.9, w.width, w.height,
                                         k + 1)
        prev_vp = gpu.mat4_mul(np.array(f.proj, np.float32), np.array(f.view, np.float32))
        so = o.stats()
        for key in ("valid_gaussians", "num_points", "culled_blocks", "passes"):
            assert sg[key] == so[key], (k, key, sg[key], so[key])
        assert_same(f"c3 fb {k}", fg, o.debug(gpu.DBG_FRAMEBUFFER))
        assert_same(f"c3 color {k}", loop.color[a].download(), c)
        assert_same(f"c3 depth {k}", loop.depth[b].download(), d)
        assert_same(f"c3 taa {k}", loop.hist[a].download(), h_hist[a])
        assert_same(f"c3 rgb {k}", loop.rgb.download(), rgb)
        culled += sum(sg["culled_blocks"][:sg["passes"]])
        second += sg["num_points"][1] if sg["passes"] == 2 else 0
    flush(f"c3 tiny: culled blocks {culled}, second-pass points {second}")
    assert culled > 0


# ------------------------------------------------------------------ resolve
def test_resolve_taa_and_progressive(gpu):
    import orc
    w, h = 160, 96
    rng = np.random.default_rng(7)
    cur = rng.uniform(-0.1, 1.2, (h, w, 3)).astype(np.float32)
    prev = rng.uniform(0.0, 1.0, (h, w, 3)).astype(np.float32)
    cdep = rng.uniform(0.9, 1.0, (h, w)).astype(np.float32)
    pdep = rng.uniform(0.9, 1.0, (h, w)).astype(np.float32)
    view0 = gpu.look_at((0.0, 0.0, -30.0), (0.0, 0.0, 0.0))
    view1 = gpu.look_at((0.5, 0.2, -29.0), (0.0, 0.0, 0.0))
    proj = gpu.perspective(45.0, w / h, 0.2, 1e4)
    M = gpu.taa_reprojection(gpu.mat4_mul(proj, view0), view1, proj)
    dc, dd, dp, dpd = (gpu.DeviceArray(a.shape, np.float32) for a in (cur, cdep, prev, pdep))
    for d_, a_ in ((dc, cur), (dd, cdep), (dp, prev), (dpd, pdep)):
        d_.upload(a_)
    out = gpu.DeviceArray(cur.shape, np.float32)
    rgb = gpu.DeviceArray(cur.shape, np.uint8)
    for fn in (0, 1, 5):
        gpu.resolve_taa(dc, dd, dp, dpd, M, 0.9, w, h, fn, out, rgb)
        oc, orgb = orc.resolve_taa(cur, cdep, prev, pdep, M, 0.9, w, h, fn)
        assert_same(f"taa color {fn}", out.download(), oc)
        assert_same(f"taa rgb {fn}", rgb.download(), orgb)
    for n in (1, 2, 7):
        dc.upload(cur)
        dp.upload(prev)
        gpu.resolve_progressive(dc, dp, w, h, n, rgb)
        c2, a2, r2 = orc.resolve_progressive(cur, prev, w, h, n)
        assert_same(f"prog color {n}", dc.download(), c2)
        assert_same(f"prog accum {n}", dp.download(), a2)
        assert_same(f"prog rgb {n}", rgb.download(), r2)


# ------------------------------------------------------------------ serving
def test_flexible_pass_matches_exact_pass_and_oracle(gpu):
    """A pass created at 320x200 spp 4 renders a 160x96 spp 1 frame exactly
    like a pass (and the oracle) created at that size."""
    import orc
    arrays = H.scene(0x5EED0501, 20_000, H.C2, clusters=40)
    big = gpu.RenderSettings(320, 200, 4, 2, flexible_frame_size=True)
    small = gpu.RenderSettings(160, 96, 1, 2)
    f = gpu.frame_params((0, 80, -300), (0, 0, 0), (0, 1, 0), 45.0, 160, 96, ss=2, frame=3, hierarchical=True,
                         spp=1)
    scene = gpu.Scene(arrays)
    flex = gpu.Pass(big, scene)
    color = gpu.DeviceArray((96, 160, 3), np.float32, zero=True)
    depth = gpu.DeviceArray((96, 160), np.float32, zero=True)
    flex.render(f, color, depth)
    fb_flex = flex.debug(gpu.DBG_FRAMEBUFFER)
    rs = H.GpuRunner(arrays, small).render(f, [gpu.DBG_FRAMEBUFFER])
    o = orc.OraclePass(small, arrays, 0, 8)
    c, d = o.render(f)
    assert_same("flex fb", fb_flex, rs.debug[gpu.DBG_FRAMEBUFFER])
    assert_same("flex fb vs oracle", fb_flex, o.debug(gpu.DBG_FRAMEBUFFER))
    assert_same("flex color", color.download(), c)
    assert_same("flex depth", depth.download(), d)
    # frames larger than the pass are refused
    with pytest.raises(gpu.GpsError):
        flex.render(gpu.frame_params((0, 0, -30), (0, 0, 0), (0, 1, 0), 45.0, 400, 100, ss=2), color, depth)


def test_serving_results_identical_alone_and_in_the_mix(gpu):
    """Requests of three resident scenes with mixed size / spp, submitted
    together to a 4-slot server, give the same images (colour floats and
    RGB8, bit for bit) as the same requests rendered one at a time on a
    1-slot server, and the same colour as the oracle."""
    import orc
    from gps_amd import workloads as W
    scenes = (W.scaled(W.WORKLOADS["c1"], 20_000), W.scaled(W.WORKLOADS["c2"], 30_000),
              W.scaled(W.WORKLOADS["c3"], 10_000))
    sw = W.ServeWorkload(scenes=scenes, radii=(30.0, 350.0, 600.0),
                         centers=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (375.0, 0.0, 0.0)),
                         resolutions=((128, 72), (200, 120), (320, 180)), spps=(1, 4))
    arrays = [W.host_scene(w) for w in scenes]
    dev = [gpu.Scene(a, w.fmt) for a, w in zip(arrays, scenes)]
    reqs = W.serve_requests(sw, count=24, rate=0.0)

    def run(slots, one_at_a_time):
        srv = gpu.Server(slots=slots, max_width=320, max_height=180, max_spp=4, max_supersampling=2)
        ids = [srv.add_scene(s) for s in dev]
        srv.start()
        outs = []
        for k, f, _ in reqs:
            rgb = np.zeros((f.height, f.width, 3), np.uint8)
            col = np.zeros((f.height, f.width, 3), np.float32)
            srv.submit(ids[k], f, -1.0, rgb, col)
            outs.append((rgb, col))
            if one_at_a_time:
                srv.drain(120_000)
        srv.drain(120_000)
        res = srv.poll(1024)
        st = srv.stats()
        srv.close()
        assert len(res) == len(reqs) and all(r["status"] == 0 for r in res)
        return outs, res, st

    flush("serving: 4 slots, all requests at once")
    mixed, res_m, st_m = run(4, False)
    flush(f"serving stats {st_m}")
    flush("serving: 1 slot, one at a time")
    alone, _, _ = run(1, True)
    assert len({r["slot"] for r in res_m}) > 1  # the mix really used several slots
    for i, ((a_rgb, a_col), (m_rgb, m_col)) in enumerate(zip(alone, mixed)):
        assert_same(f"request {i} colour", m_col, a_col)
        assert_same(f"request {i} rgb", m_rgb, a_rgb)
    # every fourth request against the oracle at its exact size
    for i in range(0, len(reqs), 4):
        k, f, _ = reqs[i]
        st = gpu.RenderSettings(f.width, f.height, f.samples_per_pixel, 2)
        o = orc.OraclePass(st, arrays[k], scenes[k].fmt, 8)
        c, _ = o.render(f)
        assert_same(f"request {i} vs oracle", mixed[i][1], c)


# ------------------------------------------------------------------ batches
def test_batch_views_identical_to_single_renders(gpu):
    """A batch of views of uneven size / spp / ss (one without block culling)
    gives every view's colour and depth bit for bit as a pass rendering that
    frame alone, for every splat variant, eagerly and replayed from a HIP
    graph with new cameras (each replay checked, eager work in between)."""
    import orc
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c2"], 30_000, fmt=gpu.FORMAT_HALF_FP8)
    arrays = W.host_scene(w)
    scene = gpu.Scene(arrays, w.fmt)
    bw = W.BatchWorkload(scene=w, views=6, radius=150.0, resolutions=((96, 64), (160, 90), (200, 120)),
                         spps=(1, 2, 4), ss=2)
    stream = gpu.Stream()
    batch = gpu.Batch(scene, 6, 200, 120, max_spp=4, max_supersampling=2)
    single = gpu.Pass(gpu.RenderSettings(200, 120, 4, 2, flexible_frame_size=True), scene)

    def outputs(frames):
        cols = [gpu.DeviceArray((f.height, f.width, 3), np.float32, zero=True) for f in frames]
        deps = [gpu.DeviceArray((f.height, f.width), np.float32, zero=True) for f in frames]
        return cols, deps

    def alone(frames):
        ref = []
        for f in frames:
            c = gpu.DeviceArray((f.height, f.width, 3), np.float32)
            d = gpu.DeviceArray((f.height, f.width), np.float32)
            single.render(f, c, d, stream)
            stream.synchronize()
            ref.append((c.download(), d.download(), single.stats(stream)))
        return ref

    def check_batch(frames, cols, deps, ref, what):
        stream.synchronize()
        for v, f in enumerate(frames):
            assert_same(f"{what} view {v} colour", cols[v].download(), ref[v][0])
            assert_same(f"{what} view {v} depth", deps[v].download(), ref[v][1])
            st = batch.view_stats(v, stream)
            for key in ("valid_gaussians", "num_points", "total_weight"):
                assert st[key][0] == ref[v][2][key][0], (what, v, key)

    frames = W.batch_views(bw, 0)
    frames[1].enable_hierarchical_culling = 0
    frames[2].supersampling_factor = 1
    frames[4].use_unbiased_2d_splatting = 0  # biased splat (atomicAddColor) in the mix
    ref = alone(frames)
    cols, deps = outputs(frames)
    for t, q in ((256, 4), (256, 2), (256, 8), (128, 4), (512, 2), (512, 4)):
        flush(f"batch eager, splat variant {t}x{q}")
        batch.set_tuning(t, q, False)
        batch.render(frames, cols, deps, stream)
        check_batch(frames, cols, deps, ref, f"eager {t}x{q}")
    # preprocess splits: views per workgroup (grid rows over the views; 5
    # leaves a last row of one view, 16 covers all six in one row)
    for vpw in (1, 4, 5, 16):
        flush(f"batch eager, preprocess views per workgroup {vpw}")
        batch.set_tuning(256, 4, False, vpw)
        batch.render(frames, cols, deps, stream)
        check_batch(frames, cols, deps, ref, f"eager vpw {vpw}")
    # oracle for three views at their exact sizes (view 4: biased)
    for v in (0, 3, 4):
        f = frames[v]
        o = orc.OraclePass(gpu.RenderSettings(f.width, f.height, f.samples_per_pixel, f.supersampling_factor),
                           arrays, w.fmt, 8)
        c, _ = o.render(f)
        assert o.biased_events() == {"zero_colour": 0, "wrapped": 0}
        assert_same(f"view {v} vs oracle", ref[v][0], c)
    tuned = batch.autotune(frames, cols, deps, stream, repeats=1)
    flush(f"autotune: {tuned}")
    assert len(tuned["preprocess_variants"]) == 3
    batch.set_tuning(tuned["splat_threads"], tuned["splat_points_per_thread"], True,
                     tuned["preprocess_views_per_workgroup"])
    for k in range(3):
        fr = W.batch_views(bw, k + 1)
        fr[1].enable_hierarchical_culling = 0
        fr[2].supersampling_factor = 1
        flush(f"batch graph replay {k}")
        batch.render(fr, cols, deps, stream)
        stream.synchronize()
        r = alone(fr)  # eager single-view work between replays
        check_batch(fr, cols, deps, r, f"graph {k}")


def test_batch_tile_scan_across_workgroups(gpu):
    """64 views of a 1.1M-Gaussian scene concatenate to about 17 200 tiles, so
    the tile scan runs three workgroups whose carries come from the steps
    before them; every view must match a pass rendering it alone (one view is
    269 tiles: one workgroup, no carry)."""
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c2"], 1_100_000, fmt=gpu.FORMAT_HALF_FP8)
    arrays = W.host_scene(w)
    scene = gpu.Scene(arrays, w.fmt)
    nv = 64
    bw = W.BatchWorkload(scene=w, views=nv, radius=150.0, resolutions=((64, 48), (96, 64)), spps=(1, 2), ss=2)
    stream = gpu.Stream()
    batch = gpu.Batch(scene, nv, 96, 64, max_spp=2, max_supersampling=2)
    single = gpu.Pass(gpu.RenderSettings(96, 64, 2, 2, flexible_frame_size=True), scene)
    frames = W.batch_views(bw, 0)
    cols = [gpu.DeviceArray((f.height, f.width, 3), np.float32, zero=True) for f in frames]
    deps = [gpu.DeviceArray((f.height, f.width), np.float32, zero=True) for f in frames]
    flush(f"batch of {nv} views, {w.n} Gaussians")
    batch.render(frames, cols, deps, stream)
    stream.synchronize()
    total = 0
    for v, f in enumerate(frames):
        c = gpu.DeviceArray((f.height, f.width, 3), np.float32)
        d = gpu.DeviceArray((f.height, f.width), np.float32)
        single.render(f, c, d, stream)
        stream.synchronize()
        ref = single.stats(stream)
        assert_same(f"view {v} colour", cols[v].download(), c.download())
        assert_same(f"view {v} depth", deps[v].download(), d.download())
        st = batch.view_stats(v, stream)
        for key in ("valid_gaussians", "num_points", "total_weight"):
            assert st[key][0] == ref[key][0], (v, key)
        total += ref["num_points"][0]
    assert total > 0


@pytest.mark.parametrize("case", ["mixed", "closeup_big", "budget", "quantized"])
def test_batch_binned_views_identical(gpu, case):
    """Binned batch views (the regions of all binned views in one job queue,
    counting and scattering per view, jobs that find their view by region)
    give every view's colour and depth bit for bit as a pass rendering the
    frame alone, with the binned form forced, automatic (K * spp >= 8, at
    most 8 samples per framebuffer word) and off; with Gaussians too large to bin and with a point budget that cuts
    views short."""
    cfg = dict(H.C1, half_extent=6.0, log_scale_mu=-2.0) if case != "closeup_big" else dict(H.C1, log_scale_mu=-0.7)
    arrays = H.scene(0x5EED0701 + len(case), 30_000, cfg).morton_sorted()
    scene = gpu.Scene(arrays, gpu.FORMAT_QUANTIZED if case == "quantized" else gpu.FORMAT_FP32)
    budget = 6_000 if case == "budget" else 250_000_000
    stream = gpu.Stream()
    batch = gpu.Batch(scene, 5, 200, 120, max_spp=4, max_supersampling=2, gaussians_per_block=64,
                      max_number_of_points=budget)
    single = gpu.Pass(gpu.RenderSettings(200, 120, 4, 2, flexible_frame_size=True, max_number_of_points=budget,
                                         gaussians_per_block=64), scene)
    # (width, height, spp, ss): K * spp = 4, 4, 16, 3, 8
    shapes = [(200, 120, 1, 2), (160, 90, 4, 1), (200, 120, 4, 2), (96, 64, 3, 1), (150, 100, 2, 2)]
    eye = (0.0, 2.0, -9.0) if case == "closeup_big" else (0.0, 3.0, -25.0)
    frames = [gpu.frame_params((eye[0] + 0.7 * v, eye[1] - 0.3 * v, eye[2]), (0, 0, 0), (0, 1, 0), 45.0, w, h,
                               ss=ss, frame=v + 3, hierarchical=(v != 1), spp=spp)
              for v, (w, h, spp, ss) in enumerate(shapes)]
    ref = []
    for f in frames:
        c = gpu.DeviceArray((f.height, f.width, 3), np.float32)
        d = gpu.DeviceArray((f.height, f.width), np.float32)
        single.render(f, c, d, stream)
        stream.synchronize()
        ref.append((c.download(), d.download(), single.stats(stream)))
    cols = [gpu.DeviceArray((f.height, f.width, 3), np.float32, zero=True) for f in frames]
    deps = [gpu.DeviceArray((f.height, f.width), np.float32, zero=True) for f in frames]
    for mode, name in ((gpu.Pass.SPLAT_BINNED, "binned"), (gpu.Pass.SPLAT_AUTO, "auto"),
                       (gpu.Pass.SPLAT_POINT_PARALLEL, "point-parallel")):
        flush(f"batch binned {case}: {name}")
        batch.set_splat_mode(mode)
        batch.render(frames, cols, deps, stream)
        stream.synchronize()
        for v, (f, (w, h, spp, ss)) in enumerate(zip(frames, shapes)):
            assert_same(f"{case} {name} view {v} colour", cols[v].download(), ref[v][0])
            assert_same(f"{case} {name} view {v} depth", deps[v].download(), ref[v][1])
            st = batch.view_stats(v, stream)
            assert st["num_points"][0] == ref[v][2]["num_points"][0], (case, name, v)
            stood_down = name == "auto" and ss * ss * spp >= 8 and _dense(st, w, h, ss)
            binned = {"binned": True, "auto": ss * ss * spp >= 8 and not stood_down, "point-parallel": False}[name]
            assert st["splat_binned"] == int(binned), (case, name, v, st)
            assert st["splat_fallback"] == (2 if stood_down else 0), (case, name, v, st)
        if name == "binned":
            st = batch.view_stats(0, stream)
            flush(f"{case}: entries {st['bin_entries']} jobs {st['bin_jobs']} big {st['bin_big']}")
            if case == "closeup_big":
                assert st["bin_big"] > 0
    if case == "budget":
        assert any(r[2]["budget_truncated"] for r in ref)


# ------------------------------------------------------------------ binned splat
@pytest.mark.parametrize("case", ["spp1_ss1", "spp2_ss2", "spp3_ss1", "spp4_ss2", "closeup_big", "budget",
                                  "occlusion_two_pass"])
def test_binned_splat_equals_point_parallel(gpu, case):
    """The binned (LDS region) splat and the point-parallel splat with global
    atomics write the same framebuffer bit for bit: single- and multi-job
    regions, Gaussians too large to bin, the point budget, and the second
    pass of occlusion culling (regions start from pass-1 contents)."""
    spp, ss = {"spp1_ss1": (1, 1), "spp2_ss2": (2, 2), "spp3_ss1": (3, 1), "spp4_ss2": (4, 2)}.get(case, (1, 2))
    cfg = dict(H.C1, half_extent=6.0, log_scale_mu=-2.0) if case != "closeup_big" else dict(H.C1, log_scale_mu=-0.7)
    arrays = H.scene(0x5EED0601 + len(case), 30_000, cfg).morton_sorted()
    w, h = 320, 200
    st = gpu.RenderSettings(w, h, spp, ss, max_number_of_points=5_000 if case == "budget" else 250_000_000,
                            gaussians_per_block=64)
    occl = case == "occlusion_two_pass"
    eye = (0.0, 2.0, -9.0) if case == "closeup_big" else (0.0, 3.0, -25.0)
    frames = [gpu.frame_params((eye[0] + i, eye[1], eye[2]), (0, 0, 0), (0, 1, 0), 45.0, w, h, ss=ss, frame=i + 1,
                               occlusion=occl, hierarchical=occl) for i in range(3 if occl else 1)]
    runs = []
    for mode in (gpu.Pass.SPLAT_BINNED, gpu.Pass.SPLAT_POINT_PARALLEL):
        r = H.GpuRunner(arrays, st)
        r.p.set_splat_mode(mode)
        r.p.set_counters(True)
        out = []
        for k, f in enumerate(frames):
            flush(f"binned {case}: mode {mode} frame {k}")
            res = r.render(f, [gpu.DBG_FRAMEBUFFER])
            out.append(res)
        runs.append(out)
    for k, (a, b) in enumerate(zip(*runs)):
        assert_same(f"{case} frame {k} fb", a.debug[gpu.DBG_FRAMEBUFFER], b.debug[gpu.DBG_FRAMEBUFFER])
        assert_same(f"{case} frame {k} colour", a.color, b.color)
        assert a.stats["splat_binned"] == 1 and a.stats["splat_fallback"] == 0
        assert b.stats["splat_binned"] == 0
        assert a.stats["num_points"] == b.stats["num_points"]
    s = runs[0][-1].stats
    flush(f"{case}: entries {s['bin_entries']} jobs {s['bin_jobs']} big {s['bin_big']} points {s['num_points']}")
    if case == "closeup_big":
        assert s["bin_big"] > 0
    if case == "spp1_ss1":
        assert s["bin_jobs"] > 6  # 3 x 2 regions of 128 px: some are split into several jobs


def _dense(st, w, h, ss):
    """More than 8 samples per framebuffer word (points * K > 8 * up_w * up_h)."""
    return st["num_points"][0] * st["K"] > 8 * (w * ss) * (h * ss)


def test_binned_density_gate(gpu):
    """The automatic splat choice bins a 4-sample frame only while it has at
    most 8 samples per framebuffer word: a dense frame (many points on few
    pixels) stands the binned kernels down on the device
    (splat_fallback = 2) and splats point-parallel, a sparse one stays
    binned, the forced binned form never stands down, and every form gives
    the same framebuffer bit for bit."""
    arrays = H.scene(0x5EED0611, 30_000, dict(H.C1, half_extent=6.0, log_scale_mu=-2.0)).morton_sorted()
    seen = set()
    for w, h in ((48, 32), (320, 200)):
        st = gpu.RenderSettings(w, h, 4, 2, gaussians_per_block=64)
        f = gpu.frame_params((0.0, 3.0, -25.0), (0, 0, 0), (0, 1, 0), 45.0, w, h, ss=2, frame=5, spp=4)
        runs = {}
        for mode, name in ((gpu.Pass.SPLAT_AUTO, "auto"), (gpu.Pass.SPLAT_BINNED, "binned"),
                           (gpu.Pass.SPLAT_POINT_PARALLEL, "point-parallel")):
            flush(f"density gate {w}x{h}: {name}")
            r = H.GpuRunner(arrays, st)
            r.p.set_splat_mode(mode)
            runs[name] = r.render(f, [gpu.DBG_FRAMEBUFFER])
        dense = _dense(runs["auto"].stats, w, h, 2)
        seen.add(dense)
        a, b, p = runs["auto"], runs["binned"], runs["point-parallel"]
        assert a.stats["splat_binned"] == 1 and a.stats["splat_fallback"] == (2 if dense else 0), (w, h, a.stats)
        assert b.stats["splat_binned"] == 1 and b.stats["splat_fallback"] == 0, (w, h, b.stats)
        assert p.stats["splat_binned"] == 0
        for name in ("auto", "binned"):
            assert_same(f"{w}x{h} {name} fb", runs[name].debug[gpu.DBG_FRAMEBUFFER], p.debug[gpu.DBG_FRAMEBUFFER])
            assert_same(f"{w}x{h} {name} colour", runs[name].color, p.color)
    assert seen == {True, False}


def test_batch_binned_density_gate(gpu):
    """In a batch the density test is per view: a dense view is splatted
    point-parallel (splat_fallback = 2) while the sparse views of the same
    batch stay binned, and every view equals a pass rendering it alone."""
    arrays = H.scene(0x5EED0612, 30_000, dict(H.C1, half_extent=6.0, log_scale_mu=-2.0)).morton_sorted()
    scene = gpu.Scene(arrays, gpu.FORMAT_FP32)
    stream = gpu.Stream()
    batch = gpu.Batch(scene, 3, 320, 200, max_spp=4, max_supersampling=2, gaussians_per_block=64)
    single = gpu.Pass(gpu.RenderSettings(320, 200, 4, 2, flexible_frame_size=True, gaussians_per_block=64), scene)
    shapes = [(48, 32), (320, 200), (64, 40)]
    frames = [gpu.frame_params((0.5 * v, 3.0, -25.0), (0, 0, 0), (0, 1, 0), 45.0, w, h, ss=2, frame=v + 7, spp=4)
              for v, (w, h) in enumerate(shapes)]
    ref = []
    for f in frames:
        c = gpu.DeviceArray((f.height, f.width, 3), np.float32)
        d = gpu.DeviceArray((f.height, f.width), np.float32)
        single.set_splat_mode(gpu.Pass.SPLAT_POINT_PARALLEL)
        single.render(f, c, d, stream)
        stream.synchronize()
        ref.append((c.download(), d.download()))
    cols = [gpu.DeviceArray((f.height, f.width, 3), np.float32, zero=True) for f in frames]
    deps = [gpu.DeviceArray((f.height, f.width), np.float32, zero=True) for f in frames]
    flush("batch density gate: auto")
    batch.render(frames, cols, deps, stream)
    stream.synchronize()
    kinds = []
    for v, (w, h) in enumerate(shapes):
        assert_same(f"view {v} colour", cols[v].download(), ref[v][0])
        assert_same(f"view {v} depth", deps[v].download(), ref[v][1])
        st = batch.view_stats(v, stream)
        dense = _dense(st, w, h, 2)
        kinds.append(dense)
        assert st["splat_binned"] == (0 if dense else 1) and st["splat_fallback"] == (2 if dense else 0), (v, st)
    assert kinds[0] and not kinds[1], kinds


def test_device_fp8_decode_matches_oracle(gpu):
    """Every FP8 E4M3 code through the hardware conversion the compact SH
    path uses equals the oracle's software decode (NaN codes: both NaN)."""
    import orc
    x = np.arange(256, dtype=np.float32)
    d_in = gpu.DeviceArray(x.shape, np.float32)
    d_out = gpu.DeviceArray(x.shape, np.float32)
    d_in.upload(x)
    gpu.check(gpu.lib.gps_selftest_math(11, d_in.ptr, d_out.ptr, x.size, None), "selftest")
    dev = d_out.download()
    host = orc.detmath(11, x)
    nan = np.isnan(host)
    assert nan.sum() == 2 and np.isnan(dev[nan]).all()
    assert_same("fp8 decode", dev[~nan], host[~nan])


def test_cli_camera_path_run_tests(gpu, tmp_path):
    """The headless driver on a camera-path file with --run-tests: the
    cameras in main.cu's order, the per-camera timing JSON of
    main.cu:211-241, and the resolution the first camera sets (width kept,
    height from its 4946 x 3286 aspect ratio)."""
    import json
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "gaussian-point-splatting_amd", "gps_render")
    fix = os.path.join(root, "tests", "golden", "truck_cameras_head.json")
    out = str(tmp_path / "frame")
    flush("cli: camera path run-tests")
    r = subprocess.run([exe, "--synthetic", "c1", "--num-gaussians", "20000", "--camera-path", fix, "--run-tests",
                        "--frames-per-camera-keyframe", "2", "--set-target-width", "320", "--out", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    text = r.stdout
    body = text[text.index("---BEGIN_TIMINGS_JSON---") + len("---BEGIN_TIMINGS_JSON---"):
                text.index("---END_TIMINGS_JSON---")]
    doc = json.loads(body)
    assert doc["algorithm"] == "GaussianPointSplatting"
    assert [t["camera"] for t in doc["timings"]] == [0, 1, 2, 3]
    assert all(len(t["samples"]) == 2 and all(s > 0 for s in t["samples"]) for t in doc["timings"])
    # frames 1 .. 3 * 4 + 1 were written at 320 x round(320 * 3286 / 4946)
    with open(out + "_0001.ppm", "rb") as fh:
        head = fh.read(32).split()
    assert head[0] == b"P6" and (int(head[1]), int(head[2])) == (320, round(320 * 3286 / 4946))
    assert os.path.exists(out + f"_{3 * 4 + 1:04d}.ppm")


def read_png_rgb(path):
    """RGB8 PNG (8-bit, colour type 2, no interlace) -> (h, w, 3) array."""
    import struct
    import zlib
    with open(path, "rb") as fh:
        d = fh.read()
    assert d[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(d):
        n = struct.unpack(">I", d[pos:pos + 4])[0]
        t, data = d[pos + 4:pos + 8], d[pos + 8:pos + 8 + n]
        assert zlib.crc32(t + data) & 0xFFFFFFFF == struct.unpack(">I", d[pos + 8 + n:pos + 12 + n])[0]
        if t == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", data)
        elif t == b"IDAT":
            idat += data
        pos += 12 + n
    w, h, depth, ctype = hdr[0], hdr[1], hdr[2], hdr[3]
    assert depth == 8 and ctype == 2
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, w * 3 + 1)
    assert (rows[:, 0] == 0).all()  # filter "none"
    return rows[:, 1:].reshape(h, w, 3)


def test_cli_evaluation_screenshots(gpu, tmp_path):
    """--create-evaluation-screenshots / --max-eval-count (main.cu:72-120,
    render_server.cu:263-315): per camera the progressive image after 1, 4,
    16, 64, 256 and 1024 frames, named <image_name>_sppNNNN.png under
    <dir>/eval/<model name>/, for cameras 0 .. max_eval_count; the first
    image equals the same frame rendered and resolved through the library,
    rows flipped as save_image_to_disk flips them."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "gaussian-point-splatting_amd", "gps_render")
    fix = os.path.join(root, "tests", "golden", "truck_cameras_head.json")
    flush("cli: evaluation screenshots")
    r = subprocess.run([exe, "--synthetic", "c1", "--num-gaussians", "20000", "--camera-path", fix,
                        "--create-evaluation-screenshots", "--max-eval-count", "1", "--set-target-width", "160",
                        "--screenshot-dir", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    cams = gpu.camera_path_load(fix)
    names = [c.img_name.decode() for c in cams]
    d = tmp_path / "eval" / "synthetic_c1"
    want = sorted(f"{names[c]}_spp{s:04d}.png" for c in (0, 1) for s in (1, 4, 16, 64, 256, 1024))
    assert sorted(os.listdir(d)) == want
    w, h = 160, round(160 * 3286 / 4946)
    img = {s: read_png_rgb(str(d / f"{names[0]}_spp{s:04d}.png")) for s in (1, 4, 256, 1024)}
    assert img[1].shape == (h, w, 3)
    # frame 1 of camera 0 through the library: same scene, pass, camera, resolve
    arrays = H.scene(0x5EED0001, 20_000, H.C1)
    scene = gpu.Scene(arrays)
    ps = gpu.Pass(gpu.RenderSettings(w, h, 1, 2), scene)
    f = gpu.frame_from_camera(cams[0], w, h, 0.0, 2, 0.2, 1e4)
    view = gpu.mat4_mul(np.array(f.view, np.float32), np.eye(4, dtype=np.float32).reshape(-1))  # scene matrix I
    for i in range(16):
        f.view[i] = float(view[i])
    planes = np.empty(24, np.float32)
    gpu._lib.gps_frustum_planes(f.fovy_degrees, view.ctypes.data_as(gpu._fp), w / h, 0.2, 1e4,
                                planes.ctypes.data_as(gpu._fp))
    for i in range(24):
        f.frustum_planes[i] = float(planes[i])
    f.frame = 1
    f.enable_occlusion_culling = 1
    f.enable_hierarchical_culling = 1
    f.cull_small_gaussians = 0
    f.reduce_point_count = 1
    f.use_unbiased_2d_splatting = 1
    color = gpu.DeviceArray((h, w, 3), np.float32, zero=True)
    accum = gpu.DeviceArray((h, w, 3), np.float32, zero=True)
    depth = gpu.DeviceArray((h, w), np.float32, zero=True)
    rgb = gpu.DeviceArray((h, w, 3), np.uint8, zero=True)
    ps.render(f, color, depth)
    gpu.resolve_progressive(color, accum, w, h, 1, rgb)
    assert_same("eval spp0001", img[1], rgb.download()[::-1])
    # the accumulation converges: 1024 frames are closer to 256 than 4 are
    err = {s: np.abs(img[s].astype(np.int32) - img[256].astype(np.int32)).mean() for s in (4, 1024)}
    assert err[1024] < err[4], err


def test_selftest_atomics_min(gpu):
    """gps_selftest_atomics (the atomic-rate probe bench.py prices the splat
    against) leaves every word at the minimum of the values hashed to it."""
    words, blocks, iters = 1000, 3, 5
    flush("selftest atomics")
    buf = gpu.DeviceArray((words,), np.uint64)
    buf.upload(np.full(words, np.iinfo(np.uint64).max, np.uint64))
    gpu.check(gpu.lib.gps_selftest_atomics(buf.ptr, words, blocks, iters, None), "atomics")
    got = buf.download()

    def h32(x):  # lowbias32
        x = np.uint32(x)
        x ^= x >> np.uint32(16)
        x = np.uint32((int(x) * 0x7FEB352D) & 0xFFFFFFFF)
        x ^= x >> np.uint32(15)
        x = np.uint32((int(x) * 0x846CA68B) & 0xFFFFFFFF)
        x ^= x >> np.uint32(16)
        return int(x)

    want = np.full(words, np.iinfo(np.uint64).max, np.uint64)
    for t in range(blocks * 256):
        x = t ^ 0x5EED
        for _ in range(iters):
            x = h32((x + 0x9E3779B9) & 0xFFFFFFFF)
            idx = (x * words) >> 32
            want[idx] = min(int(want[idx]), (h32(x) << 32) | x)
    assert np.array_equal(got, want)

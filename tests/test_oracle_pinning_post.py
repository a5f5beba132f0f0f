"""Pins for the oracle's post-splat stages and scene decoding, each against an
independent numpy computation or a case the mathematics fixes:

* combine: colour = mean of the decoded 12-bit channels over the ss x ss x spp
  samples of a pixel, depth = min NDC z, recomputed in float64 from the
  framebuffer with numpy's own Morton addressing and the projection matrix
* TAA: no history on frame 0; a static camera over a constant image is a
  fixed point; history off screen leaves the current colour; history is
  clamped to the 3x3 neighbourhood +-0.05; otherwise a lerp with the
  bilinear history
* progressive accumulation: running sum / n, rounded to 8 bits
* compact fp16 / fp8 scene: decoding equals numpy's float16 and torch's
  float8_e4m3fn decoding (same framebuffer bit for bit)
* block culling: frustum part — a block is culled iff its 3-sigma AABB lies
  entirely on the negative side of one of the six planes; Hi-Z part — a block
  is culled only if every on-screen Gaussian centre in it lies behind the
  finest Hi-Z texel under it
"""
import numpy as np
import pytest

import gps_amd as G
import helpers as H
import orc


def morton_np(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    code = np.zeros_like(x, dtype=np.int64)
    for b in range(6):
        code |= ((x >> b) & 1) << (2 * b)
        code |= ((y >> b) & 1) << (2 * b + 1)
    return code


def fb_index_np(x, y, layer, spp, up_w):
    tiles_per_row = -(-up_w // 64)
    return ((((y // 64) * tiles_per_row + (x // 64)) * 4096 + morton_np(x % 64, y % 64)) * spp + layer)


def test_combine_is_sample_mean_and_min_depth():
    arrays = H.scene(0x5EED0401, 8000, H.C1)
    w, h, ss, spp = 96, 64, 2, 3
    st = G.RenderSettings(w, h, spp, ss, fully_disable_hierarchical_culling=True)
    o = orc.OraclePass(st, arrays, threads=4)
    near, far = 0.2, 1e4
    f = G.frame_params((0, 0, -30), (0, 0, 0), (0, 1, 0), 45.0, w, h, ss=ss, near=near, far=far)
    color, depth = o.render(f)
    fb = o.debug(G.DBG_FRAMEBUFFER)
    ys, xs = np.mgrid[0:h, 0:w]
    col = np.zeros((h, w, 3))
    dmin = np.full((h, w), np.inf)
    proj = np.asarray(G.perspective(45.0, w / h, near, far), np.float64).reshape(4, 4).T
    for dx in range(ss):
        for dy in range(ss):
            for layer in range(spp):
                words = fb[fb_index_np(xs * ss + dx, ys * ss + dy, layer, spp, w * ss)]
                for c, shift in enumerate((24, 12, 0)):
                    col[..., c] += ((words >> np.uint64(shift)) & np.uint64(0xFFF)).astype(np.float64) / 255.0
                d28 = ((words >> np.uint64(36)) & np.uint64((1 << 28) - 1)).astype(np.float64)
                z_view = -(d28 / ((1 << 28) - 1) * (far - near) + near)
                clip = np.einsum("ij,...j->...i", proj, np.stack([0 * z_view, 0 * z_view, z_view, 1 + 0 * z_view], -1))
                dmin = np.minimum(dmin, clip[..., 2] / clip[..., 3])
    col /= ss * ss * spp
    assert np.abs(color - col).max() < 1e-5
    assert np.abs(depth - dmin).max() < 1e-5 * np.abs(dmin).max()
    # something was drawn, and background pixels keep the background word
    bg_word = np.uint64(0xFFFFFFF000000000)
    assert (fb != bg_word).any() and (fb == bg_word).any()


def test_combine_background_only():
    arrays = H.scene(0x5EED0402, 10, H.C1)
    st = G.RenderSettings(64, 40, 2, 2, fully_disable_hierarchical_culling=True)
    o = orc.OraclePass(st, arrays, threads=2)
    # camera looking away from the scene: nothing splats
    f = G.frame_params((0, 0, -30), (0, 0, -60), (0, 1, 0), 45.0, 64, 40, ss=2, background=(0.25, 0.5, 1.0))
    color, _ = o.render(f)
    # 36-bit colour: 12-bit channels of the truncated float32 product c * 255,
    # averaged back / 255
    expect = np.floor(np.array([0.25, 0.5, 1.0], np.float32) * np.float32(255)).astype(np.float64) / 255.0
    assert list(expect * 255) == [63.0, 127.0, 255.0]
    assert np.abs(color - expect).max() < 1e-6


def taa(cur, cdep, prev, M, blend, fn):
    h, w = cdep.shape
    return orc.resolve_taa(cur, cdep, prev, cdep, M, blend, w, h, fn)


def test_taa_fixed_cases():
    w, h = 48, 32
    rng = np.random.default_rng(11)
    cur = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    dep = rng.uniform(0.95, 0.99, (h, w)).astype(np.float32)
    prev = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    ident = np.eye(4, dtype=np.float32).reshape(-1)
    # frame 0: no history
    out, rgb = taa(cur, dep, prev, ident, 0.9, 0)
    assert np.array_equal(out, cur)
    assert np.array_equal(rgb, np.round(np.clip(cur, 0, 1) * 255).astype(np.uint8))
    # constant image under a static camera is a fixed point
    const = np.full((h, w, 3), 0.3, np.float32)
    out, _ = taa(const, dep, const, ident, 0.9, 5)
    assert np.abs(out - 0.3).max() < 1e-6
    # history reprojected off screen: current colour kept
    shift = np.eye(4, dtype=np.float32)
    shift[0, 3] = 5.0  # x_ndc += 5
    out, _ = taa(cur, dep, prev, shift.T.reshape(-1), 0.9, 5)
    assert np.array_equal(out, cur)
    # identity reprojection: history = prev at the same pixel, clamped to the
    # 3x3 neighbourhood of the current frame widened by 0.05, then lerped
    out, _ = taa(cur, dep, prev, ident, 0.9, 5)
    pad = np.pad(cur, ((1, 1), (1, 1), (0, 0)), mode="edge")
    win = np.stack([pad[dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)])
    lo, hi = win.min(0) - 0.05, win.max(0) + 0.05
    expect = cur * 0.1 + np.clip(prev, lo, hi) * 0.9
    # (row / column 0 can reproject to -1e-7 and count as off screen)
    assert np.abs(out - expect)[1:, 1:].max() < 1e-5
    edge = np.zeros((h, w), bool)
    edge[0, :] = edge[:, 0] = True
    assert (np.isclose(out, expect, atol=1e-5).all(-1) | (out == cur).all(-1))[edge].all()


def test_progressive_accumulation():
    w, h = 20, 12
    rng = np.random.default_rng(12)
    frames = [rng.uniform(0, 1.2, (h, w, 3)).astype(np.float32) for _ in range(5)]
    accum = np.zeros((h, w, 3), np.float32)
    total = np.zeros((h, w, 3), np.float64)
    for n, c in enumerate(frames, start=1):
        col, acc, rgb = orc.resolve_progressive(c, accum, w, h, n)
        total += c
        assert np.abs(col - total).max() < 1e-5  # running sum kept in the colour buffer
        assert not acc.any()
        expect = np.clip(np.round(total / n * 255), 0, 255)
        assert np.abs(rgb.astype(np.int64) - expect).max() <= 1
        accum = col


def test_half_fp8_decode_matches_numpy_and_torch():
    torch = pytest.importorskip("torch")
    arrays = H.scene(0x5EED0403, 6000, H.C2, clusters=30)
    e = G.encode_half_fp8(arrays)
    dec = G.SceneArrays(
        arrays.means.copy(),
        e["rotations"].view(np.float16).astype(np.float32),
        e["scales"].view(np.float16).astype(np.float32),
        e["opacities"].view(np.float16).astype(np.float32),
        np.concatenate([e["sh_dc"].view(np.float16).astype(np.float32),
                        torch.from_numpy(e["sh_rest"]).view(torch.float8_e4m3fn).float().numpy()
                        * e["sh_scale"][:, None]], axis=1).astype(np.float32),
        3)
    st = G.RenderSettings(160, 100, 2, 2, fully_disable_hierarchical_culling=True)
    f = G.frame_params((0, 50, -300), (0, 0, 0), (0, 1, 0), 45.0, 160, 100, ss=2)
    a = orc.OraclePass(st, arrays, G.FORMAT_HALF_FP8, threads=4)
    b = orc.OraclePass(st, dec, G.FORMAT_FP32, threads=4)
    a.render(f)
    b.render(f)
    H.assert_same("fb", a.debug(G.DBG_FRAMEBUFFER), b.debug(G.DBG_FRAMEBUFFER))
    H.assert_same("weights", a.debug(G.DBG_WEIGHTS), b.debug(G.DBG_WEIGHTS))


def test_block_frustum_culling_is_aabb_plane_test():
    # Morton order so that blocks are spatially compact
    arrays = H.scene(0x5EED0404, 40_000, dict(H.C1, half_extent=40.0, log_scale_mu=-1.5)).morton_sorted()
    gpb = 16
    st = G.RenderSettings(200, 120, 1, 1, gaussians_per_block=gpb)
    o = orc.OraclePass(st, arrays, threads=4)
    f = G.frame_params((10, 5, -35), (30, 0, 10), (0, 1, 0), 45.0, 200, 120, hierarchical=True)
    o.render(f)
    mask = o.debug(G.DBG_CULL_MASK)
    nb = -(-arrays.n // gpb)
    culled = np.unpackbits(mask.view(np.uint8), bitorder="little")[:nb].astype(bool)
    # 3-sigma AABB of each Gaussian from the diagonal of R S S^T R^T, in float64
    q = arrays.rotations.astype(np.float64)
    w_, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w_ * z), 2 * (x * z + w_ * y)], -1),
                  np.stack([2 * (x * y + w_ * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w_ * x)], -1),
                  np.stack([2 * (x * z - w_ * y), 2 * (y * z + w_ * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    M = R * arrays.scales.astype(np.float64)[:, None, :]
    ext = 3.0 * np.sqrt((M * M).sum(-1))
    lo = arrays.means - ext
    hi = arrays.means + ext
    pad = nb * gpb - arrays.n
    lo = np.concatenate([lo, np.full((pad, 3), np.inf)]).reshape(nb, gpb, 3).min(1)
    hi = np.concatenate([hi, np.full((pad, 3), -np.inf)]).reshape(nb, gpb, 3).max(1)
    planes = np.array(f.frustum_planes, np.float64).reshape(6, 4)
    # p-vertex: the corner furthest along each plane normal
    pv = np.where(planes[None, :, :3] >= 0, hi[:, None, :], lo[:, None, :])
    dist = (pv * planes[None, :, :3]).sum(-1) + planes[None, :, 3]
    outside = (dist < 0).any(1)
    margin = np.abs(dist).min(1) < 1e-3
    assert 0.1 < outside.mean() < 0.9, outside.mean()
    assert np.array_equal(culled[~margin], outside[~margin])


def test_block_occlusion_culling_is_conservative():
    """A block may be culled against the Hi-Z only if every Gaussian centre of
    it that lands on screen lies behind the finest Hi-Z texel under it
    (culled => centre depth >= min corner depth > max over the block's
    rectangle >= that texel)."""
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 20_000, width=160, height=90, gaussians_per_block=32)
    arrays = W.host_scene(w)
    o = orc.OraclePass(w.settings(), arrays, w.fmt, 4)
    f = W.frame(w, 0, 1)  # eye at x=-300 looking +x: the copies hide each other
    o.render(f)           # leaves the Hi-Z of this view in the pass
    hiz = o.debug(G.DBG_HIZ)[-w.width * w.height:].reshape(w.height, w.width)
    nb = -(-arrays.n // 32)
    bits = lambda m: np.unpackbits(m.view(np.uint8), bitorder="little")[:nb].astype(bool)
    occl = bits(o.cull(f))
    f_nocc = W.frame(w, 0, 1)
    f_nocc.enable_occlusion_culling = 0
    frustum = bits(o.cull(f_nocc))
    assert (occl >= frustum).all()
    by_hiz = occl & ~frustum
    assert by_hiz.sum() > nb // 10, (by_hiz.sum(), nb)
    V = np.asarray(f.view, np.float64).reshape(4, 4).T
    P = np.asarray(f.proj, np.float64).reshape(4, 4).T
    clip = np.c_[arrays.means.astype(np.float64), np.ones(arrays.n)] @ (P @ V).T
    front = clip[:, 3] > 0
    ndc = clip[:, :3] / np.where(front, clip[:, 3], 1.0)[:, None]
    px = np.floor((ndc[:, 0] * 0.5 + 0.5) * w.width).astype(np.int64)
    py = np.floor((ndc[:, 1] * 0.5 + 0.5) * w.height).astype(np.int64)
    on = front & (px >= 0) & (px < w.width) & (py >= 0) & (py < w.height) & (np.abs(ndc[:, 2]) <= 1)
    in_culled = np.repeat(occl, 32)[:arrays.n]
    sel = on & in_culled
    assert sel.sum() > 1000
    assert (ndc[sel, 2] >= hiz[py[sel], px[sel]] - 1e-6).all()

"""Host-side checks of libgps_amd.so that need no GPU: the C ABI exports,
camera / matrix math, the compact-format encoders, the synthetic scene
generator and argument validation of the entry points."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import gps_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "gps.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gps_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r" T (gps_\w+)", out.stdout))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(G.lib, n)


def test_abi_version_and_status_strings():
    assert G.lib.gps_abi_version() == 3
    for s in range(7):
        assert G.lib.gps_status_str(s)


def test_entry_points_reject_bad_arguments():
    h = C.c_void_p()
    assert G.lib.gps_scene_create(None, C.byref(h)) == 1
    d = G.SceneDesc()
    d.num_gaussians = 10
    d.sh_degree = 5
    assert G.lib.gps_scene_create(C.byref(d), C.byref(h)) == 1
    d.sh_degree = 0
    d.format = 0
    assert G.lib.gps_scene_create(C.byref(d), C.byref(h)) == 1  # null arrays
    assert G.lib.gps_pass_create(None, None, C.byref(h)) == 1
    assert G.lib.gps_pass_render(None, None, None, None, None) == 1
    assert G.lib.gps_resolve_progressive(None, None, 4, 4, 1, None, None) == 1
    assert G.lib.gps_selftest_math(99, None, None, 0, None) == 1
    s = G.RenderSettingsC()
    s.width, s.height, s.render_pass_count, s.supersampling_factor = 100, 100, 0, 1
    assert G.lib.gps_pass_create(C.byref(s), C.c_void_p(1), C.byref(h)) == 1
    # a framebuffer of 2^31 words or more (32-bit word indices) is refused
    # before anything is allocated: 40000^2 pixels at 2 samples = 3.2e9 words
    s.width, s.height, s.render_pass_count, s.supersampling_factor = 40000, 40000, 2, 1
    assert G.lib.gps_pass_create(C.byref(s), C.c_void_p(1), C.byref(h)) == 4
    # load-time BVH order: argument checks come before any device work
    assert G.lib.gps_bvh_order(None, None, None, 5, None) == 1
    assert G.lib.gps_bvh_order(None, None, None, 0, None) == 0  # nothing to order
    assert G.lib.gps_selftest_atomics(None, 16, 1, 1, None) == 1
    assert G.lib.gps_selftest_atomics(C.c_void_p(8), 0, 1, 1, None) == 1
    assert G.lib.gps_selftest_math(13, None, None, 4, None) == 1  # null buffers
    assert G.lib.gps_selftest_math(14, C.c_void_p(8), C.c_void_p(8), 4, None) == 1  # unknown kind
    # per-pass and per-batch switches: a null handle is refused
    assert G.lib.gps_pass_set_splat_mode(None, 0) == 1
    assert G.lib.gps_pass_set_ordered_splat(None, 0) == 1
    assert G.lib.gps_pass_set_group_culling(None, 1) == 1
    assert G.lib.gps_batch_set_tuning(None, None) == 1


def lookat_ref(eye, center, up):
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = [-s @ eye, -u @ eye, f @ eye]
    return m


def colmajor(a):
    return np.asarray(a, np.float64).reshape(4, 4).T


def test_look_at_and_perspective():
    eye, ctr = (3.0, 2.0, -10.0), (0.5, -1.0, 2.0)
    v = colmajor(G.look_at(eye, ctr, (0, 1, 0)))
    assert np.allclose(v, lookat_ref(eye, ctr, (0, 1, 0)), atol=1e-6)
    p = colmajor(G.perspective(45.0, 16 / 9, 0.2, 1e4))
    t = np.tan(np.radians(45.0) / 2)
    assert np.isclose(p[0, 0], 1 / (16 / 9 * t), rtol=1e-6) and np.isclose(p[1, 1], 1 / t, rtol=1e-6)
    assert p[3, 2] == -1.0 and np.isclose(p[2, 3], -2 * 1e4 * 0.2 / (1e4 - 0.2), rtol=1e-6)
    # a point on the view axis at depth z maps to NDC z = A + B / z_view
    ndc = p @ np.array([0, 0, -5.0, 1])
    assert np.isclose(ndc[2] / ndc[3], (1e4 + 0.2) / (1e4 - 0.2) - 2 * 1e4 * 0.2 / (1e4 - 0.2) / 5.0, rtol=1e-6)


def test_frustum_planes_contain_visible_points():
    f = G.frame_params((0, 0, -30), (0, 0, 0), (0, 1, 0), 45.0, 640, 360)
    planes = np.array(f.frustum_planes).reshape(6, 4)
    assert np.allclose(np.linalg.norm(planes[:, :3], axis=1), 1.0, atol=1e-6)
    inside = np.array([[0, 0, 0], [5, 3, 0], [0, 0, 500]], np.float64)
    outside = np.array([[0, 0, -31], [100, 0, 0], [0, 0, 2e4]], np.float64)
    assert (inside @ planes[:, :3].T + planes[:, 3] >= 0).all()
    assert ((outside @ planes[:, :3].T + planes[:, 3]) < 0).any(axis=1).all()


def test_taa_reprojection_identity_for_static_camera():
    v = G.look_at((1, 2, -20), (0, 0, 0))
    p = G.perspective(45.0, 1.5, 0.2, 1e4)
    vp = G.mat4_mul(p, v)
    m = colmajor(G.taa_reprojection(vp, v, p))
    for clip in ([0.1, -0.2, 0.95, 1.0], [-0.7, 0.4, 0.99, 1.0]):
        r = m @ np.array(clip)
        assert np.allclose(r / r[3], clip, atol=1e-4)


def test_half_encoder_matches_numpy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-5, 5000), rng.normal(0, 3e4, 2000),
                        [0.0, -0.0, 65504.0, 65520.0, 1e9, 2 ** -24, 2 ** -25, 3 * 2 ** -26, 6.1e-5]]).astype(
        np.float32)
    got = np.array([G.lib.gps_float_to_half(float(v)) for v in x], np.uint16)
    assert np.array_equal(got, x.astype(np.float16).view(np.uint16))


def test_fp8_encoder_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-448, 448, 20000), rng.normal(0, 0.01, 5000),
                        [0.0, 448.0, -448.0, 2 ** -9, 2 ** -10, 3 * 2 ** -10, 2 ** -6, 240.0, 0.0017]]).astype(
        np.float32)
    got = np.array([G.lib.gps_float_to_fp8e4m3(float(v)) for v in x], np.uint8)
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(got, ref)
    assert G.lib.gps_float_to_fp8e4m3(1e6) == 0x7E and G.lib.gps_float_to_fp8e4m3(-1e6) == 0xFE


def test_half_fp8_scene_encoding_error():
    p = G.synth_params(3, 5000, kind="clustered", sh_degree=3, clusters=20, half_extent=100.0, log_scale_mu=-3.0,
                       log_scale_sigma=0.7, opacity_sigma=1.5, rest_sigma=0.05)
    s = G.synth_scene(p)
    e = G.encode_half_fp8(s)
    rest = s.shs[:, 3:]
    scale = e["sh_scale"][:, None]
    assert np.allclose(scale[:, 0], np.abs(rest).max(axis=1) / 448.0, rtol=1e-6)
    torch = pytest.importorskip("torch")
    dec = torch.from_numpy(e["sh_rest"]).view(torch.float8_e4m3fn).float().numpy() * scale
    err = np.abs(dec - rest) / scale / 448.0
    assert err.max() <= 2.0 ** -4 + 1e-7  # 3 mantissa bits, relative to the per-Gaussian max
    assert np.allclose(e["rotations"].view(np.float16).astype(np.float32), s.rotations, atol=2 ** -11)
    assert np.array_equal(e["sh_dc"], s.shs[:, :3].astype(np.float16).view(np.uint16))


def test_synthetic_generator_deterministic_and_chunk_independent():
    p = G.synth_params(7, 10000, kind="clustered", sh_degree=3, clusters=50, half_extent=100.0)
    a = G.synth_scene(p)
    b = G.synth_scene(p, 6000, 100)
    assert np.array_equal(a.means[6000:6100], b.means) and np.array_equal(a.shs[6000:6100], b.shs)
    assert np.allclose(np.linalg.norm(a.rotations, axis=1), 1.0, atol=1e-6)
    assert ((a.opacities > 0) & (a.opacities < 1)).all()
    u = G.synth_scene(G.synth_params(8, 20000, kind="uniform", half_extent=10.0))
    assert np.abs(u.means).max() <= 10.0
    assert abs(np.log(u.scales).mean() + 4.0) < 0.02 and abs(np.log(u.scales).std() - 0.5) < 0.02


def test_synthetic_gather_and_positions_only():
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 5000)
    p = W.params(w)
    base = G.synth_scene(p)
    idx = np.array([4999, 0, 17, 17, 5000 + 3, 3 * 5000 + 42], np.uint64)
    got = W._generate(w, idx)
    g = (idx % 5000).astype(np.int64)
    shift = (idx // 5000).astype(np.float32) * np.float32(250.0)
    assert np.array_equal(got.means[:, 0], base.means[g, 0] + shift)
    assert np.array_equal(got.means[:, 1:], base.means[g, 1:])
    for name in ("rotations", "scales", "opacities", "shs"):
        assert np.array_equal(getattr(got, name), getattr(base, name)[g]), name
    assert np.array_equal(W._generate(w, idx, means_only=True), got.means)
    # whole-scene Morton order is a permutation and is spatially coherent
    perm = W.order(w)
    assert np.array_equal(np.sort(perm), np.arange(w.n, dtype=np.uint64))
    m = W._generate(w, perm, means_only=True)
    step = np.linalg.norm(np.diff(m, axis=0), axis=1)
    rnd = np.linalg.norm(np.diff(m[np.random.default_rng(0).permutation(w.n)], axis=0), axis=1)
    assert np.median(step) < 0.1 * np.median(rnd)
    bad = np.array([w.n], np.uint64)
    with pytest.raises(G.GpsError):
        G.check(G.lib.gps_synth_generate_indices(C.byref(p), G._ptr(bad % np.uint64(10**9)), 1,
                                                 G._ptr(np.empty((1, 3), np.float32)), None, None, None, None),
                "synth")


def test_streamed_scene_upload_validation():
    h = C.c_void_p()
    assert G.lib.gps_scene_create_empty(10, 4, 0, C.byref(h)) == 1
    assert G.lib.gps_scene_create_empty(10, 0, 7, C.byref(h)) == 1
    assert G.lib.gps_scene_upload_range(None, 0, None) == 1

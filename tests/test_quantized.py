"""The reference's compressed scene (GPS_FORMAT_QUANTIZED): host encoder and
the oracle's decode (gaussian_types.cuh:123-374, 421-541; math/sh.cuh:28-96,
257-273), pinned by what the format fixes rather than by restating it:

  * the scene box is the componentwise min / max of the data (position box
    widened by 0.001),
  * decode(encode(x)) is within one quantization step of x for every field
    (a mis-split bit field, a swapped quaternion component or an inverted
    mix fails by orders of magnitude),
  * a zero-width box decodes exactly to its single value,
  * the Sloan degree-3 basis equals the 3DGS basis on unit directions, so the
    compressed colour evaluation agrees with the independently written
    computeColorFromSHFloat for the same coefficients,
  * quat_iem at the corner code (1023, 1023, 1023) against its closed form.
"""
import math

import numpy as np
import pytest

import helpers as H
import gps_amd as G


def _scene(n=4000, seed=0x5EED0901):
    return H.scene(seed, n, H.C2, clusters=40)


def test_bounds_are_the_data_box():
    s = _scene()
    b = G.quantized_bounds(s)
    f32 = np.float32
    assert np.array_equal(np.array(b.min_pos, f32), s.means.min(0) - f32(0.001))
    assert np.array_equal(np.array(b.max_pos, f32), s.means.max(0) + f32(0.001))
    ls = np.log(s.scales.astype(np.float64))
    assert np.allclose(np.array(b.min_log_scale), ls.min(0), rtol=0, atol=2e-6)
    assert np.allclose(np.array(b.max_log_scale), ls.max(0), rtol=0, atol=2e-6)
    rest = s.shs[:, 3:]
    assert np.array_equal(np.array(b.min_sh, f32), rest.min(0))
    assert np.array_equal(np.array(b.max_sh, f32), rest.max(0))


def test_round_trip_within_one_step():
    import orc
    s = _scene()
    enc = G.encode_quantized(s)
    b = enc["bounds"]
    dec = orc.quantized_decode(enc)
    mean, rot, scale, op, sh = dec[:, 0:3], dec[:, 3:7], dec[:, 7:10], dec[:, 10], dec[:, 11:59]
    # position: truncation to 23 / 22 / 23 bits of the widened box
    span = np.array(b.max_pos, np.float64) - np.array(b.min_pos, np.float64)
    step = span / np.array([2 ** 23 - 1, 2 ** 22 - 1, 2 ** 23 - 1])
    err = np.abs(mean.astype(np.float64) - s.means)
    assert (err <= step * 1.001 + 4e-6 * np.abs(s.means).max()).all(), err.max(0) / step
    # log-scale: 10 bits of the log box
    lstep = (np.array(b.max_log_scale, np.float64) - np.array(b.min_log_scale, np.float64)) / 1023
    lerr = np.abs(np.log(scale.astype(np.float64)) - np.log(s.scales.astype(np.float64)))
    assert (lerr <= lstep * 1.001 + 1e-5).all(), lerr.max(0) / lstep
    # rotation: a unit quaternion for the same rotation; the angle error of a
    # truncated 10-bit component is at most pi * 2 / 1023, so <= 0.011 rad
    q0 = s.rotations / np.linalg.norm(s.rotations, axis=1, keepdims=True)
    assert np.allclose(np.linalg.norm(rot, axis=1), 1.0, atol=2e-6)
    cosang = np.abs((q0 * rot).sum(1))
    assert (cosang > math.cos(0.012 / 2)).all(), cosang.min()
    # opacity: one byte, rounded
    assert (np.abs(op.astype(np.float64) - s.opacities) <= 0.5 / 255 + 1e-7).all()
    # SH: DC kept in fp32, the rest truncated to bytes of each coefficient's range
    assert np.array_equal(sh[:, :3], s.shs[:, :3])
    rstep = (np.array(b.max_sh, np.float64) - np.array(b.min_sh, np.float64)) / 255
    rerr = np.abs(sh[:, 3:].astype(np.float64) - s.shs[:, 3:])
    assert (rerr <= rstep * 1.001 + 1e-7).all(), (rerr / rstep).max()


def test_single_gaussian_zero_width_boxes():
    import orc
    s = _scene(n=1)
    enc = G.encode_quantized(s)
    assert enc["geometry"][0, 3] >> 2 == 0  # 0 / 0 scale codes encode as 0
    assert not enc["sh_codes"].any()
    dec = orc.quantized_decode(enc)[0]
    assert np.allclose(dec[0:3], s.means[0], rtol=0, atol=2e-3 / 2 ** 22)
    assert np.allclose(dec[7:10], s.scales[0], rtol=3e-7)
    assert np.array_equal(dec[14:59], s.shs[0, 3:])  # mix(c, c, 0) = c exactly


def test_quat_iem_corner_code_closed_form():
    import orc
    s = _scene(n=2)
    enc = G.encode_quantized(s)
    g = enc["geometry"].copy()
    g[0, 2] = (g[0, 2] & 0xF) | (1023 << 4) | (1023 << 14) | (0xFF << 24)
    g[0, 3] = (g[0, 3] & ~np.uint32(3)) | 3
    enc2 = dict(enc, geometry=g)
    q = orc.quantized_decode(enc2)[0, 3:7].astype(np.float64)
    d = 3.0
    t = 1 / math.sqrt(d + 1e-8)
    a = 1.57079632679 * t * d
    want = np.array([math.sin(a) * t] * 3 + [math.cos(a)])
    assert np.allclose(q, want, rtol=0, atol=4e-7), (q, want)


def test_sloan_basis_equals_3dgs_basis():
    import orc
    rng = np.random.default_rng(7)
    for _ in range(200):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = rng.normal(0, 0.3, 48).astype(np.float32)
        a = orc.sh_eval(d, c, fp64=True, compact=True)
        b = orc.sh_eval(d, c, fp64=True, compact=False)
        assert np.allclose(a, b, rtol=0, atol=5e-7), (a, b)
        f = orc.sh_eval(d, c, fp64=False, compact=True)
        assert np.allclose(f, a, rtol=0, atol=3e-6), (f, a)


def test_encoder_rejects_bad_input():
    s = _scene(n=10)
    with pytest.raises(G.GpsError):
        G.Scene.empty(10, 3, G.FORMAT_QUANTIZED).upload_range(0, s)  # no scene box yet


def test_box_of_parts_is_box_of_whole():
    """Streamed uploads take the scene box from per-chunk boxes: merging them
    gives exactly the box of the whole scene."""
    s = _scene(n=5000)
    parts = [slice(0, 1200), slice(1200, 3900), slice(3900, 5000)]
    boxes = [G.quantized_bounds(G.SceneArrays(*(np.ascontiguousarray(getattr(s, k)[p]) for k in
                                                ("means", "rotations", "scales", "opacities", "shs")), s.sh_degree))
             for p in parts]
    m = G.merge_quantized_bounds(G.merge_quantized_bounds(boxes[0], boxes[1]), boxes[2])
    assert np.array_equal(m.as_array(), G.quantized_bounds(s).as_array())

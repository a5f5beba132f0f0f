"""Pins for the CPU oracle, each from something the mathematics or a
published algorithm fixes rather than a retyped formula:

* makeSeed / hash32 (lowbias32) inverted with the published inverse constants
* pcg2d uniformity, u = x * 2^-32 endpoint
* dilog against the Li2 power series; inv_dilog as its inverse
* corrected Box-Muller against its defining CDF (Li2(alpha*a) = (1-u) Li2(alpha))
* Poisson QN3 mean / variance; stochastic rounding expectation
* framebuffer addressing: Morton bijection, tile order; depth quantization
  closed forms (including the far-plane wrap)
* workload distribution: scatter + max-scan == searchsorted (independent algorithm)
* projection / EWA: isotropic closed form, FP32 vs FP64 within 1e-5
* SH basis orthonormality (Monte Carlo), real SH constants
* Hi-Z pyramid against numpy max pooling; occlusion query conservativeness
"""
import ctypes as C

import numpy as np
import pytest

import gps_amd as G
import helpers as H
import orc

MASK = 0xFFFFFFFF


def inv_hash32(x: int) -> int:
    """Inverse of lowbias32 (modular inverses 0x1d69e2a5 = 0x7feb352d^-1,
    0x43021123 = 0x846ca68b^-1 mod 2^32)."""
    x ^= x >> 16
    x = (x * 0x43021123) & MASK
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & MASK
    x ^= x >> 16
    return x


def probe(idx, frame, draws):
    out = np.empty(2 + 2 * draws, np.uint32)
    orc.lib.orc_rng_probe(idx, frame, draws, out.ctypes.data_as(C.c_void_p))
    return out


def test_make_seed_inverts_to_inputs():
    rng = np.random.default_rng(0)
    for idx, frame in zip(rng.integers(0, 2 ** 32, 300), rng.integers(0, 2 ** 32, 300)):
        idx, frame = int(idx), int(frame)
        h1p, h2p = (int(v) for v in probe(idx, frame, 0)[:2])
        h2 = h2p ^ ((h1p * 0xA54FF53A) & MASK)
        h1 = h1p ^ ((h2 * 0x3C6EF372) & MASK)
        assert inv_hash32(h1) == (idx ^ ((frame * 0x9E3779B9) & MASK))
        assert inv_hash32(h2) == (((idx * 0xBB67AE85) & MASK) ^ frame)


def test_pcg2d_uniform():
    draws = np.stack([probe(i, 17, 8)[2:].astype(np.float64) for i in range(4000)]).reshape(-1)
    u = draws / 2.0 ** 32
    hist, _ = np.histogram(u, bins=32, range=(0, 1))
    exp = u.size / 32
    chi2 = float(((hist - exp) ** 2 / exp).sum())
    assert chi2 < 70.0, chi2  # 31 dof, p ~ 1e-4
    assert abs(u.mean() - 0.5) < 0.005
    assert np.float32(np.float32(0xFFFFFFFF) * np.float32(2.3283064365387e-10)) == 1.0


def li2(x):
    """Dilogarithm: power series for x <= 1/2, Euler's reflection
    Li2(x) = pi^2/6 - ln(x) ln(1-x) - Li2(1-x) above."""
    k = np.arange(1, 80, dtype=np.float64)

    def series(v):
        return float(np.sum(v ** k / k ** 2))

    out = []
    for v in np.atleast_1d(x).astype(np.float64):
        if v <= 0.5:
            out.append(series(v))
        elif v >= 1.0:
            out.append(np.pi ** 2 / 6)
        else:
            out.append(np.pi ** 2 / 6 - np.log(v) * np.log1p(-v) - series(1.0 - v))
    return np.array(out)


def test_dilog_fits_li2():
    x = np.linspace(0.0, 1.0, 401).astype(np.float32)
    d = np.array([orc.lib.orc_dilog(float(v)) for v in x])
    ref = li2(x.astype(np.float64))
    # the reference's degree-7 fit + (1-x)log(1-x) term: |error| < 3e-4 on [0, 1]
    assert np.max(np.abs(d - ref)) < 3e-4
    assert abs(d[-1] - np.pi ** 2 / 6) < 3e-4 and abs(d[0]) < 1e-4


def test_inv_dilog_inverts_dilog():
    x = np.linspace(0.05, 1.0, 300).astype(np.float32)
    back = np.array([orc.lib.orc_inv_dilog(orc.lib.orc_dilog(float(v))) for v in x])
    assert np.max(np.abs(back - x)) < 5e-3


@pytest.mark.parametrize("alpha", [0.1, 0.5, 0.99])
def test_corrected_box_muller_cdf(alpha):
    rng = np.random.default_rng(5)
    out = (C.c_float * 2)()
    u1 = rng.uniform(0.02, 0.98, 400)
    u2 = rng.uniform(0.0, 1.0, 400)
    li2a = li2(np.array([alpha]))[0]
    for a, b in zip(u1, u2):
        orc.lib.orc_corrected_box_muller(float(a), float(b), alpha, out)
        x, y = out[0], out[1]
        r2 = x * x + y * y
        t = li2(np.array([alpha * np.exp(-0.5 * r2)]))[0] / li2a
        assert abs(t - (1.0 - a)) < 5e-3
        ang = np.arctan2(y, x) % (2 * np.pi)
        assert min(abs(ang - 2 * np.pi * b), 2 * np.pi - abs(ang - 2 * np.pi * b)) < 1e-4 or r2 < 1e-8


@pytest.mark.parametrize("lam", [10.0, 60.0, 400.0])
def test_poisson_moments(lam):
    n = 20000
    s = np.stack([probe(i, 99, 0)[:2] for i in range(n)])
    k = np.array([orc.lib.orc_poisson(int(a), int(b), lam) for a, b in s], np.float64)
    assert abs(k.mean() - lam) < 4 * np.sqrt(lam / n) + 0.05
    assert abs(k.var() / lam - 1.0) < 0.06


def test_stochastic_round():
    sr = orc.lib.orc_stochastic_round
    assert sr(2.0, 0.0) == 2 and sr(2.0, 0.999) == 2
    assert sr(2.3, 0.2) == 3 and sr(2.3, 0.4) == 2
    assert sr(-1.5, 0.1) == 0
    u = np.random.default_rng(6).uniform(0, 1, 20000)
    assert abs(np.mean([sr(3.25, float(v)) for v in u]) - 3.25) < 0.02


def test_framebuffer_addressing():
    p2m = orc.lib.orc_pos_to_memory
    assert [p2m(x, y, 0, 1, 64, 64) for x, y in ((0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (63, 63))] == \
        [0, 1, 2, 3, 4, 4095]
    # Morton code is a bijection of a 64x64 tile
    codes = {p2m(x, y, 0, 1, 64, 64) for x in range(64) for y in range(64)}
    assert codes == set(range(4096))
    # tiles in row-major order, layers innermost
    assert p2m(64, 0, 0, 1, 200, 100) == 4096
    assert p2m(0, 64, 0, 1, 200, 100) == 4 * 4096  # 4 tiles per row at width 200
    assert p2m(5, 7, 2, 3, 200, 100) == 3 * p2m(5, 7, 0, 1, 200, 100) + 2
    # brute force: every (x, y, layer) of a 130x70 buffer with 2 layers maps to a distinct word
    w, h, L = 130, 70, 2
    idx = {p2m(x, y, l, L, w, h) for x in range(w) for y in range(h) for l in range(L)}
    assert len(idx) == w * h * L and max(idx) < 3 * 2 * 4096 * L


def test_depth_quantization():
    d = orc.lib.orc_depth_bits
    n, f = 0.2, 1e4
    assert d(-n, n, f) == 0
    assert d(-f, n, f) == 0  # float(2^28-1) == 2^28: the far plane wraps to 0
    assert d(-(n + 0.5 * (f - n)), n, f) == 2 ** 27
    assert d(-1e6, n, f) == 0 and d(-0.1, n, f) == 0
    zs = -np.linspace(0.3, 9000, 500).astype(np.float32)
    q = [d(float(z), n, f) for z in zs]
    assert all(a <= b for a, b in zip(q, q[1:]))


def test_distribution_equals_searchsorted():
    arrays = H.scene(0x5EED0401, 30_000, H.C1)
    st = G.RenderSettings(320, 200, 2, 2, fully_disable_hierarchical_culling=True)
    o = orc.OraclePass(st, arrays, threads=4)
    f = G.frame_params((0, 0, -30), (0, 0, 0), (0, 1, 0), 45.0, 320, 200, ss=2)
    o.render(f)
    w = o.debug(G.DBG_WEIGHTS)
    ids = o.debug(G.DBG_VALID_IDS)
    offs = o.debug(G.DBG_VALID_OFFSETS)
    ind = o.debug(G.DBG_INDICES)
    assert np.array_equal(ids, np.flatnonzero(w))
    assert np.array_equal(offs, np.concatenate([[0], np.cumsum(w[ids])[:-1]]).astype(np.uint32))
    p = np.arange(ind.size)
    assert np.array_equal(ind, ids[np.searchsorted(offs, p, side="right") - 1])
    mask = o.debug(G.DBG_MASK_CURRENT)
    bits = np.unpackbits(mask.view(np.uint8), bitorder="little")[:w.size]
    assert np.array_equal(bits.astype(bool), w > 0)


def one_gaussian(mean, scale, rot=(1.0, 0.0, 0.0, 0.0), sh=None):
    sh = np.zeros((1, 3), np.float32) if sh is None else sh
    return G.SceneArrays(np.array([mean], np.float32), np.array([rot], np.float32), np.array([scale], np.float32),
                         np.array([0.8], np.float32), sh, 0 if sh.shape[1] == 3 else 3)


def test_projection_isotropic_closed_form():
    W, Hh, ss, dist, s = 640, 480, 2, 25.0, 0.3
    rot = np.random.default_rng(8).normal(size=4)
    rot /= np.linalg.norm(rot)
    arrays = one_gaussian((0.0, 0.0, 0.0), (s, s, s), tuple(rot))
    st = G.RenderSettings(W, Hh, 1, ss, fully_disable_hierarchical_culling=True)
    o = orc.OraclePass(st, arrays)
    f = G.frame_params((0, 0, -dist), (0, 0, 0), (0, 1, 0), 45.0, W, Hh, ss=ss)
    r = o.project(f, fp64=True)[0]
    focal_y = Hh * ss / (2 * np.tan(np.radians(45.0) / 2))
    assert r[0] == 1.0 and abs(r[1] + dist) < 1e-9
    assert abs(r[2] - W * ss / 2) < 1e-9 and abs(r[3] - Hh * ss / 2) < 1e-9
    expect = (focal_y * s / dist) ** 2 + 0.3 * ss * ss
    assert abs(r[4] / expect - 1) < 1e-7 and abs(r[6] / expect - 1) < 1e-7 and abs(r[5]) < 1e-6 * expect


@pytest.mark.parametrize("cfg,ss", [(H.C1, 1), (H.C2, 2)])
def test_projection_fp32_vs_fp64(cfg, ss):
    arrays = H.scene(0x5EED0402, 20_000, cfg, clusters=50)
    W, Hh = 480, 270
    st = G.RenderSettings(W, Hh, 1, ss, fully_disable_hierarchical_culling=True)
    o = orc.OraclePass(st, arrays, threads=4)
    eye = (0, 0, -30) if cfg is H.C1 else (0, 100, -330)
    f = G.frame_params(eye, (0, 0, 0), (0, 1, 0), 45.0, W, Hh, ss=ss)
    p32 = o.project(f, fp64=False)
    p64 = o.project(f, fp64=True)
    ok = (p32[:, 0] == 1) & (p64[:, 0] == 1)
    assert ok.sum() > 0.5 * arrays.n
    assert np.mean(p32[:, 0] == p64[:, 0]) > 0.999
    a, b = p32[ok], p64[ok]
    assert np.max(np.abs(a[:, 1] - b[:, 1]) / np.abs(b[:, 1])) < 1e-6
    scale_px = max(W, Hh) * ss
    assert np.max(np.abs(a[:, 2:4] - b[:, 2:4])) / scale_px < 1e-5
    covmax = np.maximum(np.abs(b[:, 4]), np.abs(b[:, 6]))
    assert np.max(np.abs(a[:, 4:7] - b[:, 4:7]) / covmax[:, None]) < 1e-5
    cmax = np.maximum(np.abs(b[:, 7:10]).max(axis=1), 1e-3)
    assert np.max(np.abs(a[:, 7:10] - b[:, 7:10]) / cmax[:, None]) < 1e-5
    # importance 2*pi*sqrt(det)*dilog(alpha); dilog's (1-x)log(1-x) cancels near x -> 1
    assert np.max(np.abs(a[:, 10] - b[:, 10]) / np.abs(b[:, 10])) < 1e-4


def test_sh_basis_orthonormal():
    rng = np.random.default_rng(10)
    dirs = rng.normal(size=(20000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    c = 0.05
    Y = np.empty((dirs.shape[0], 16))
    for i in range(16):
        coef = np.zeros(48, np.float32)
        coef[3 * i] = c
        for k, d in enumerate(dirs):
            Y[k, i] = (orc.sh_eval(d, coef, True)[0] - 0.5) / c
    gram = 4 * np.pi * (Y.T @ Y) / dirs.shape[0]
    assert np.max(np.abs(gram - np.eye(16))) < 0.06
    assert abs(Y[0, 0] - 0.5 / np.sqrt(np.pi)) < 1e-7  # Y00 = 1/(2 sqrt(pi))


def test_hiz_is_max_pyramid_and_query_is_conservative():
    arrays = H.scene(0x5EED0403, 20_000, dict(H.C1, log_scale_mu=-2.0))
    W, Hh = 200, 120
    st = G.RenderSettings(W, Hh, 1, 1)
    o = orc.OraclePass(st, arrays, threads=4)
    f = G.frame_params((0, 0, -30), (0, 0, 0), (0, 1, 0), 45.0, W, Hh, ss=1, occlusion=True, hierarchical=False)
    _, depth = o.render(f)
    mip = o.debug(G.DBG_HIZ)
    levels = int(np.ceil(np.log2(max(W, Hh)))) + 1
    lv = [depth]
    for _ in range(levels - 1):
        a = lv[-1]
        h2, w2 = (a.shape[0] + 1) // 2, (a.shape[1] + 1) // 2
        p = np.full((2 * h2, 2 * w2), -np.inf, np.float32)
        p[:a.shape[0], :a.shape[1]] = a
        lv.append(p.reshape(h2, 2, w2, 2).max(axis=(1, 3)))
    flat = np.concatenate([l.reshape(-1) for l in reversed(lv)])
    assert np.array_equal(flat, mip)

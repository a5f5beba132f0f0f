"""The shared deterministic scalar math (gps_detmath.h) against FP64 libm.

These functions replace CUDA's fast-math intrinsics on both the GPU and the
oracle, so their accuracy is pinned here against numpy's float64 functions
(an independent implementation), in ulps of the float32 result.
"""
import numpy as np
import pytest

import orc


def ulp_err(approx: np.ndarray, exact: np.ndarray) -> np.ndarray:
    exact32 = exact.astype(np.float32)
    spacing = np.spacing(np.abs(exact32)).astype(np.float64)
    spacing = np.where(spacing == 0, np.spacing(np.float32(0)), spacing)
    return np.abs(approx.astype(np.float64) - exact) / spacing


def test_exp_accuracy():
    rng = np.random.default_rng(1)
    x = rng.uniform(-87.0, 88.0, 400_000).astype(np.float32)
    x = np.concatenate([x, rng.uniform(-1, 1, 100_000).astype(np.float32),
                        np.array([0.0, -0.0, 1.0, -1.0, 1e-30, -1e-30], np.float32)])
    e = ulp_err(orc.detmath(0, x), np.exp(x.astype(np.float64)))
    assert e.max() <= 2.0, e.max()


def test_exp_special_values():
    x = np.array([0.0, np.nan, -200.0, -104.0, 89.0, np.inf, -np.inf, -103.0], np.float32)
    r = orc.detmath(0, x)
    assert r[0] == 1.0 and np.isnan(r[1]) and r[2] == 0.0 and r[3] == 0.0
    assert np.isinf(r[4]) and np.isinf(r[5]) and r[6] == 0.0
    assert 0 < r[7] < 1e-44  # subnormal result kept, not flushed


def test_log_accuracy():
    rng = np.random.default_rng(2)
    x = np.exp(rng.uniform(np.log(1e-37), np.log(1e37), 400_000)).astype(np.float32)
    x = np.concatenate([x, rng.uniform(0.5, 2.0, 200_000).astype(np.float32),
                        np.array([1.0, 2.0, 0.5, 1e-37, 1.5e-45, 3.4e38], np.float32)])
    exact = np.log(x.astype(np.float64))
    approx = orc.detmath(1, x)
    e = ulp_err(approx, exact)
    # log(1 +- tiny) is tiny: compare those relative to ulp(1) instead
    near1 = np.abs(exact) < 1e-3
    assert e[~near1].max() <= 2.0, e[~near1].max()
    assert np.max(np.abs(approx[near1] - exact[near1])) < 1e-9
    assert orc.detmath(1, np.array([1.0], np.float32))[0] == 0.0


def test_log_special_values():
    r = orc.detmath(1, np.array([0.0, -1.0, np.inf, np.nan], np.float32))
    assert np.isneginf(r[0]) and np.isnan(r[1]) and np.isposinf(r[2]) and np.isnan(r[3])


@pytest.mark.parametrize("kind,fn", [(2, np.sin), (3, np.cos)])
def test_sincos_accuracy(kind, fn):
    rng = np.random.default_rng(3)
    # the renderer only evaluates 2*pi*u, u in [0, 1] (and quaternion angles < 3)
    x = rng.uniform(-7.0, 7.0, 400_000).astype(np.float32)
    exact = fn(x.astype(np.float64))
    approx = orc.detmath(kind, x)
    abs_err = np.abs(approx.astype(np.float64) - exact)
    # absolute error bound: 2 ulp of 1.0 (zeros of sin/cos make relative error meaningless)
    assert abs_err.max() <= 2 * 2.0 ** -24 * 2, abs_err.max()
    big = np.abs(exact) > 0.25
    assert ulp_err(approx[big], exact[big]).max() <= 2.0


def test_sincos_special_values():
    s = orc.detmath(2, np.array([0.0, np.pi / 2, np.pi], np.float32))
    c = orc.detmath(3, np.array([0.0, np.pi / 2, np.pi], np.float32))
    assert s[0] == 0.0 and c[0] == 1.0
    assert abs(s[1] - 1.0) < 1e-7 and abs(c[2] + 1.0) < 1e-7


def test_saturating_round_to_int():
    x = np.array([2.5, 3.5, -2.5, 1e10, -1e10, np.nan, 0.49999997, 2147483520.0], np.float32)
    r = orc.detmath(8, x)
    assert list(r[:6]) == [2.0, 4.0, -2.0, 2147483647.0, -2147483648.0, 0.0]
    assert r[6] == 0.0 and r[7] == 2147483520.0


def test_sqrt_and_reciprocal_correctly_rounded():
    rng = np.random.default_rng(4)
    x = np.exp(rng.uniform(-80, 80, 200_000)).astype(np.float32)
    assert np.array_equal(orc.detmath(4, x), np.sqrt(x))
    assert np.array_equal(orc.detmath(5, x), np.float32(1.0) / x)

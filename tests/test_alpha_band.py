"""Two shortcuts of the splat's per-sample tests against the exact forms, on
the CPU: the banded alpha test, and the pixel bounds test in float.

The splat decides a sample's alpha test (opacity * exp(sigma) < 1/255 rejects,
gaussian_point_splatting.cu:118-123) from the exponent alone when it lies
outside a band of 2^-8 around t = ln((1/255) / opacity), and evaluates the
exponential only inside the band.  alpha_band_check.cpp compiles the
renderer's own header (gps_rng.h) with g++ and compares the two decisions on
random opacities and exponents, concentrated around the threshold and on
consecutive floats across the band edges: no disagreement is allowed, and
with the band set to zero the harness must find disagreements (the exact
threshold differs from t by the rounding of exp and of the product).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("alpha_band") / "check")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-Wno-unknown-pragmas",
           "-I", os.path.join(ROOT, "gaussian-point-splatting_amd", "csrc"),
           os.path.join(ROOT, "tests", "alpha_band_check.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return dict((k, int(v)) for k, v in re.findall(r"(\w+) (\d+)", r.stdout.strip().splitlines()[-1]))


@pytest.mark.parametrize("seed", [1, 2])
def test_banded_alpha_test_equals_exact(checker, seed):
    m = run(checker, 200_000, seed)
    assert m["mismatches"] == 0, m
    assert m["checked"] > 600_000 and 0 < m["in_band"] < m["checked"]


def test_harness_detects_a_zero_band(checker):
    assert run(checker, 200_000, 3, 0.0)["mismatches"] > 0


@pytest.fixture(scope="module")
def pixel_checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("pixel_bounds") / "check")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-Wno-unknown-pragmas",
           "-I", os.path.join(ROOT, "gaussian-point-splatting_amd", "csrc"),
           os.path.join(ROOT, "tests", "pixel_bounds_check.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def test_pixel_bounds_test_equals_conversion(pixel_checker):
    """The splat's pixel bounds test in float (det_pixel_in) makes the same
    decision and gives the same pixel as converting with det_f2i_rz and
    comparing integers (NaN converts to 0; infinities and out-of-range values
    saturate), for every integral float in [-70000, 70000] and the special
    values, against tile-aligned bounds up to 65536."""
    m = run(pixel_checker)
    assert m["mismatches"] == 0 and m["checked"] > 1_400_000, m

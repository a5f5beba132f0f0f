"""The renderer's projection (gps_project.h) against the oracle's, on the CPU.

project_gaussian leaves out the products with exact zeros of the reference's
M = R * S (S diagonal) and T = W^T * J (two zero Jacobian entries), which can
only change the sign of a zero intermediate.  projection_check.cpp compiles
the renderer's header with g++ and writes its projection of adversarial
Gaussians; every value (validity, view z, pixel mean, dilated 2D covariance,
NDC z, tile bounds) must equal the oracle's, which keeps the reference's
sums with their zero products, bit for bit: axis-aligned cameras (view
matrices with exact zeros), quaternions with zero and negative-zero
components, zero, tiny and large scales, and random Gaussians.
"""
import os
import subprocess

import numpy as np
import pytest

import gps_amd as G
import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def projector(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("projection") / "project")
    libdir = os.path.dirname(G.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-Wno-unknown-pragmas",
           "-I", os.path.join(ROOT, "gaussian-point-splatting_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "projection_check.cpp"), "-L", libdir, "-lgps_amd",
           "-Wl,-rpath," + libdir, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def adversarial_scene(seed: int, n_random: int) -> G.SceneArrays:
    rng = np.random.default_rng(seed)
    h = np.sqrt(0.5)
    quats = [(1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1), (h, h, 0, 0), (h, 0, h, 0), (h, 0, 0, h),
             (0, h, h, 0), (-0.0, 1, -0.0, 0), (1, -0.0, 0, -0.0), (0.5, 0.5, 0.5, 0.5), (0.5, -0.5, 0.5, -0.5),
             (2, 0, 0, 0), (0, 0, 0, 0), (0.9, 0, 0, 0)]
    scales = [(0.05, 0.05, 0.05), (0.5, 0.01, 0.001), (0, 0.1, 0.1), (0, 0, 0), (1e-20, 0.2, 3.0),
              (-0.1, 0.2, 0.05), (3.0, 3.0, 0.0), (1e-3, 1e-3, 1e-3)]
    means = [(0, 0, 0), (1.5, -2.0, 0.25), (0, 0, 5), (-3, 0, 0), (0.0, -0.0, 1.0)]
    rows = [(m, q, s) for m in means for q in quats for s in scales]
    m = np.array([r[0] for r in rows], np.float32)
    q = np.array([r[1] for r in rows], np.float32)
    s = np.array([r[2] for r in rows], np.float32)
    qr = rng.normal(size=(n_random, 4)).astype(np.float32)
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    qr[rng.random((n_random, 4)) < 0.15] = 0.0  # exact zero components
    m = np.concatenate([m, rng.uniform(-8, 8, (n_random, 3)).astype(np.float32)])
    q = np.concatenate([q, qr])
    s = np.concatenate([s, np.exp(rng.normal(-2.5, 1.5, (n_random, 3))).astype(np.float32)])
    n = m.shape[0]
    return G.SceneArrays(np.ascontiguousarray(m), np.ascontiguousarray(q), np.ascontiguousarray(s),
                         np.full(n, 0.5, np.float32), np.zeros((n, 3), np.float32), 0)


CAMERAS = [((0.0, 0.0, -30.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)),     # axis aligned
           ((30.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)),      # axis aligned, looking -x
           ((0.0, 25.0, 0.0001), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)),   # looking down
           ((7.0, -4.0, -18.0), (0.5, 0.3, 0.0), (0.1, 1.0, 0.05))]   # generic


@pytest.mark.parametrize("cam", range(len(CAMERAS)))
@pytest.mark.parametrize("ss", [1, 2])
def test_projection_equals_oracle_bit_for_bit(projector, tmp_path, cam, ss):
    arrays = adversarial_scene(cam, 3000)
    eye, center, up = CAMERAS[cam]
    w, h = 320, 200
    f = G.frame_params(eye, center, up, 50.0, w, h, ss=ss)
    settings = G.RenderSettings(w, h, 1, ss)
    o = orc.OraclePass(settings, arrays, G.FORMAT_FP32, 4)
    ref = o.project(f, fp64=False)
    o.close()
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        fh.write(bytes(f))
        fh.write(np.uint64(arrays.n).tobytes())
        fh.write(np.concatenate([arrays.means, arrays.rotations, arrays.scales], axis=1).astype(np.float32).tobytes())
    subprocess.run([projector, str(inp), str(out)], check=True)
    mine = np.fromfile(out, np.float32).reshape(-1, 12)
    valid = ref[:, 0] != 0
    assert np.array_equal(valid, mine[:, 0] != 0)
    assert valid.sum() > 1000
    cols_ref = ref[valid][:, [1, 2, 3, 4, 5, 6, 11, 12, 13, 14, 15]].astype(np.float32)
    cols_mine = mine[valid][:, 1:]
    bad = cols_ref.view(np.uint32) != cols_mine.view(np.uint32)
    assert not bad.any(), (np.argwhere(bad)[:5], cols_ref[bad][:5], cols_mine[bad][:5])

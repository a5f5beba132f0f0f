"""The C++ RenderPass adapter a maintainer adds to the reference
(tests/adapter/gps_render_pass.h, shown verbatim in the integration guide), compiled
against include/gps.h and run as the reference's RenderServer would run it.

The frame carries a non-identity scene transform folded into
cameraViewMatrix the way main.cu:148 folds it (view = flip * camera view *
sceneMatrix); the adapter must pass that matrix on unchanged, so its images
must equal the oracle's rendering of the same matrices bit for bit, and its
Statistics must be the reference's (gaussian_point_splatting.cu:675-685).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import gps_amd as G
import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADAPTER = os.path.join(ROOT, "tests", "adapter")


def test_integration_doc_shows_the_tested_adapter():
    """The integration guide's adapter listing is the compiled file, line for line."""
    with open(os.path.join(ROOT, "INTEGRATION.md")) as fh:
        doc = fh.read()
    with open(os.path.join(ADAPTER, "gps_render_pass.h")) as fh:
        src = fh.read().strip()
    blocks = re.findall(r"```cpp\n(.*?)```", doc, re.S)
    assert any(b.strip() == src for b in blocks), "the integration guide does not show tests/adapter/gps_render_pass.h"
    assert os.path.exists(os.path.join(ADAPTER, "adapter_demo")), "build() compiles the adapter demo"


def scene_matrix(translation, axis, angle_deg, scale) -> np.ndarray:
    """T * R * S (Transform::getModelMatrix order), row-major float32."""
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    t = np.radians(angle_deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * (K @ K)
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(scale)[None, :]
    M[:3, 3] = translation
    return M.astype(np.float32)


def cpu_gaussians(arrays) -> np.ndarray:
    """CPUGaussian records (gaussian_types.cuh:89-100): mean, sh[48],
    opacity, scale, rotation (w, x, y, z) = 59 floats."""
    n = arrays.n
    rec = np.zeros((n, 59), np.float32)
    rec[:, 0:3] = arrays.means
    nc = (arrays.sh_degree + 1) ** 2
    rec[:, 3:3 + 3 * nc] = arrays.shs.reshape(n, 3 * nc)
    rec[:, 51] = arrays.opacities
    rec[:, 52:55] = arrays.scales
    rec[:, 55:59] = arrays.rotations
    return rec


@pytest.mark.gpu
@pytest.mark.parametrize("occl", [False, True])
def test_adapter_renders_scene_transform_like_the_oracle(gpu, tmp_path, occl):
    import orc
    H.flush(f"adapter occl={occl}")
    arrays = H.scene(0x5EEDADA0 + int(occl), 20_000, H.C2, clusters=60)
    w, h, spp, ss = 240, 160, 2, 2
    base = G.frame_params((40.0, 60.0, -330.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, w, h, ss=ss, frame=3,
                          occlusion=occl, hierarchical=occl, spp=spp)
    V = np.array(base.view, np.float32).reshape(4, 4).T
    S = scene_matrix((12.0, -7.0, 30.0), (0.3, 1.0, -0.2), 33.0, (1.3, 0.8, 1.1))
    view = (V.astype(np.float64) @ S.astype(np.float64)).astype(np.float32)
    col_major = np.ascontiguousarray(view.T).reshape(-1)
    planes = np.empty(24, np.float32)
    G._lib.gps_frustum_planes(45.0, col_major.ctypes.data_as(G._fp), w / h, 0.2, 1e4,
                              planes.ctypes.data_as(G._fp))
    for i in range(16):
        base.view[i] = float(col_major[i])
    for i in range(24):
        base.frustum_planes[i] = float(planes[i])
    hdr = np.array([arrays.n, arrays.sh_degree, w, h, spp, ss, 3, int(occl), int(occl), 1], np.int32)
    mats = np.concatenate([col_major, np.array(base.proj, np.float32), planes,
                           np.array([45.0, 0.2, 1e4], np.float32)])
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        fh.write(hdr.tobytes() + mats.astype(np.float32).tobytes() + cpu_gaussians(arrays).tobytes())
    r = subprocess.run([os.path.join(ADAPTER, "adapter_demo"), str(inp), str(out)], capture_output=True, text=True,
                       timeout=120)
    H.flush(r.stdout + r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(out, np.uint8)
    color = raw[:w * h * 12].view(np.float32).reshape(h, w, 3)
    depth = raw[w * h * 12:w * h * 16].view(np.float32).reshape(h, w)
    st = raw[w * h * 16:].view(np.int32)
    total_weight = float(raw[w * h * 16:w * h * 16 + 4].view(np.float32)[0])
    o = orc.OraclePass(G.RenderSettings(w, h, spp, ss, fully_disable_hierarchical_culling=not occl), arrays,
                       G.FORMAT_FP32, 8)
    oc, od = o.render(base)
    os_ = o.stats()
    H.assert_same("adapter colour", color, oc)
    H.assert_same("adapter depth", depth, od)
    assert sum(os_["num_points"][:os_["passes"]]) > 1000
    # Statistics: numPointsOne / Two, numGaussiansOne (= scene size), totalWeight
    assert st[2] == os_["num_points"][0] and st[6] == arrays.n
    if occl:
        assert os_["passes"] == 2 and st[3] == os_["num_points"][1] and st[7] == arrays.n
    assert total_weight == float(os_["num_points"][os_["passes"] - 1])
    # the scene transform matters: the same frame without it differs
    plain = G.frame_params((40.0, 60.0, -330.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, w, h, ss=ss, frame=3,
                           occlusion=occl, hierarchical=occl, spp=spp)
    pc, _ = orc.OraclePass(G.RenderSettings(w, h, spp, ss, fully_disable_hierarchical_culling=not occl), arrays,
                           G.FORMAT_FP32, 8).render(plain)
    assert not np.array_equal(pc, oc)

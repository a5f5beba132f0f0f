"""Load-time BVH order on the GPU (gps_bvh_order) against the oracle
(orc_bvh.cpp, pinned to the reference's tinybvh in test_bvh_order.py) and
against the tinybvh golden leaves.  Integer output: bit-exact.
"""
import os

import numpy as np
import pytest

import bvh_scenes as B
import orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bvh_tinybvh_leaves.npz")


def _gpu_order(G, m, q, s):
    a = G.SceneArrays(m, q, s, np.zeros(len(m), np.float32), np.zeros((len(m), 3), np.float32), 0)
    return a.bvh_order()


@pytest.mark.parametrize("kind", B.KINDS)
@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 64, 1000, 49999, 50000, 300000])
def test_bvh_order_matches_oracle(gpu, kind, n):
    m, q, s = B.scene(n, 1000 + n, kind)
    print(f"bvh_order kind={kind} n={n}", flush=True)
    got = _gpu_order(gpu, m, q, s)
    want = orc.bvh_order(m, q, s)
    assert np.array_equal(got, want), (kind, n, int((got != want).sum()))


def test_bvh_order_matches_tinybvh_golden(gpu):
    g = np.load(GOLDEN)
    for key in g.files:
        kind, n, seed = key.rsplit("_", 2)
        n, seed = int(n), int(seed)
        m, q, s = B.scene(n, seed, kind)
        count = g[key].astype(np.int64)
        first = np.concatenate([[0], np.cumsum(count)[:-1]])
        print(f"bvh_order golden {key}", flush=True)
        assert np.array_equal(_gpu_order(gpu, m, q, s), B.perm_from_leaves(n, first, count)), key


def test_bvh_order_large_clusters(gpu):
    # 4M Gaussians: many levels of the level-synchronous build
    n = 4_000_000
    m, q, s = B.scene(n, 77, "clusters")
    print(f"bvh_order clusters n={n}", flush=True)
    got = _gpu_order(gpu, m, q, s)
    want = orc.bvh_order(m, q, s)
    assert np.array_equal(got, want), int((got != want).sum())


def test_bvh_order_synthetic_scene(gpu):
    # the renderer's own synthetic Gaussians (Gaussian-distributed positions,
    # log-normal scales, random rotations)
    p = gpu.synth_params(0x5EED0042, 200_000, kind="clusters", clusters=16, half_extent=20.0)
    a = gpu.synth_scene(p)
    print("bvh_order synth n=200000", flush=True)
    assert np.array_equal(a.bvh_order(), orc.bvh_order(a.means, a.rotations, a.scales))


def test_cli_ply_bvh_reorder(gpu, tmp_path):
    """The headless driver on a PLY without "bvh_sorted_order": it reorders
    on load (gaussian_point_splatting.cu:791-797), and
    --sorted-model-output-path writes the reordered scene with the comment
    (main.cu:266-273); a file carrying the comment is not reordered again."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "gaussian-point-splatting_amd", "gps_render")
    p = gpu.synth_params(0x5EED0077, 20_000, kind="clusters", clusters=8, sh_degree=3, half_extent=10.0)
    src = str(tmp_path / "scene.ply")
    gpu.write_ply(src, gpu.synth_scene(p))
    loaded = gpu.load_ply(src)
    perm = orc.bvh_order(loaded.means, loaded.rotations, loaded.scales)
    out = str(tmp_path / "sorted.ply")
    print("cli: ply bvh reorder", flush=True)
    r = subprocess.run([exe, "--model-path", src, "--sorted-model-output-path", out, "--resolution", "160", "90",
                        "--frames", "1", "--out", str(tmp_path / "a")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "is pre-sorted: false" in r.stdout
    assert gpu.ply_info(out)["bvh_sorted"]
    assert np.array_equal(gpu.load_ply(out).means, loaded.means[perm])
    r = subprocess.run([exe, "--model-path", out, "--resolution", "160", "90", "--frames", "1",
                        "--out", str(tmp_path / "b")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "is pre-sorted: True" in r.stdout

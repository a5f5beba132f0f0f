"""Writes the golden fixtures of the CPU oracle (tests/golden/*.npz).

The reference repository holds no golden vectors for this path, so these
fixtures are produced by the oracle after it passed its pinning tests
(tests/test_oracle_pinning.py); they pin the oracle (and through the GPU
tests, the kernels) against accidental change.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "gaussian-point-splatting_amd")]

import gps_amd as G  # noqa: E402
import helpers as H  # noqa: E402

CASES = {
    "c1_tiny": dict(seed=0x5EED1001, n=3000, cfg=H.C1, W=96, H=64, ss=1, spp=1, eye=(0.0, 0.0, -30.0),
                    frames=1, occlusion=False, hier=False, fmt=0),
    "c2_tiny_culling": dict(seed=0x5EED1002, n=3000, cfg=H.C2, W=64, H=48, ss=2, spp=2, eye=(0.0, 60.0, -250.0),
                            frames=2, occlusion=True, hier=True, fmt=0),
    "c2_tiny_half_fp8": dict(seed=0x5EED1003, n=3000, cfg=H.C2, W=64, H=48, ss=2, spp=1, eye=(0.0, 60.0, -250.0),
                             frames=1, occlusion=False, hier=False, fmt=1),
}


def render_case(name, runner_cls):
    c = CASES[name]
    arrays = H.scene(c["seed"], c["n"], c["cfg"], clusters=20)
    st = G.RenderSettings(c["W"], c["H"], c["spp"], c["ss"], fully_disable_hierarchical_culling=not c["hier"],
                          gaussians_per_block=16)
    r = runner_cls(arrays, st, c["fmt"])
    out = {}
    for i in range(c["frames"]):
        f = G.frame_params(c["eye"], (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, c["W"], c["H"], ss=c["ss"],
                           frame=i + 1, occlusion=c["occlusion"], hierarchical=c["hier"])
        res = r.render(f, [G.DBG_WEIGHTS, G.DBG_INDICES, G.DBG_FRAMEBUFFER])
        out[f"f{i}_color"] = res.color
        out[f"f{i}_depth"] = res.depth
        out[f"f{i}_weights"] = res.debug[G.DBG_WEIGHTS]
        out[f"f{i}_indices"] = res.debug[G.DBG_INDICES]
        out[f"f{i}_fb"] = res.debug[G.DBG_FRAMEBUFFER]
    return out


def main():
    for name in CASES:
        out = render_case(name, H.OracleRunner)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()

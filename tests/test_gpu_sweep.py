"""Seeded sweep of small two-frame sequences, GPU against the oracle bit for
bit (framebuffer, weights, masks, compacted ids and offsets, point indices,
colour, depth): supersampling 1-4 (3 takes the division path of the
supersampled bounds, 1, 2 and 4 the shifts), 1-4 samples per pixel, with
and without occlusion + block culling, uniform and clustered scenes, so
both splat forms (point-parallel with its LDS word table below 8 samples
per point, binned at 8 and above) run under varied loads.
"""
import numpy as np
import pytest

import helpers as H
from test_gpu_parity import moving_frames, run_pair

pytestmark = pytest.mark.gpu

# (supersampling, samples per pixel): point-parallel below 8 samples per
# point (ss 2 spp 1 is the 100M fly-through's setting), binned above.
SS_SPP = [(2, 1), (2, 1), (1, 1), (1, 3), (2, 1), (3, 1), (1, 4), (2, 1),
          (4, 1), (2, 2), (3, 3), (2, 1), (4, 4), (1, 2), (3, 2), (2, 1)]
CASES = list(range(len(SS_SPP)))


@pytest.mark.parametrize("case", CASES)
def test_sweep_frames(gpu, case):
    r = np.random.default_rng(0x5EED5000 + case)
    ss, spp = SS_SPP[case]
    culled = bool(case % 2)
    clustered = bool((case // 2) % 2)
    n = int(r.integers(5_000, 40_000))
    w, h = int(r.integers(120, 260)), int(r.integers(80, 180))
    cfg = dict(H.C2, half_extent=8.0, log_scale_mu=-2.5) if clustered else dict(H.C1, half_extent=8.0,
                                                                                    log_scale_mu=-2.5)
    arrays = H.scene(0x5EED5100 + case, n, cfg, clusters=40)
    H.flush(f"sweep case {case}: n={n} {w}x{h} ss={ss} spp={spp} culled={culled} clustered={clustered}")
    if culled:
        st = gpu.RenderSettings(w, h, spp, ss, gaussians_per_block=64)
    else:
        st = gpu.RenderSettings(w, h, spp, ss, fully_disable_hierarchical_culling=True)
    frames = moving_frames(2, w, h, ss, radius=25.0, occlusion=culled, hierarchical=culled)
    which = H.INTERMEDIATES + ([H.G.DBG_MASK_PREVIOUS, H.G.DBG_HIZ, H.G.DBG_CULL_MASK] if culled else [])
    res = run_pair(arrays, st, frames, which=which)
    assert res[0][0].stats["K"] == ss * ss
    if ss * ss * spp < 8:
        assert res[0][0].stats["splat_binned"] == 0  # point-parallel with the word table
    assert sum(rg.stats["num_points"][0] for rg, _ in res) > 0

"""The preprocess reject-reason counters (GPS_COUNT_REJECTS, diagnostic):

* counting changes no result: frames with the counters on equal the oracle
  bit for bit (the counting form of the kernel is a separate instantiation);
* every count equals the oracle's (orc_pass_rejects), per pass and reason;
* the reasons partition the Gaussians of unculled blocks: their sum is the
  number tested, which is N minus the Gaussians of culled blocks;
* "valid" equals the pass's valid-Gaussian count, and in the second pass
  "prev_valid" equals the first pass's valid count minus those of blocks the
  second pass culls (a second pass never re-tests a first-pass survivor).
"""
import numpy as np
import pytest

import helpers as H
from helpers import assert_same, flush

pytestmark = pytest.mark.gpu


def test_reject_counters_partition_and_change_nothing(gpu):
    import orc
    from gps_amd import workloads as W
    w = W.scaled(W.WORKLOADS["c3"], 30_000, width=320, height=180, gaussians_per_block=64)
    perm = W.order(w)
    arrays = W.host_scene(w, perm)
    scene = W.device_scene(w, perm)
    loop = W.FrameLoop(w, scene)
    loop.pass_.set_counters(True, count_rejects=True)
    o = orc.OraclePass(w.settings(), arrays, w.fmt, 8)
    seen_zero = seen_second = 0
    for k, pos in enumerate([0, 1, 2, 150, 151]):
        f = W.frame(w, pos, k + 1)
        flush(f"rejects: frame {k} (camera {pos})")
        loop.step(f)
        st = loop.pass_.stats(loop.stream)
        rj = loop.pass_.rejects(loop.stream)
        fb = loop.pass_.debug(gpu.DBG_FRAMEBUFFER, loop.stream)
        o.render(f)
        assert_same(f"fb {k}", fb, o.debug(gpu.DBG_FRAMEBUFFER))
        assert rj[:st["passes"]] == o.rejects()[:st["passes"]], (k, rj, o.rejects())
        mask = loop.pass_.debug(gpu.DBG_CULL_MASK, loop.stream)
        for ps in range(st["passes"]):
            r = rj[ps]
            parts = sum(v for key, v in r.items() if key != "tested")
            assert parts == r["tested"], (k, ps, r)
            assert r["valid"] == st["valid_gaussians"][ps], (k, ps, r, st)
            if ps == 0:
                assert r["prev_valid"] == 0
            seen_zero += r["zero_count"]
        # tested = Gaussians of unculled blocks (blocks of gpb, the last one ragged)
        gpb, nb = st["gpb"], st["num_blocks"]
        bits = np.unpackbits(mask.view(np.uint8), bitorder="little")[:nb].astype(bool)
        sizes = np.full(nb, gpb, np.int64)
        sizes[-1] = w.n - gpb * (nb - 1)
        assert rj[st["passes"] - 1]["tested"] == int(sizes[~bits].sum())  # the mask is the last pass's
        if st["passes"] == 2:
            seen_second += 1
            assert rj[1]["prev_valid"] <= rj[0]["valid"]
    assert seen_zero > 0 and seen_second > 0

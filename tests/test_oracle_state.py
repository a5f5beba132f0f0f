"""The oracle's carried-state injection (orc_pass_set_state): a frame rendered
on a fresh pass that is handed the Hi-Z chain and the untouched valid mask of
another pass equals the same frame rendered in order on that pass.  This is
what lets the full-size parity check compare path positions deep into the
fly-through without the oracle rendering every frame before them.

Checked on the 100M workload's construction at test size (4 translated
copies, Morton order, compact storage, blocks + Hi-Z, two-pass culling) at
path positions in order and after a camera cut.  The mask handed over is the
one the next first pass leaves in place (the last pass's own mask when that
pass flipped the ping-pong, else the one before it): a second pass reads it
as its previous mask whenever the first pass yields no points."""
import numpy as np

import gps_amd as G
import helpers as H
import orc
from gps_amd import workloads as W


def _w():
    return W.scaled(W.WORKLOADS["c3"], 12_000, width=192, height=108, gaussians_per_block=64)


def test_injected_state_equals_rendering_in_order():
    w = _w()
    arrays = W.host_scene(w)
    a = orc.OraclePass(w.settings(), arrays, w.fmt, 4)
    seq = [0, 1, 2, 150, 151]
    fbs = []
    states = []
    for k, pos in enumerate(seq):
        states.append(H.carried_state(a))
        a.render(W.frame(w, pos, k + 1))
        fbs.append(a.debug(G.DBG_FRAMEBUFFER))
    for k in (2, 3, 4):
        b = orc.OraclePass(w.settings(), arrays, w.fmt, 4)
        hiz, other = states[k]
        b.set_state(hiz, other)
        b.render(W.frame(w, seq[k], k + 1))
        H.assert_same(f"injected frame {k}", b.debug(G.DBG_FRAMEBUFFER), fbs[k])
        b.close()
    # the fresh pass without the state renders frame 3 differently (the
    # check above is not vacuous)
    c = orc.OraclePass(w.settings(), arrays, w.fmt, 4)
    c.render(W.frame(w, seq[3], 4))
    assert not np.array_equal(c.debug(G.DBG_FRAMEBUFFER), fbs[3])


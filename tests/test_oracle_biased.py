"""Pins for the oracle's biased splat (use_unbiased_2d_splatting = 0), the
reference's pre-check + atomicAddColor (gaussian_point_splatting.cu:130-138,
post_processing_kernels.cuh:13-42), against cases the word format fixes:

* equal depth: the 36 colour bits add modulo 2^36 (a 12-bit channel that
  overflows carries into the next one; the top one's carry is dropped)
* a nearer sample replaces the word, a farther one leaves it
* the pre-check skips a sample whose depth equals a stored word of colour 0
  (atomicAddColor alone would add); with add_if_equal false atomicAddColor
  returns false at equal depth and leaves the word
* the resulting order dependence, shown on two samples
* with no zero colours and no carry out of bit 35 the word is the nearest
  depth with the sum of the colours at that depth, for every order (brute
  force over permutations)
* a frame of one large Gaussian: every touched word holds k times the
  Gaussian's packed colour (k >= 1, and k >= 2 somewhere), where the unbiased
  frame holds it once
"""
import itertools

import numpy as np

import gps_amd as G
import helpers as H
import orc

DMAX = (1 << 28) - 1
CMASK = (1 << 36) - 1


def D(d: int) -> int:
    return d << 36


def col(r: int, g: int, b: int) -> int:
    return (r << 24) | (g << 12) | b


def apply(word: int, samples, precheck=True) -> int:
    for v in samples:
        word, _ = orc.biased_sample(word, v, precheck)
    return word


def test_equal_depth_adds_with_carry_across_channels():
    w = apply(D(100) | col(1, 0xFFF, 0xFFF), [D(100) | col(0, 0, 1)])
    assert w == D(100) | col(2, 0, 0)
    w = apply(D(100) | col(3, 0x800, 0x7FF), [D(100) | col(0, 0x800, 0x001)])
    assert w == D(100) | col(4, 0x000, 0x800)


def test_top_channel_carry_is_dropped():
    w = apply(D(100) | col(0xFFF, 5, 0), [D(100) | col(1, 0, 0)])
    assert w == D(100) | col(0, 5, 0)
    w = apply(D(7) | col(0xFFF, 0xFFF, 0xFFF), [D(7) | col(0, 0, 1)])
    assert w == D(7)  # colour wraps to 0, depth kept


def test_nearer_replaces_farther_is_ignored():
    assert apply(D(200) | col(7, 7, 7), [D(100) | col(1, 2, 3)]) == D(100) | col(1, 2, 3)
    assert apply(D(100) | col(7, 7, 7), [D(200) | col(1, 2, 3)]) == D(100) | col(7, 7, 7)
    # background word (all-ones depth) is replaced by any nearer sample
    assert apply(D(DMAX), [D(DMAX - 1) | col(9, 0, 0)]) == D(DMAX - 1) | col(9, 0, 0)


def test_precheck_skips_equal_depth_on_zero_colour():
    word = D(100)
    assert apply(word, [D(100) | col(1, 2, 3)], precheck=True) == word
    w, r = orc.biased_sample(word, D(100) | col(1, 2, 3), precheck=False, add_if_equal=True)
    assert (w, r) == (D(100) | col(1, 2, 3), 1)
    w, r = orc.biased_sample(D(100) | col(4, 4, 4), D(100) | col(1, 2, 3), precheck=False, add_if_equal=False)
    assert (w, r) == (D(100) | col(4, 4, 4), 0)
    # a black background: samples at the background depth are always skipped
    assert apply(D(DMAX), [D(DMAX) | col(1, 1, 1)] * 3) == D(DMAX)
    # a coloured background: they add to it
    assert apply(D(DMAX) | col(1, 1, 1), [D(DMAX) | col(1, 0, 0)] * 2) == D(DMAX) | col(3, 1, 1)


def test_zero_colour_first_makes_the_word_order_dependent():
    a, b = D(100), D(100) | col(5, 0, 0)
    assert apply(D(DMAX), [a, b]) == D(100)
    assert apply(D(DMAX), [b, a]) == D(100) | col(5, 0, 0)


def test_order_independent_without_zero_colours_or_carry():
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 7))
        depths = rng.choice([50, 60, 70], size=n)
        cols = [col(*[int(c) for c in rng.integers(0, 200, 3)]) | 1 for _ in range(n)]
        samples = [D(int(d)) | c for d, c in zip(depths, cols)]
        start = D(DMAX) if trial % 2 == 0 else D(int(rng.choice([55, 65, 80]))) | col(1, 1, 1)
        d_star = min(int(depths.min()), start >> 36)
        total = sum(c for d, c in zip(depths, cols) if d == d_star)
        if d_star == start >> 36:
            total += start & CMASK
        expect = D(d_star) | (total & CMASK)
        assert total < (1 << 36)
        perms = list(itertools.permutations(samples))
        for k in rng.choice(len(perms), size=min(len(perms), 6), replace=False):
            assert apply(start, perms[int(k)]) == expect


def _one_gaussian_fb(unbiased: bool):
    arrays = H.scene(0x5EED0B1A, 1, H.C1)
    arrays.means[:] = 0.0
    arrays.scales[:] = 2.0
    arrays.rotations[:] = (1.0, 0.0, 0.0, 0.0)
    arrays.opacities[:] = 0.9
    st = G.RenderSettings(160, 120, 1, 1, fully_disable_hierarchical_culling=True)
    o = orc.OraclePass(st, arrays, threads=1)
    f = G.frame_params((0.0, 0.0, -30.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 160, 120, ss=1,
                       unbiased=unbiased)
    o.render(f)
    return o.debug(G.DBG_FRAMEBUFFER), o.biased_events()


def test_frame_of_one_gaussian_sums_repeated_samples():
    fb_u, _ = _one_gaussian_fb(True)
    fb_b, ev = _one_gaussian_fb(False)
    bg = np.uint64(DMAX << 36)
    hit_u = fb_u[fb_u != bg]
    assert np.unique(hit_u).size == 1
    word = int(hit_u[0])
    c = word & CMASK
    assert c != 0
    hit_b = fb_b[fb_b != bg].astype(object)
    assert hit_b.size > 50
    assert all((int(v) >> 36) == (word >> 36) for v in hit_b)
    k = np.array([(int(v) & CMASK) // c for v in hit_b])
    rem = np.array([(int(v) & CMASK) % c for v in hit_b])
    assert (rem == 0).all() and (k >= 1).all()
    assert (k >= 2).sum() > 10
    assert ev == {"zero_colour": 0, "wrapped": 0}

"""Colouring index of the Menger sponge and the Mandelbox in the oracle
(Compute<colouring>, src/fractal.cpp:1168-1189), pinned by closed forms
for points whose orbit escapes after one iteration."""
import math

import pytest

from mandelbulber_amd import scene
from oracle import oracle


def test_menger_colour_is_1000_times_the_smallest_folded_radius():
    """The Menger colour is 1000 x the smallest |z| after the abs/sort folds;
    abs and sort keep |z|, and a far orbit only grows, so for a far point it
    is 1000 |p|."""
    s = scene.mandelbulb_plain(8, 8)
    s.formula = 7
    s.N = 20
    for p in [(400.0, -150.0, 90.0), (-80.0, 700.0, 33.0), (12.0, 13.0, -500.0)]:
        v, _ = oracle.compute(s, p, mode=1)
        assert v == pytest.approx(1000.0 * math.sqrt(sum(c * c for c in p)), rel=1e-12)


def test_mandelbox_colour_counts_the_folded_axes():
    """Mandelbox colour = 100 x (1 + the colour weight of every folded axis
    and sphere fold) + |z| x mandelbox_color_R at escape.  A far point folds
    all three axes once, takes no sphere fold (r^2 above the fixed radius),
    and escapes after one iteration: z1 = scale * (val - |p_i|, signs
    flipped) + p."""
    s = scene.mandelbulb_plain(8, 8)
    scene.apply(s, scene.MANDELBOX_KEYS)
    s.mb_color_X, s.mb_color_Y, s.mb_color_Z, s.mb_color_R = 0.3, 0.5, 0.7, 0.011
    cX, cY, cZ, cR = s.mb_color_X, s.mb_color_Y, s.mb_color_Z, s.mb_color_R
    lim, val, sc = s.mb_folding_limit, s.mb_folding_value, s.mb_scale
    for p in [(1000.0, -800.0, 700.0), (-1200.0, 900.0, -650.0)]:
        z = [(val - c) if c > lim else (-val - c) for c in p]
        z1 = [sc * zc + pc for zc, pc in zip(z, p)]
        r1 = math.sqrt(sum(c * c for c in z1))
        assert r1 > 1024.0
        v, _ = oracle.compute(s, p, mode=1)
        assert v == pytest.approx((1.0 + cX + cY + cZ) * 100.0 + r1 * cR, rel=1e-12)

"""The oracle's screen-space ambient occlusion (PostRendering_SSAO /
ThreadSSAO, src/image.cpp:249-415) pinned by cases worked out by hand:
a flat depth plane (every direction's horizon is level: AO 1/2), border
pixels (a direction with no sample in the image counts as open: 1),
a one-sample step whose horizon angle is atan(4), and the pixels the
effect must leave alone (background, full opacity, the 16-bit clamp)."""
import math

import numpy as np

from mandelbulber_amd import scene as S
from oracle import oracle


def _scene(q, persp=1.0, intensity=1.0, ptype=0):
    s = S.defaults()
    s.ssao_quality = q
    s.persp = persp
    s.global_ilum = intensity
    s.perspective_type = ptype
    return s


def _planes(w, h, z, colour=(100, 50, 200)):
    zb = np.full((h, w), z, np.float32) if np.isscalar(z) else np.asarray(z, np.float32)
    op = np.zeros((h, w), np.uint16)
    col = np.stack([np.full((h, w), c, np.uint8) for c in colour])
    img = np.zeros((h, w, 3), np.uint16)
    return zb, op, col, img


def test_flat_plane_interior_is_half_open():
    w, h = 64, 48
    zb, op, col, img = _planes(w, h, 2.0)
    out = oracle.ssao(_scene(4), zb, op, col, img)
    # q = 4: samples at r^2 * w / (2 q^2) = 2, 8, 18 pixels along 0, 90, 180, 270 degrees;
    # a pixel at least 2 from every border has a sample in every direction
    inner = out[2:h - 2, 2:w - 2]
    for c, v in enumerate((100, 50, 200)):
        assert np.all(inner[..., c] == int(65535.0 * v / 256.0 * 0.5)), c


def test_border_directions_without_samples_count_as_open():
    w, h = 64, 48
    zb, op, col, img = _planes(w, h, 2.0, colour=(255, 255, 255))
    out = oracle.ssao(_scene(4), zb, op, col, img)
    one_open = int(65535.0 * 255 / 256.0 * (0.5 * 3 + 1.0) / 4)
    two_open = int(65535.0 * 255 / 256.0 * (0.5 * 2 + 2.0) / 4)
    assert out[20, 1, 0] == one_open                                      # 180 degrees open
    # column 0: cos(3 pi / 2) is -1.8e-16 in double, so the 270-degree samples land at x < 0 and
    # that direction is open too (the table holds the library values, not exact zeros)
    assert out[20, 0, 0] == two_open
    assert out[0, 30, 0] == one_open                                      # top edge: 270 degrees open
    assert out[0, 0, 0] == two_open and out[h - 1, w - 1, 0] == two_open  # corners


def test_single_sample_step_horizon():
    # q = 2: one sample per direction at 2 pixels (16 / (2 * 4)); directions 0 and 180 degrees.
    w, h = 16, 8
    z = np.ones((h, w), np.float32)
    z[4, 10] = 0.5                      # two pixels right of the centre pixel (8, 4): closer by 0.5
    zb, op, col, img = _planes(w, h, z, colour=(255, 255, 255))
    img[:] = 1000
    out = oracle.ssao(_scene(2, persp=1.0), zb, op, col, img)
    # three-point plane coordinates: x2 = ((x / w - 0.5) * w / h) * z * persp, y2 = (y / h - 0.5) * z * persp;
    # centre (0, 0) at z 1, sample (0.125, 0) at z 0.5: the horizon rises by 0.5 over 0.125
    ao = (0.5 + (0.5 - math.atan(0.5 / 0.125) / math.pi)) / 2
    assert out[4, 8, 0] == int(1000 + 65535.0 * 255 / 256.0 * ao)
    # pixel (12, 4) sees the step in its 180-degree sample (10, 4): the sample's plane x shrinks with its
    # own depth, x2 = 0.5 at z 1 against xx2 = 0.125 at z 0.5, so the rise is 0.5 over 0.375
    ao_r = (0.5 + (0.5 - math.atan(0.5 / 0.375) / math.pi)) / 2
    assert out[4, 12, 0] == int(1000 + 65535.0 * 255 / 256.0 * ao_r)


def test_background_opacity_and_clamp():
    w, h = 32, 24
    z = np.full((h, w), 1.0, np.float32)
    z[5, 5] = 1e20                       # background: no AO term
    zb, op, col, img = _planes(w, h, z, colour=(255, 255, 255))
    op[10, 10] = 65535                   # fully fogged: no AO term
    op[12, 12] = 32768
    img[:] = 7
    img[20, 20] = 65000                  # clamps at 65535
    out = oracle.ssao(_scene(4, intensity=2.0), zb, op, col, img)
    assert np.all(out[5, 5] == 7) and np.all(out[10, 10] == 7)
    assert np.all(out[20, 20] == 65535)
    assert out[12, 12, 0] == int(7 + 65535.0 * 255 / 256.0 * 0.5 * 2.0 * (1.0 - 32768 / 65535.0))

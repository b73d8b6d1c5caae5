"""The oracle's depth of field (PostRendering_DOF, src/image.cpp:57-247)
pinned by hand-worked cases: an in-focus frame is left alone (blur 0 gives
an undefined opacity that never passes the op > 0 test), one defocused
pixel among in-focus ones splats a disc whose weights follow
op = (blur - r) / (blur^2.5 pi / 3), and the processing order of a
three-pixel frame traced by hand through the z sort (from index 1) and the
randomisation with glibc rand() after srand(1)."""
import math

import numpy as np

from mandelbulber_amd import scene as S
from oracle import oracle


def _scene(focus=21.7, radius=10.0):
    s = S.defaults()
    s.dof_focus = focus
    s.dof_radius = radius
    return s


def _neutral(s):
    return 10 ** (s.dof_focus / 10.0 - 16.0)


def test_in_focus_frame_is_unchanged():
    w, h = 24, 16
    s = _scene(focus=160.0)                 # neutral = 10^0 = 1
    z = np.full((h, w), 1.0, np.float32)
    rng = np.random.default_rng(1)
    img = rng.integers(0, 65535, (h, w, 3)).astype(np.uint16)
    a = rng.integers(0, 65535, (h, w)).astype(np.uint16)
    out, aout = oracle.dof(s, z, img, a, 1)
    assert np.array_equal(out, img) and np.array_equal(aout, a)


def test_single_defocused_pixel_splats_a_disc():
    w, h = 32, 32
    s = _scene(focus=160.0, radius=250.0)   # neutral 1; deep = 250 * 32 / 1000 = 8
    z = np.full((h, w), 1.0, np.float32)
    z[16, 16] = 2.0                         # blur = |2 - 1| / 2 * 8 = 4
    img = np.zeros((h, w, 3), np.uint16)
    img[16, 16] = (60000, 30000, 0)
    a = np.zeros((h, w), np.uint16)
    a[16, 16] = 50000
    out, aout = oracle.dof(s, z, img, a, 1)
    blur = 4.0
    factor = blur ** 2.5 * math.pi / 3.0
    for (dy, dx) in [(0, 1), (1, 1), (0, 3), (2, 3), (3, 3), (0, 4)]:
        r = float(np.float32(math.sqrt(dx * dx + dy * dy)))   # the float sqrt of the reference
        op = min(1.0, max(0.0, (blur - r) / factor))
        want = (int(60000 * op), int(30000 * op), 0, int(50000 * op))
        got = tuple(int(v) for v in out[16 + dy, 16 + dx]) + (int(aout[16 + dy, 16 + dx]),)
        assert got == want, (dx, dy, got, want)
    # outside the disc (r >= blur) and beyond size = int(blur) nothing moves
    assert np.all(out[16, 21] == 0) and np.all(out[12, 12] == 0)
    # the centre blends with its own original value: unchanged
    assert tuple(out[16, 16]) == (60000, 30000, 0)


def test_processing_order_by_hand():
    # z = [5, 3, 4]: the sort skips index 0 and leaves [3, 4]; with srand(1) the first
    # glibc rand() values are 1804289383 (% 3 = 1) and 846930886 (% 2 = 0):
    # i = 2 swaps with 1, i = 1 and i = 0 stay; the frame is processed from the end
    s = _scene()
    z = np.array([[5.0, 3.0, 4.0]], np.float32)
    assert list(oracle.dof_order(s, z, 1)) == [1, 2, 0]


def test_processing_order_is_far_to_near_when_blur_varies():
    w, h = 40, 30
    rng = np.random.default_rng(3)
    z = rng.uniform(1.0, 5.0, (h, w)).astype(np.float32)
    # focus at z = 3: blur sizes (z - 3) / z * deep vary in sign and size across the frame, and the
    # randomisation only swaps with pixels of similar size: the sequence stays far to near
    s = _scene(focus=10 * (16 + math.log10(3.0)))
    order = oracle.dof_order(s, z, 1234)
    assert sorted(order) == list(range(w * h))
    assert np.corrcoef(np.arange(w * h), z.reshape(-1)[order])[0, 1] < -0.9
    # focus near 0: every pixel has the same size, every first draw is accepted and the
    # randomisation is a uniform shuffle of the sorted frame
    order = oracle.dof_order(_scene(), z, 1234)
    assert sorted(order) == list(range(w * h))
    assert abs(np.corrcoef(np.arange(w * h), z.reshape(-1)[order])[0, 1]) < 0.2

"""Raytraced reflections in the oracle (MainThread's reflection loop,
src/Render3D.cpp:440-589): with every reflection coefficient 0 the loop
ends at the primary ray and the frame is the plain one; a reflecting
surface blends the mirrored ray's result into its own shade by its
coefficient, so the frame is affine in that coefficient (glow and fogs are
affine in the pixel too); a mirror adds the reflected bulb to the floor."""
import numpy as np

from mandelbulber_amd import scene
from oracle import oracle


def _floor_scene(w=48, h=36):
    s = scene.mandelbulb_plain(w, h)
    s.N = 20
    s.shadow = 1
    s.view_distance_max = 12.0
    # a floor plane under the bulb, seen from above at an angle
    s.plane_enabled = 1
    s.plane_centre[:] = (0.0, 0.0, -1.3)
    s.plane_normal[:] = (0.0, 0.0, 1.0)
    s.vp[:] = (0.0, -3.2, 1.4)
    s.beta = -0.35
    s.raytraced_reflections = 1
    s.reflections_max = 2
    return s


def test_no_reflecting_surface_gives_the_plain_frame():
    s = _floor_scene()
    s.reflect = 0.0
    s.plane_reflect = 0.0
    r, _ = oracle.render(s)
    s.raytraced_reflections = 0
    p, _ = oracle.render(s)
    assert np.array_equal(r.rgb, p.rgb) and np.array_equal(r.z, p.z)
    assert np.array_equal(r.alpha, p.alpha) and np.array_equal(r.colour, p.colour)


def test_reflection_is_affine_in_the_coefficient():
    s = _floor_scene()
    frames = {}
    for a in (0.0, 0.5, 1.0):
        s.plane_reflect = a
        frames[a], _ = oracle.render(s)
    f0, fh, f1 = (frames[a].rgb.astype(np.float64) for a in (0.0, 0.5, 1.0))
    floor = (frames[0.0].z < 1e19) & np.all(np.abs(f1 - f0) > 1e-4, axis=0)
    assert floor.sum() > 50  # the floor reflects something
    # (FP32 planes: their rounding is the tolerance)
    np.testing.assert_allclose(fh, 0.5 * (f0 + f1), rtol=0, atol=3e-6)
    # depth and colour index come from the primary ray
    assert np.array_equal(frames[0.5].z, frames[0.0].z)
    assert np.array_equal(frames[0.5].colour, frames[0.0].colour)


def test_mirror_floor_shows_the_bulb_below_it():
    """With a perfect mirror floor the floor pixels right below the bulb's
    image take the bulb's shade: their mirrored rays hit the bulb, so they
    differ from the plain floor by more than where the mirrored ray leaves
    to the background."""
    s = _floor_scene(64, 48)
    s.plane_reflect = 1.0
    m, _ = oracle.render(s)
    s.plane_reflect = 0.0
    p, _ = oracle.render(s)
    d = np.abs(m.rgb.astype(np.float64) - p.rgb).max(axis=0)
    floor_rows = np.where(((p.z < 1e19) & (d > 1e-3)).any(axis=1))[0]
    assert floor_rows.size > 5

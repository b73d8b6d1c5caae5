"""The soft-shadow branch of MainShadow (src/shaders.cpp:503-535, cone angle
> 0) in the oracle: a point whose shadow ray meets nothing is fully lit
whatever the cone; a ray through the bulb is fully shadowed; a ray passing
just beside the bulb is in the penumbra, darker the wider the cone."""
import numpy as np
import pytest

from mandelbulber_amd import scene
from oracle import oracle


def _shadow(s, p, thr=1e-3):
    o = oracle.object_shader(s, tuple(p), (0.0, 1.0, 0.0), thr, 3.0)
    return o.shadow


@pytest.fixture(scope="module")
def bulb():
    s = scene.mandelbulb_plain(16, 16)
    s.N = 20
    s.shadow = 1
    s.penetrating_lights = 0
    s.main_light_alpha = 0.4
    s.main_light_beta = 0.3
    s.view_distance_max = 20.0
    _, _, light = oracle.camera(s)
    return s, light / np.linalg.norm(light)


def test_open_sky_is_lit_for_every_cone(bulb):
    s, L = bulb
    p = 3.0 * L  # beyond the bulb, looking away from it
    for cone in (0.0, 1.0, 5.0, 20.0):
        s.shadow_cone_angle = cone
        assert _shadow(s, p) == 1.0


def test_ray_through_the_bulb_is_dark(bulb):
    s, L = bulb
    p = -2.5 * L  # the bulb lies between the point and the light
    for cone in (0.0, 2.0, 10.0):
        s.shadow_cone_angle = cone
        assert _shadow(s, p) == 0.0


def test_penumbra_darkens_with_the_cone(bulb):
    s, L = bulb
    side = np.cross(L, [0.0, 0.0, 1.0])
    side /= np.linalg.norm(side)
    p = -2.5 * L + 1.45 * side  # the shadow ray passes beside the bulb
    s.shadow_cone_angle = 0.0
    assert _shadow(s, p) == 1.0  # hard shadow: the ray misses
    vals = []
    for cone in (1.0, 5.0, 15.0, 40.0):
        s.shadow_cone_angle = cone
        vals.append(_shadow(s, p))
    assert all(0.0 <= v <= 1.0 for v in vals)
    assert all(a >= b for a, b in zip(vals, vals[1:])), vals
    assert vals[-1] < vals[0]


def test_penetrating_main_light_keeps_the_path_fraction(bulb):
    """With penetrating lights MainShadow marches to factor = delta /
    resolution and a hit after a length i keeps 1 - (factor - i)/factor
    (src/shaders.cpp:487-488, :548-553).  Beyond the bulb along the light, a
    sphere across the path at i = 2.2: threshold 0.5, quality 1 and a
    16-pixel width give factor 0.5 x 16 = 8, so the shadow is 2.2/8."""
    s, L = bulb
    s = scene.copy(s)
    s.shadow_cone_angle = 0.0
    s.penetrating_lights = 1
    p = 3.0 * L
    s.sphere_enabled = 1
    s.sphere_centre[:] = tuple(p + 2.5 * L)
    s.sphere_radius = 0.3
    assert s.image_width == 16 and s.quality == 1.0
    assert _shadow(s, p, thr=0.5) == pytest.approx(2.2 / 8.0, rel=1e-9)

"""Scene helpers: camera placement and the batched-flight generator."""
import numpy as np

from mandelbulber_amd import scene
from oracle import oracle


def test_look_at_origin_points_the_camera_at_the_origin():
    """The oracle's camera matrix (pinned to the reference CRotationMatrix)
    maps the forward axis (0,1,0) onto the direction from vp to the origin."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        vp = rng.normal(size=3) * rng.uniform(1.0, 5.0)
        s = scene.look_at_origin(scene.mandelbulb_plain(64, 48), vp, roll=rng.uniform(-1, 1))
        R, start, _ = oracle.camera(s)
        fwd = R @ np.array([0.0, 1.0, 0.0])
        assert np.allclose(fwd, -vp / np.linalg.norm(vp), atol=1e-9)
        # the central pixel's ray (x = W/2, y = H/2) is the forward axis
        d, _ = oracle.view_vector(s, 32, 24)
        assert np.allclose(d, fwd, atol=1e-12)


def test_flight_cycles_formulas_and_draws_its_parameters():
    fl = scene.flight(64, seed=3)
    assert [s.formula for s in fl[:6]] == [2, 8, 7, 2, 8, 7]
    for k, s in enumerate(fl):
        assert (s.image_width, s.image_height) in scene.FLIGHT_RESOLUTIONS
        dist = np.linalg.norm(np.array(s.vp[:])) / (2.0 if s.formula == 8 else 1.0)
        assert 1.5 <= dist <= 5.0
        # no camera starts inside the solid
        R, start, _ = oracle.camera(s)
        d, max_iter, _, _ = oracle.calculate_distance(s, start)
        assert d > 0 and not max_iter
        if s.formula == 2:
            assert 6.0 <= s.power <= 10.0
        if s.formula == 7:
            assert s.N == 15
        if s.formula == 8:
            assert s.mb_scale == 2.0 and s.mb_min_radius == 0.5 and s.mb_fixed_radius == 1.0
    # all four resolutions occur, the draws are reproducible
    assert {(s.image_width, s.image_height) for s in fl} == set(scene.FLIGHT_RESOLUTIONS)
    assert [s.power for s in scene.flight(64, seed=3)] == [s.power for s in fl]

"""The hybrid building blocks 22-36 (src/fractal.cpp:984-1103) in the
oracle, one block per hybrid slot, through the orbit radius after one
iteration (the delta-DE mode of Compute returns it): each block's invariant
(angle multiplication and abs-free sign folds keep the radius where they
must, scale and offset move one axis, invertR keeps the direction and maps
r to r^2/(r+p), the spherical fold leaves r >= 2p alone, powXYZ at power 1
is the identity)."""
import math

import numpy as np
import pytest

from mandelbulber_amd import scene
from oracle import oracle

P0 = (0.3, -0.4, 0.5)


def _radius(block, power, p=P0, iters=1):
    s = scene.mandelbulb_plain(8, 8)
    s.formula = 12
    s.N = iters
    s.bailout = 1e9
    s.hybrid_formula[:] = (block, 0, 0, 0, 0)
    s.hybrid_iterations[:] = (1, 0, 0, 0, 0)
    s.hybrid_power[:] = (power, 0.0, 0.0, 0.0, 0.0)
    s.hybrid_cyclic = 1
    r, _ = oracle.compute(s, p, mode=3, n_iter=iters)
    return r


def _norm(v):
    return math.sqrt(sum(c * c for c in v))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_scale_and_offset_move_one_axis(axis):
    z = list(P0)
    z[axis] *= 2.5
    assert _radius(28 + axis, 2.5) == pytest.approx(_norm(z), rel=1e-15)
    z = list(P0)
    z[axis] += 0.75
    assert _radius(31 + axis, 0.75) == pytest.approx(_norm(z), rel=1e-15)
    # three iterations compose
    z = list(P0)
    z[axis] *= 2.0 ** 3
    assert _radius(28 + axis, 2.0, iters=3) == pytest.approx(_norm(z), rel=1e-15)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_invert_axis_shrinks_that_axis_keeping_its_sign(axis):
    for p in (0.5, 2.0):
        r = _radius(22 + axis, p)
        c = P0[axis]
        others = sum(P0[k] ** 2 for k in range(3) if k != axis)
        c2 = math.sqrt(r * r - others)  # |new component|
        assert c2 == pytest.approx(c * c / (abs(c) + p), rel=1e-12)
        assert c2 < abs(c)


def test_invert_r_keeps_direction_and_maps_the_radius():
    r0 = _norm(P0)
    for p in (0.1, 1.0, 3.0):
        assert _radius(25, p) == pytest.approx(r0 * r0 / (r0 + p), rel=1e-15)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_angle_multiply_keeps_the_radius(axis):
    r0 = _norm(P0)
    for p in (1.0, 2.0, 3.7):
        assert _radius(34 + axis, p) == pytest.approx(r0, rel=1e-14)


def test_spherical_fold_leaves_far_points_alone():
    r0 = _norm(P0)
    assert _radius(26, 0.3) == pytest.approx(r0, rel=1e-15)  # r >= 2p: unchanged
    # inside the inner radius every component becomes 1/p^2
    assert _radius(26, 0.9) == pytest.approx(math.sqrt(3.0) / 0.81, rel=1e-15)


def test_pow_xyz_at_power_one_is_the_identity_and_odd():
    assert _radius(27, 1.0) == pytest.approx(_norm(P0), rel=1e-15)
    expect = _norm([abs(c) ** 3 for c in P0])
    assert _radius(27, 3.0) == pytest.approx(expect, rel=1e-14)
    assert _radius(27, 3.0, p=tuple(-c for c in P0)) == pytest.approx(expect, rel=1e-14)

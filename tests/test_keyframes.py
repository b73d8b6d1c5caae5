"""Keyframe interpolation (CMorph::CatmullRom, src/morph.cpp:49-109) through
mbr_keyframe_scene, pinned by the spline's own algebra: it passes through
the keys, reproduces linear data exactly, runs in log space (so geometric
data stays geometric) when the four support values share a sign and vary by
more than 10% of their mean, and clamps its support at the first key."""
import math

import pytest

from mandelbulber_amd import scene as S


def _keys(values, field="zoom", index=None):
    keys = []
    for v in values:
        k = S.defaults()
        if index is None:
            setattr(k, field, v)
        else:
            getattr(k, field)[index] = v
        keys.append(k)
    return keys


def test_passes_through_keys_and_is_linear_on_linear_data():
    keys = _keys([-1.0, 0.0, 1.0, 2.0, 3.0], field="vp", index=0)   # mixed signs: no log space
    base = S.defaults()
    # (key 3's support 1, 2, 3, 3 is all positive and varies: it would run in log space)
    for frame, want in [(0, -1.0), (4, 0.0), (8, 1.0), (12, 2.0), (6, 0.5), (9, 1.25), (11, 1.75)]:
        assert S.keyframe(keys, 4, frame, base).vp[0] == pytest.approx(want, abs=1e-15), frame


def test_geometric_data_stays_geometric():
    keys = _keys([1.0, 2.0, 4.0, 8.0, 16.0], field="power")
    base = S.defaults()
    s = S.keyframe(keys, 2, 3, base)        # key 1, factor 0.5: 2^1.5
    assert s.power == pytest.approx(2.0 ** 1.5, rel=1e-12)


def test_support_is_clamped_at_the_first_key():
    keys = _keys([0.0, 1.0, 2.0], field="vp", index=2)
    s = S.keyframe(keys, 2, 1, S.defaults())  # key 0, factor 0.5, v1 = v2 = 0, v3 = 1, v4 = 2
    want = 0.5 * ((-0.0 + 1.0) * 0.5 + (0.0 - 0.0 + 4.0 - 2.0) * 0.25 + (0.0 + 0.0 - 3.0 + 2.0) * 0.125)
    assert s.vp[2] == pytest.approx(want, abs=1e-15)


def test_other_fields_come_from_the_base_and_n_truncates():
    keys = _keys([10.0, 20.0], field="power")
    for k, n in zip(keys, (100, 201)):
        k.N = n
        k.image_width = 7
    base = S.defaults()
    base.image_width = 333
    s = S.keyframe(keys, 4, 2, base)
    assert s.image_width == 333
    # N: 100 -> 201 in log space (same sign, deviation > 10%), read back as int
    l1, l2 = math.log(100.0), math.log(201.0)
    want = math.exp(0.5 * (2 * l1 + (-l1 + l2) * 0.5 + (2 * l1 - 5 * l1 + 4 * l2 - l2) * 0.25
                           + (-l1 + 3 * l1 - 3 * l2 + l2) * 0.125))
    assert s.N == int(want)

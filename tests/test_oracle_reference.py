"""Pins the oracle against the reference's own compiled arithmetic.

oracle/_ref/libmbref.so is built by oracle/Makefile from the reference's
src/algebra.cpp, src/common_math.cpp and src/primitives.cpp, compiled where
they lie, plus a C shim (oracle/ref_shim.cpp).  These are the reference
files on the render path that build without GTK; they fix the camera
rotation convention, the projection, the Mandelbox rotation matrices, the
primitive SDFs and Random().
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from mandelbulber_amd import scene as S
from oracle import oracle as O

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libmbref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference tree absent)")

D3 = C.c_double * 3


@pytest.fixture(scope="module")
def ref():
    lib = C.CDLL(REF)
    lib.ref_camera_matrix.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double * 9]
    lib.ref_camera_rotate.argtypes = [C.c_double, C.c_double, C.c_double, D3, D3]
    lib.ref_set_rotation3.argtypes = [D3, C.c_double * 9]
    lib.ref_inv_projection.argtypes = [D3, D3, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_double, D3]
    for n in ("ref_primitive_plane",):
        getattr(lib, n).argtypes = [D3, D3, D3]
        getattr(lib, n).restype = C.c_double
    lib.ref_primitive_box.argtypes = [D3, D3, D3, C.c_int]
    lib.ref_primitive_box.restype = C.c_double
    lib.ref_primitive_sphere.argtypes = [D3, D3, C.c_double, C.c_int]
    lib.ref_primitive_sphere.restype = C.c_double
    lib.ref_primitive_water.argtypes = [D3, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                        C.c_int]
    lib.ref_primitive_water.restype = C.c_double
    lib.ref_random_sequence.argtypes = [C.c_uint, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.ref_dmax.argtypes = [C.c_double] * 3
    lib.ref_dmax.restype = C.c_double
    return lib


def _angles(rng, n):
    a = rng.uniform(-math.pi, math.pi, size=(n, 3))
    a[::5, 0] = 0.0  # exercise the skipped rotations (angle == 0)
    a[::7, 1] = 0.0
    a[::3, 2] = 0.0
    return a


def test_camera_matrix_matches_reference(ref):
    rng = np.random.default_rng(1)
    for al, be, ga in _angles(rng, 60):
        sc = S.defaults()
        sc.alpha, sc.beta, sc.gamma = al, be, ga
        R, _, _ = O.camera(sc)
        m = (C.c_double * 9)()
        ref.ref_camera_matrix(al, be, ga, m)
        assert np.array_equal(R.ravel(), np.array(m[:])), (al, be, ga)


def test_light_vector_matches_reference(ref):
    rng = np.random.default_rng(2)
    for al, be, ga in _angles(rng, 30):
        sc = S.defaults()
        sc.alpha, sc.beta, sc.gamma = al, be, ga
        la, lb = rng.uniform(-3, 3, size=2)
        sc.main_light_alpha, sc.main_light_beta = la, lb
        _, _, light = O.camera(sc)
        v = D3(math.cos(la - 0.5 * math.pi) * math.cos(-lb), math.sin(la - 0.5 * math.pi) * math.cos(-lb),
               math.sin(-lb))
        out = D3()
        ref.ref_camera_rotate(al, be, ga, v, out)
        assert np.array_equal(light, np.array(out[:]))


def test_mandelbox_rotation_matches_reference(ref):
    rng = np.random.default_rng(3)
    for a in _angles(rng, 60):
        m = (C.c_double * 9)()
        ref.ref_set_rotation3(D3(*a), m)
        assert np.array_equal(O.rotation3(a).ravel(), np.array(m[:]))


@pytest.mark.parametrize("w,h", [(320, 240), (64, 48), (100, 37)])
def test_primary_rays_project_back_to_their_pixel(ref, w, h):
    """InvProjection3D (the reference's own inverse of the camera) maps a
    point on the ray of pixel (x, y) back to screen position (x, y).  It
    offsets the ray start along the unrotated y axis (src/common_math.cpp:
    117-125), so it inverts MainThread's camera exactly when the zoom offset
    vanishes or the rotation is the identity: both cases are covered."""
    rng = np.random.default_rng(4)
    sc = S.defaults()
    sc.image_width, sc.image_height = w, h
    for i in range(40):
        sc.persp = float(rng.uniform(0.3, 2.0))
        sc.vp[:] = rng.uniform(-2, 2, size=3)
        if i % 2:
            sc.alpha, sc.beta, sc.gamma = rng.uniform(-3, 3, size=3)
            sc.zoom = 1e-15
        else:
            sc.alpha, sc.beta, sc.gamma = 0.0, 0.0, 0.0
            sc.zoom = float(rng.uniform(0.1, 3.0))
        _, start, _ = O.camera(sc)
        x, y = int(rng.integers(0, w)), int(rng.integers(0, h))
        d, _ = O.view_vector(sc, x, y)
        t = float(rng.uniform(0.5, 5.0))
        p = start + d * t
        out = D3()
        ref.ref_inv_projection(D3(*p), D3(*sc.vp), sc.alpha, sc.beta, sc.gamma, 0, sc.persp, sc.zoom, w, h, out)
        assert abs(out[0] - x) < 1e-7 and abs(out[1] - y) < 1e-7, (x, y, out[:])
        # depth of InvProjection3D is the planar (camera-forward) distance
        # along the unnormalised view vector: the scan value t
        assert abs(out[2] - t) < 1e-9


def _only(sc, **kw):
    sc.primitive_only_plane = 1
    for k, v in kw.items():
        setattr(sc, k, v)
    return sc


def test_primitives_match_reference(ref):
    rng = np.random.default_rng(5)
    for _ in range(200):
        p = rng.uniform(-4, 4, size=3)
        c = rng.uniform(-1, 1, size=3)
        n = rng.uniform(-1, 1, size=3)
        size = rng.uniform(0.2, 3, size=3)
        rad = float(rng.uniform(0.2, 2))
        sc = _only(S.defaults(), plane_enabled=1)
        sc.plane_centre[:] = c
        sc.plane_normal[:] = n
        d, _, _, _ = O.calculate_distance(sc, p, detail=1e-3)
        assert d == min(10.0, ref.ref_primitive_plane(D3(*p), D3(*c), D3(*n)))
        for inv in (0, 1):
            sc = _only(S.defaults(), **({"inv_box_enabled": 1} if inv else {"box_enabled": 1}))
            (sc.inv_box_centre if inv else sc.box_centre)[:] = c
            (sc.inv_box_size if inv else sc.box_size)[:] = size
            d, _, _, _ = O.calculate_distance(sc, p, detail=1e-3)
            assert d == min(10.0, ref.ref_primitive_box(D3(*p), D3(*c), D3(*size), inv))
            sc = _only(S.defaults(), **({"inv_sphere_enabled": 1} if inv else {"sphere_enabled": 1}))
            (sc.inv_sphere_centre if inv else sc.sphere_centre)[:] = c
            if inv:
                sc.inv_sphere_radius = rad
            else:
                sc.sphere_radius = rad
            d, _, _, _ = O.calculate_distance(sc, p, detail=1e-3)
            assert d == min(10.0, ref.ref_primitive_sphere(D3(*p), D3(*c), rad, inv))


def test_water_matches_reference(ref):
    rng = np.random.default_rng(6)
    for _ in range(200):
        p = rng.uniform(-2, 2, size=3)
        p[2] = rng.uniform(-0.3, 0.3)
        sc = _only(S.defaults(), water_enabled=1)
        sc.water_height = float(rng.uniform(-0.1, 0.1))
        sc.water_amplitude = float(rng.uniform(0.01, 0.1))
        sc.water_length = float(rng.uniform(0.1, 0.5))
        sc.water_rotation = float(rng.uniform(-90, 90))
        sc.water_iterations = int(rng.integers(1, 8))
        sc.water_anim_speed = float(rng.uniform(0, 0.2))
        sc.frame_no = int(rng.integers(0, 100))
        d, _, _, _ = O.calculate_distance(sc, p, detail=1e-3)
        r = ref.ref_primitive_water(D3(*p), sc.water_height, sc.water_amplitude, sc.water_length,
                                    sc.water_rotation, sc.water_iterations, sc.water_anim_speed, sc.frame_no)
        assert d == min(10.0, r)


def test_glibc_jitter_is_reference_random(ref):
    """The oracle's reference-exact jitter mode draws rand() % 1001, which
    is Random(1000) of src/common_math.cpp:25-32."""
    libc = C.CDLL("libc.so.6")
    out = (C.c_int * 500)()
    ref.ref_random_sequence(1234, 1000, 500, out)
    libc.srand(1234)
    mine = [libc.rand() % 1001 for _ in range(500)]
    assert list(out) == mine


def test_new_palette_from_reference_random(ref, tmp_path):
    """A scene without a palette gets NewPalette (src/image.cpp:28-55) drawn
    from srand(coloring_random_seed) (src/settings.cpp:1185-1195)."""
    f = tmp_path / "nopal.fract"
    f.write_text("formula 2;\ncoloring_random_seed 987;\ncoloring_saturation 0.5;\n")
    sc = S.load(str(f))
    draws = (C.c_int * 1020)()
    ref.ref_random_sequence(987, 255, 1020, draws)
    sat = 0.5
    for i in range(255):
        y, r, g, b = draws[4 * i:4 * i + 4]
        Y = int((y - 128) / (1.0 + sat))
        rgb = [int(Y + 128 + (v - 128) * sat) for v in (r, g, b)]
        rgb = [min(255, max(0, v)) for v in rgb]
        if i < 253:  # 253 and 254 are overwritten after the loop
            assert (sc.palette[i].R, sc.palette[i].G, sc.palette[i].B) == tuple(rgb), i
    assert (sc.palette[255].R, sc.palette[255].G, sc.palette[255].B) == (255, 255, 255)
    assert (sc.palette[254].R, sc.palette[254].G, sc.palette[254].B) == (0, 0, 0)
    assert (sc.palette[253].R, sc.palette[253].G, sc.palette[253].B) == (0, 20, 40)


def test_limits_box_uses_reference_dmax(ref):
    rng = np.random.default_rng(7)
    sc = S.defaults()
    sc.limits_enabled = 1
    sc.limit_min[:] = (-0.5, -0.7, -0.2)
    sc.limit_max[:] = (0.6, 0.3, 0.9)
    for _ in range(100):
        p = rng.uniform(-3, 3, size=3)
        da = max(p[0] - 0.6, -(p[0] + 0.5))
        db = max(p[1] - 0.3, -(p[1] + 0.7))
        dc = max(p[2] - 0.9, -(p[2] + 0.2))
        lb = ref.ref_dmax(da, db, dc)
        d, _, _, _ = O.calculate_distance(sc, p, detail=1e-3)
        if lb > 1e-3:
            assert d == lb
        else:
            assert d >= lb

"""The oracle's generic formulas (src/fractal.cpp:189-320): trig_DE,
mandelbulb2/3/4 and xenodreambuie, and the delta DE of CalculateDistance
(:1307-1342).  Pinned by what the mathematics fixes:
- at power 1 the maps reduce to z -> (x, y, -z) + c (trig_DE),
  (x, y, |z|) + c (xenodreambuie) and z + c (mandelbulb4);
- mandelbulb2/3 are three (two) planar complex squarings, (u + iv)^2 / |u + iv|,
  which numpy's complex arithmetic states independently;
- on the positive x axis trig_DE's derivative vector stays on the axis, so its
  distance estimate equals trig_optim's (already pinned) there;
- the delta DE's linear and logarithmic forms differ by the factor ln r, and a
  point that never escapes has distance 0;
- foldingIntPow2 (:794-836): its map in complex form, (x + iy)^2 / rho^2
  Re(B), Im(B), B = (rho - iz)^2, with the z factor applied after c; the
  squared norm; the box fold as a mirror at +-F (same orbit in Julia mode);
  the sphere fold as z *= 4 and as the inversion p / r^2;
- aexion (:916-947): the tetrahedral start and the 4-D map restated in the
  sum/difference coordinates in which it decouples, one and two iterations;
- benesi (:948-970): the planar squaring -2|x| (y + iz)^2 / rho, both
  branches of the x rule, and the squared norm.
Each of these fails when a sign, a term, the fold offset, the sphere-fold
radius, the z-factor order or the branch rule is changed in the oracle."""
import cmath
import math

import numpy as np
import pytest

from mandelbulber_amd import scene as S
from oracle import oracle

MODE_DELTA1 = 3


def _scene(formula, power, julia=None, n=250):
    s = S.mandelbulb_plain(64, 48)
    s.formula = formula
    s.power = power
    s.N = n
    if julia is not None:
        s.julia_mode = 1
        for i in range(3):
            s.julia[i] = julia[i]
    return s


def _r1(s, p):
    """|z| after one iteration (the delta-DE run returns the orbit radius)."""
    r, _ = oracle.compute(s, p, mode=MODE_DELTA1, n_iter=1)
    return r


POINTS = [(0.3, -0.7, 0.45), (-1.1, 0.2, -0.6), (0.05, 0.9, -0.2), (-0.4, -0.3, 0.8)]
J = (0.13, -0.27, 0.31)


@pytest.mark.parametrize("p", POINTS)
def test_power_one_reductions(p):
    x, y, z = p
    want = {1: (x + J[0], y + J[1], -z + J[2]), 11: (x + J[0], y + J[1], abs(z) + J[2]),
            15: (x + J[0], y + J[1], z + J[2])}
    for formula, z1 in want.items():
        got = _r1(_scene(formula, 1.0, julia=J), p)
        assert got == pytest.approx(math.sqrt(sum(v * v for v in z1)), rel=1e-12), formula


def _sq(u, v):
    w = complex(u, v)
    w = w * w / abs(w)
    return w.real, w.imag


@pytest.mark.parametrize("p", POINTS)
def test_mandelbulb2_is_three_planar_squarings(p):
    x, y, z = p
    r0 = math.sqrt(x * x + y * y + z * z)
    x, y = _sq(x, y)
    y, z = _sq(y, z)
    x, z = _sq(x, z)
    z1 = np.array([x, y, z]) * r0 + np.array(J)
    assert _r1(_scene(13, 2.0, julia=J), p) == pytest.approx(np.linalg.norm(z1), rel=1e-12)


@pytest.mark.parametrize("p", POINTS)
def test_mandelbulb3_is_two_signed_planar_squarings(p):
    x, y, z = p
    r0 = math.sqrt(x * x + y * y + z * z)
    sign2 = -1.0 if x < 0 else 1.0
    x, y = _sq(x, y)
    sign = -1.0 if x < 0 else 1.0
    xs, zs = _sq(x, z)
    z1 = np.array([xs * sign, y, zs * sign2]) * r0 + np.array(J)
    assert _r1(_scene(14, 2.0, julia=J), p) == pytest.approx(np.linalg.norm(z1), rel=1e-12)


@pytest.mark.parametrize("x", [0.6, 0.9, 1.05, 1.2, 1.6, 3.0])
def test_trig_de_equals_trig_optim_on_positive_x_axis(x):
    a = _scene(1, 8.0)
    b = _scene(2, 8.0)
    d1, _, it1, _ = oracle.calculate_distance(a, (x, 0.0, 0.0))
    d2, _, it2, _ = oracle.calculate_distance(b, (x, 0.0, 0.0))
    assert it1 == it2
    assert d1 == pytest.approx(d2, rel=1e-12, abs=1e-300)


def test_delta_de_linear_and_log_forms():
    p = (1.3, 0.4, -0.5)
    s = _scene(13, 2.0)
    d_log, mi, it, _ = oracle.calculate_distance(s, p)
    s.linear_DE_mode = 1
    d_lin, _, _, _ = oracle.calculate_distance(s, p)
    assert not mi and d_log > 0
    r, _ = oracle.compute(s, p, mode=MODE_DELTA1)
    assert d_log == pytest.approx(d_lin * math.log(r), rel=1e-12)
    # a point of the set never escapes: L == N and distance 0
    s.linear_DE_mode = 0
    d0, mi0, it0, _ = oracle.calculate_distance(s, (0.01, 0.02, -0.01))
    assert mi0 and it0 == s.N and d0 == 0.0


@pytest.mark.parametrize("formula,power", [(11, 5.0), (13, 2.0), (14, 2.0), (15, 8.0)])
def test_delta_de_far_field_is_a_distance_scale(formula, power):
    # far from the set the estimate is of the order of the distance (within 2x of |p|)
    s = _scene(formula, power)
    for p in [(3.0, 1.0, 0.5), (-2.5, -2.0, 1.0), (0.5, 4.0, -2.0)]:
        d, mi, _, _ = oracle.calculate_distance(s, p)
        rp = math.sqrt(sum(v * v for v in p))
        assert not mi and 0.1 * rp < d < 2.0 * rp, (formula, p, d)


def _mandelbrot_escape(c, n, bailout=100.0):
    # the orbit starts at the point itself (z0 = c), as Compute's does
    z = c
    for k in range(n):
        z = z * z + c
        if abs(z) > bailout:
            return k
    return n


@pytest.mark.parametrize("formula", [3, 4, 5, 6, 21])
def test_z0_plane_is_the_mandelbrot_set(formula):
    """fast_trig, hypercomplex, quaternion, minus_fast_trig and bristorbrot
    keep z = 0 (and w = 0) in the z = 0 plane, where each is z -> z^2 + c
    on x + iy: the escape iteration equals the textbook Mandelbrot one."""
    s = _scene(formula, 2.0, n=60)
    rng = np.random.default_rng(formula)
    checked = 0
    for _ in range(400):
        x, y = rng.uniform(-2.2, 1.2), rng.uniform(-1.4, 1.4)
        want = _mandelbrot_escape(complex(x, y), s.N, s.bailout)
        # points whose |z| lands within 1e-6 of the bailout could round either way
        _, it = oracle.compute(s, (x, y, 0.0), mode=0)
        if want < s.N:
            z = complex(x, y)
            for k in range(want + 1):
                z = z * z + complex(x, y)
            if abs(abs(z) - s.bailout) < 1e-6 * s.bailout:
                continue
        assert it == want, (formula, x, y, it, want)
        checked += 1
    assert checked > 350


def test_smooth_mandelbox_tends_to_the_mandelbox():
    """With a very sharp smooth step the smooth folds are the box and sphere
    folds: the distance estimate equals tglad's away from the fold edges."""
    b = S.mandelbox_view(64, 48, shaded=False)
    a = S.copy(b)
    a.formula = 17
    a.mb_sharpness = 1e6
    rng = np.random.default_rng(5)
    n = 0
    for _ in range(200):
        p = tuple(rng.uniform(-4.0, 4.0, 3))
        da, _, ia, _ = oracle.calculate_distance(a, p)
        db, _, ib, _ = oracle.calculate_distance(b, p)
        assert ia == ib, p
        assert da == pytest.approx(db, rel=1e-6, abs=1e-300), p
        n += db > 1e-3
    assert n > 20


def _menger_ifs(s):
    """The Menger sponge as a kaleidoscopic IFS: |x|, |y|, |z|, the folds
    through the planes x = y, x = z, y = z, scale 3 about (1, 1, 1) in Menger
    mode."""
    s.formula = 10
    s.ifs_abs[0] = s.ifs_abs[1] = s.ifs_abs[2] = 1
    s.ifs_scale = 3.0
    for c in range(3):
        s.ifs_offset[c] = 1.0
    s.ifs_menger_sponge_mode = 1
    for i, d in enumerate([(1, -1, 0), (1, 0, -1), (0, 1, -1)]):
        for c in range(3):
            s.ifs_direction[i][c] = d[c]
        s.ifs_distance[i] = 0.0
        s.ifs_intensity[i] = 1.0
        s.ifs_enabled[i] = 1
    return s


def test_kaleidoscopic_ifs_reproduces_the_menger_sponge():
    """A fold through a plane is a reflection: for x < y the fold through
    x = y swaps x and y, which is menger_sponge's sort step; the rest of the
    IFS step (scale 3, offset 2 in x and y, z shifted above 1/3) is its
    affine step.  The distance estimates agree up to the rounding of the
    reflections (the orbit grows by 3 per iteration)."""
    m = S.menger_view(64, 48, shaded=False)
    ifs = _menger_ifs(S.copy(m))
    rng = np.random.default_rng(9)
    n = 0
    for _ in range(300):
        p = tuple(rng.uniform(-1.6, 1.6, 3))
        d7, _, i7, _ = oracle.calculate_distance(m, p)
        d10, _, i10, _ = oracle.calculate_distance(ifs, p)
        assert i7 == i10, p
        assert d10 == pytest.approx(d7, rel=1e-6, abs=1e-12), p
        n += d7 > 0
    assert n > 100


def _hybrid(slots, cyclic, n, julia=None):
    """slots: {index 0..4: (formula, iterations, power)}; the rest none."""
    s = _scene(12, 2.0, julia=julia, n=n)
    for k in range(5):
        f, it, pw = slots.get(k, (0, 1, 2.0))
        s.hybrid_formula[k], s.hybrid_iterations[k], s.hybrid_power[k] = f, it, pw
    s.hybrid_cyclic = 1 if cyclic else 0
    return s


def test_hybrid_with_one_slot_is_that_formula():
    """The default hybrid (slot 5 = trig_optim) iterates exactly as
    trig_optim with slot 5's power."""
    h = _hybrid({4: (2, 1, 8.0)}, cyclic=False, n=40)
    b = _scene(2, 8.0, n=40)
    for p in POINTS:
        assert oracle.compute(h, p, mode=MODE_DELTA1) == oracle.compute(b, p, mode=MODE_DELTA1)


@pytest.mark.parametrize("cyclic,want", [(True, lambda x: 4 * x + 1.5), (False, lambda x: 2 * x + 1.5)])
def test_hybrid_sequence_by_hand(cyclic, want):
    """Slots: scaleX (power 2) once, then offsetX (power 0.5).  On the x axis
    with a zero Julia constant, four iterations give x -> 2x -> 2x + 0.5 ->
    (cyclic) 4x + 1 -> 4x + 1.5, or (slot 5 for the rest) 2x + 1 -> 2x + 1.5."""
    s = _hybrid({0: (28, 1, 2.0), 4: (31, 1, 0.5)}, cyclic=cyclic, n=4, julia=(0.0, 0.0, 0.0))
    for x in (0.3, -0.7, 1.9):
        r, it = oracle.compute(s, (x, 0.0, 0.0), mode=MODE_DELTA1)
        assert it == 4 and r == pytest.approx(abs(want(x)), rel=1e-15)


# --- foldingIntPow2 (src/fractal.cpp:794-836) ---------------------------------
# Box fold at +-F, sphere fold (min radius^2 0.25, fixed 1), doubling, then the
# power-2 map (x, y, z) -> ((x^2 - y^2) t, 2xy t, -2 z rho), t = 1 - z^2/rho^2,
# rho = |(x, y)|, plus c, and finally z *= Zf.  In complex terms the map is
# (x + iy)^2 / rho^2 * Re(B) and Im(B) with B = (rho - iz)^2: numpy's complex
# arithmetic states it independently of the oracle's scalar form.

def _fip_scene(F=10.0, zf=1.0, julia=J):
    s = _scene(16, 2.0, julia=julia)
    s.folding_int_pow_factor = F
    s.folding_int_pow_z_factor = zf
    return s


def _fip_map(w, c, zf):
    x, y, z = w
    rho = math.hypot(x, y)
    a = complex(x, y) ** 2 / rho ** 2
    b = complex(rho, -z) ** 2
    xy = a * b.real
    return np.array([xy.real + c[0], xy.imag + c[1], (b.imag + c[2]) * zf])


@pytest.mark.parametrize("p", [(0.7, -0.5, 0.6), (-0.6, 0.3, -0.9), (0.3, 0.95, 0.2), (-0.8, -0.7, -0.2)])
@pytest.mark.parametrize("zf", [1.0, 5.0])
def test_folding_int_pow2_map_outside_the_folds(p, zf):
    """|p| >= 1 and every |p_i| <= F: no fold applies, z1 = map(2p) (with
    z *= Zf after adding c, so a Zf applied before the constant fails)."""
    assert math.sqrt(sum(v * v for v in p)) >= 1.0
    z1 = _fip_map(2.0 * np.array(p), J, zf)
    assert _r1(_fip_scene(zf=zf), p) == pytest.approx(np.linalg.norm(z1), rel=1e-12)


def test_folding_int_pow2_map_preserves_the_squared_norm():
    """With c = 0 the map squares the radius: |z1| = |2p|^2 (Zf = 1)."""
    for p in [(0.7, -0.5, 0.6), (-0.6, 0.3, -0.9), (1.1, 0.2, -0.3)]:
        assert sum(v * v for v in p) >= 1.0
        r = 2.0 * math.sqrt(sum(v * v for v in p))
        assert _r1(_fip_scene(julia=(0.0, 0.0, 0.0)), p) == pytest.approx(r * r, rel=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_folding_int_pow2_box_fold_mirrors_at_plus_minus_F(axis, sign):
    """Past +-F a component becomes +-2F - z_i: in Julia mode (fixed c) the
    point and its mirror image inside the box have the same orbit."""
    F = 1.2
    p = np.array([0.5, -0.6, 0.7])
    p[axis] = sign * 1.5
    q = p.copy()
    q[axis] = sign * 2.0 * F - p[axis]
    s = _fip_scene(F=F)
    assert _r1(s, tuple(p)) == pytest.approx(_r1(s, tuple(q)), rel=1e-14)
    # two iterations as well (the fold acts on every iteration's input)
    a, _ = oracle.compute(s, tuple(p), mode=MODE_DELTA1, n_iter=2)
    b, _ = oracle.compute(s, tuple(q), mode=MODE_DELTA1, n_iter=2)
    assert a == pytest.approx(b, rel=1e-13)


def test_folding_int_pow2_sphere_fold():
    """r^2 < 0.25: z *= 4 (then no fold applies to 4p when 1 <= |4p| < F), so
    p and 4p have the same orbit; 0.25 <= r^2 < 1: z *= 1/r^2, the inversion
    p / r^2."""
    s = _fip_scene()
    small = (0.2, -0.25, 0.15)           # |p| ~ 0.35
    assert _r1(s, small) == _r1(s, tuple(4.0 * np.array(small)))
    mid = np.array([0.5, 0.4, -0.45])    # |p| ~ 0.78
    r2 = float(mid @ mid)
    assert _r1(s, tuple(mid)) == pytest.approx(_r1(s, tuple(mid / r2)), rel=1e-13)


# --- aexion (src/fractal.cpp:916-947) ---------------------------------------
# The first iteration maps z and c to the four tetrahedral projections
# |(+x+y+z, -x-y+z, -x+y-z, +x-y-z)| + c_add; every iteration is then the 4-D
# map whose sum and difference coordinates s1 = x + y, d1 = x - y,
# s2 = z + w, d2 = z - w decouple:
#   s1' = s2^2 - d2^2 + (cx + cy),  d1' = 2 s1 d1 + (cx - cy),
#   s2' = s1^2 - d1^2 + (cz + cw),  d2' = 2 s2 d2 + (cz - cw),
# and |(x, y, z, w)|^2 = (s1^2 + d1^2 + s2^2 + d2^2) / 2.

_TETRA = np.array([[1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]], dtype=float)


def _aexion_orbit(p, c, cadd, n):
    z = np.abs(_TETRA @ np.array(p)) + cadd
    k = np.abs(_TETRA @ np.array(c)) + cadd
    s1, d1, s2, d2 = z[0] + z[1], z[0] - z[1], z[2] + z[3], z[2] - z[3]
    for _ in range(n):
        s1, d1, s2, d2 = (s2 * s2 - d2 * d2 + k[0] + k[1], 2 * s1 * d1 + k[0] - k[1],
                          s1 * s1 - d1 * d1 + k[2] + k[3], 2 * s2 * d2 + k[2] - k[3])
    return math.sqrt((s1 * s1 + d1 * d1 + s2 * s2 + d2 * d2) / 2.0)


@pytest.mark.parametrize("p", POINTS)
@pytest.mark.parametrize("cadd", [0.0, -0.35])
def test_aexion_in_sum_and_difference_coordinates(p, cadd):
    s = _scene(19, 2.0, julia=J)
    s.c_add = cadd
    for n in (1, 2):
        got, _ = oracle.compute(s, p, mode=MODE_DELTA1, n_iter=n)
        assert got == pytest.approx(_aexion_orbit(p, J, cadd, n), rel=1e-12), n


def test_aexion_constant_is_the_point_outside_julia_mode():
    """Without Julia mode c is the point itself, transformed like z."""
    p = (0.31, -0.22, 0.18)
    s = _scene(19, 2.0)
    s.c_add = -0.2
    got, _ = oracle.compute(s, p, mode=MODE_DELTA1, n_iter=1)
    assert got == pytest.approx(_aexion_orbit(p, p, -0.2, 1), rel=1e-12)


# --- benesi (src/fractal.cpp:948-970) ---------------------------------------
# With rho = |(y, z)|: x' = x^2 - rho^2 when c.x < 0 or x < rho, else
# rho^2 - x^2; (y' + i z') = -2|x| (y + iz)^2 / rho; plus c.  The map keeps
# |z1 - c| = |p|^2.

def _benesi(p, c):
    x, y, z = p
    rho = math.hypot(y, z)
    nx = x * x - rho * rho if (c[0] < 0 or x < rho) else rho * rho - x * x
    w = -2.0 * abs(x) * complex(y, z) ** 2 / rho
    return np.array([nx + c[0], w.real + c[1], w.imag + c[2]])


@pytest.mark.parametrize("p", [(0.3, -0.7, 0.45), (-1.1, 0.2, -0.6), (0.9, 0.3, -0.2), (1.4, -0.5, 0.6)])
@pytest.mark.parametrize("c", [(0.13, -0.27, 0.31), (-0.2, 0.1, -0.05)])
def test_benesi_planar_squaring_and_sign_rule(p, c):
    """Both branches of the x rule: c.x < 0 always takes x^2 - rho^2; with
    c.x >= 0 the points with x >= rho (0.9, 1.4 here) take rho^2 - x^2."""
    s = _scene(20, 2.0, julia=c)
    assert _r1(s, p) == pytest.approx(np.linalg.norm(_benesi(p, c)), rel=1e-12)


def test_benesi_squares_the_radius():
    s = _scene(20, 2.0, julia=(0.0, 0.0, 0.0))
    for p in [(0.3, -0.7, 0.45), (1.4, -0.5, 0.6), (-0.2, 0.05, 0.9)]:
        assert _r1(s, p) == pytest.approx(sum(v * v for v in p), rel=1e-12)


# --- generalizedFoldBox (src/fractal.cpp:531-749) ----------------------------
# With the cube's six face normals the fold is the Mandelbox box fold: a point
# beyond one face is mirrored in it, beyond an edge rotated half a turn about
# the edge (= mirrored in both faces), beyond a corner inverted through it (=
# mirrored in all three), so with solid = the fold limit and melt 0 the
# formula is tglad with folding value 2 x limit, up to the reference's
# 2 |d| / (|d| + 1e-8) factor.

def _box_pair(limit=1.0, offset=(0.0, 0.0, 0.0)):
    t = S.mandelbox_view(64, 48, shaded=False)
    t.mb_folding_limit, t.mb_folding_value = limit, 2.0 * limit
    for i in range(3):
        t.mb_offset[i] = offset[i]
    g = S.copy(t)
    g.formula = 37
    g.mb_fold_mode = 1
    g.mb_solid, g.mb_melt = limit, 0.0
    return t, g


@pytest.mark.parametrize("p", [(1.6, 0.3, -0.2), (1.5, 1.4, 0.2), (1.3, -1.6, 1.2), (0.4, 0.2, 0.1), (2.5, 2.2, -3.0)])
def test_generalized_fold_box_cube_is_the_mandelbox_box_fold(p):
    t, g = _box_pair(offset=(0.15, -0.2, 0.05))   # the sphere fold about the offset as well
    for n in (1, 2):
        rt, _ = oracle.compute(t, p, mode=MODE_DELTA1, n_iter=n)
        rg, _ = oracle.compute(g, p, mode=MODE_DELTA1, n_iter=n)
        assert rg == pytest.approx(rt, rel=1e-6), (p, n)


def test_generalized_fold_box_cube_distance_equals_the_mandelbox():
    t, g = _box_pair()
    rng = np.random.default_rng(37)
    checked = 0
    for _ in range(200):
        p = tuple(rng.uniform(-40, 40, 3))
        dt, mt, it, _ = oracle.calculate_distance(t, p)
        if mt or it > 10:
            continue  # a chaotic orbit near the set magnifies the 1e-8 factor
        dg, mg, ig, _ = oracle.calculate_distance(g, p)
        assert not mg and ig == it, p
        assert dg == pytest.approx(dt, rel=1e-5), p
        checked += 1
    assert checked > 40


def test_generalized_fold_box_inside_the_solid_is_only_the_sphere_fold():
    """A point inside every fold plane is not folded: with scale 1 and no
    sphere fold (radius >= fixed radius) z1 = z + c."""
    _, g = _box_pair(limit=3.0)
    g.mb_scale = 1.0
    p = (1.1, -1.2, 0.9)
    for mode in range(8):
        g.mb_fold_mode = mode
        assert _r1(g, p) == pytest.approx(2.0 * math.sqrt(sum(v * v for v in p)), rel=1e-14), mode


@pytest.mark.parametrize("mode,sides", [(0, 4), (1, 6), (2, 8), (3, 12), (4, 14), (5, 20), (6, 8), (7, 7)])
def test_generalized_fold_box_mirror_in_each_face(mode, sides):
    """Just beyond the middle of a face only that face cuts the segment from
    the origin: the point is mirrored in the face plane x.n = solid, so
    |z1 - c| = |p - 2 (p.n - solid) n| (scale 1, melt 0, the 1e-8 factor
    aside).  The faces are taken from the reference's polyhedra: every
    normal is a unit vector."""
    _, g = _box_pair(limit=1.0)
    g.mb_fold_mode = mode
    g.mb_scale = 1.0
    g.mb_fixed_radius, g.mb_min_radius = 0.5, 0.25   # sphere fold only below 0.5
    if mode == 3:
        aa = (1 + 5 ** 0.5) / 2
        bb = 1 / math.sqrt(aa * aa + 1)
        n = np.array([0.0, bb, aa * bb])
    elif mode in (0, 2, 4, 5):
        n = np.array([1.0, 1.0, -1.0]) / math.sqrt(3.0)
    else:
        n = np.array([0.0, 0.0, 1.0])
    assert abs(np.linalg.norm(n) - 1) < 1e-15
    p = n * 1.4
    want = p - 2 * (p @ n - 1.0) * n
    z1, _ = oracle.compute(g, tuple(p), mode=MODE_DELTA1, n_iter=1)
    # constant c = p (no Julia): |z1| = |want + p|
    assert z1 == pytest.approx(np.linalg.norm(want + p), rel=1e-7), (mode, sides)


# --- mandelboxVaryScale4D (src/fractal.cpp:887-915) ---------------------------
# |z + f| - |z - f| - z is the box fold at limit f (value 2f) on each of x, y,
# z and w; with scale_vary 0 the scale stays fixed, with rpower 1 the radius
# test is on r^2, with w_add 0 w stays 0: the formula is the Mandelbox with
# fold limit f, min radius minR and fixed radius 1.

def _vary_pair():
    t = S.mandelbox_view(64, 48, shaded=False)
    t.mb_fixed_radius = 1.0
    v = S.copy(t)
    v.formula = 18
    v.mb_vary_scale_vary, v.mb_vary_fold, v.mb_vary_minr = 0.0, t.mb_folding_limit, t.mb_min_radius
    v.mb_vary_rpower, v.mb_vary_wadd = 1.0, 0.0
    t.mb_folding_value = 2.0 * t.mb_folding_limit
    return t, v


def test_vary_scale_4d_reduces_to_the_mandelbox():
    t, v = _vary_pair()
    rng = np.random.default_rng(18)
    checked = 0
    for _ in range(120):
        p = tuple(rng.uniform(-40, 40, 3) if _ % 2 else rng.uniform(-5, 5, 3))
        for n in (1, 3):
            a, _ = oracle.compute(t, p, mode=MODE_DELTA1, n_iter=n)
            b, _ = oracle.compute(v, p, mode=MODE_DELTA1, n_iter=n)
            assert b == pytest.approx(a, rel=1e-12), (p, n)
        dt, mt, it, _ = oracle.calculate_distance(t, p)
        if mt or it > 12:
            continue  # chaotic orbits near the set magnify last-bit differences
        dv, mv, iv, _ = oracle.calculate_distance(v, p)
        assert not mv and iv == it, p
        assert dv == pytest.approx(dt, rel=1e-10)
        checked += 1
    assert checked > 20


def test_vary_scale_4d_scale_and_w():
    """scale_vary s: the scale becomes k + s (|k| - 1) before the first
    iteration; w_add a: w1 = a (w starts at 0, folds to 0), so
    |z1|^2 = |k' p + c|^2 + a^2 for a point inside the fold box whose radius
    needs no sphere fold (1 <= r^2)."""
    _, v = _vary_pair()
    v.mb_scale = 2.5
    v.mb_vary_scale_vary = 0.2
    v.mb_vary_wadd = 0.3
    v.mb_vary_fold = 2.0
    p = np.array([0.7, -0.6, 0.5])
    k = 2.5 + 0.2 * (2.5 - 1.0)
    z1 = k * p + p
    want = math.sqrt(z1 @ z1 + 0.3 ** 2)
    got, _ = oracle.compute(v, tuple(p), mode=MODE_DELTA1, n_iter=1)
    assert got == pytest.approx(want, rel=1e-14)


# --- IFS folding mode (src/fractal.cpp:100-126) --------------------------------
# The pre-fold is the kaleidoscopic IFS step (intensity 1, no edge fold, no
# Menger mode) run before the formula; followed by the identity formula
# (offset block with p = 0) the orbit is the kaleidoscopic formula's.

def test_ifs_folding_mode_is_the_kaleidoscopic_step():
    k = S.menger_view(64, 48, shaded=False)
    k.formula = 10
    k.ifs_menger_sponge_mode = 0
    k.ifs_scale = 2.2
    for i, d in enumerate([(1, -1, 0), (1, 0, -1), (0, 1, -1)]):
        for c in range(3):
            k.ifs_direction[i][c] = d[c]
        k.ifs_enabled[i] = 1
        k.ifs_intensity[i] = 1.0
    k.ifs_abs[0] = k.ifs_abs[1] = k.ifs_abs[2] = 1
    k.ifs_rotation[0], k.ifs_rotation[1] = 0.3, -0.2
    k.ifs_angles[1][2] = 0.25
    for c in range(3):
        k.ifs_offset[c] = 1.0
    f = S.copy(k)
    f.formula = 31        # offsetX
    f.power = 0.0         # + 0: the identity
    f.ifs_folding_mode = 1
    for p in [(0.3, -0.7, 0.45), (-1.1, 0.2, -0.6), (0.9, 0.3, -0.2), (1.4, -0.5, 0.6)]:
        for n in (1, 2, 3):
            a, _ = oracle.compute(k, p, mode=MODE_DELTA1, n_iter=n)
            b, _ = oracle.compute(f, p, mode=MODE_DELTA1, n_iter=n)
            assert b == a, (p, n)

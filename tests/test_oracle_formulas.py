"""Pins the oracle's Compute<Mode> and CalculateDistance (src/fractal.cpp)
with checks that do not retype its formulas:

* closed forms: the far-field Mandelbulb DE after one escaping iteration,
  computed with de Moivre complex powers instead of asin/atan2/pow/sin/cos;
* textbook reductions: the power-2 Mandelbulb on the z = 0 plane is the
  Mandelbrot set (Python complex arithmetic); the Menger sponge orbit
  stays bounded exactly on the Menger sponge (base-3 digit test); the
  Mandelbox with folding value 2*limit is Tglad's textbook clamp-fold box;
* invariants: mirror and (p-1)-fold rotational symmetry of the bulb,
  octahedral symmetry of sponge and box;
* the CalculateDistance rules (clamp, minN, iteration threshold, interior
  mode, primitive-only scenes) on points where Compute's value is known.
"""
import cmath
import math

import numpy as np
import pytest

from mandelbulber_amd import scene as S
from oracle import oracle as O


def bulb(power=8.0, N=250, bailout=100.0):
    s = S.defaults()
    s.formula, s.power, s.N, s.bailout, s.constant_factor = 2, power, N, bailout, 1.0
    return s


def demoivre_iter(z, c, p):
    """One trig_optim step for integer p with complex powers: (rho + i z)^p
    and (x + i y)^p give cos/sin of p*theta and p*phi."""
    x, y, zz = z
    r = math.sqrt(x * x + y * y + zz * zz)
    rho = math.hypot(x, y)
    u = complex(rho, zz) ** p / r ** p
    w = complex(x, y) ** p / rho ** p
    rp = r ** p
    return np.array([u.real * w.real * rp + c[0], u.real * w.imag * rp + c[1], u.imag * rp + c[2]])


def demoivre_de(c, p, N, bailout):
    z = np.array(c, dtype=float)
    rdz = 1.0
    for L in range(N):
        r = float(np.linalg.norm(z))
        rdz = p * r ** (p - 1) * rdz + 1.0
        z = demoivre_iter(z, c, p)
        rn = float(np.linalg.norm(z))
        if rn > bailout:
            return 0.5 * rn * math.log(rn) / rdz, L
    return 0.0, N


@pytest.mark.parametrize("p,rmin", [(8, 4.0), (3, 5.0), (2, 11.0)])
def test_bulb_far_field_closed_form(p, rmin):
    rng = np.random.default_rng(10 + p)
    s = bulb(power=p)
    for _ in range(200):
        v = rng.normal(size=3)
        c = v / np.linalg.norm(v) * rng.uniform(rmin, 3 * rmin)
        de, L = O.compute(s, c)
        z1 = demoivre_iter(c, c, p)
        r1 = float(np.linalg.norm(z1))
        r0 = float(np.linalg.norm(c))
        expect = 0.5 * r1 * math.log(r1) / (p * r0 ** (p - 1) + 1.0)
        assert L == 0
        assert de == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("p", [8, 5, 3])
def test_bulb_matches_complex_power_form(p):
    """Full DE on random points near the set, integer powers: the trig form
    and the de Moivre form agree (to rounding) for orbits of a few steps."""
    rng = np.random.default_rng(20 + p)
    s = bulb(power=p)
    checked = 0
    for _ in range(400):
        c = rng.uniform(-1.4, 1.4, size=3)
        de, L = O.compute(s, c)
        de2, L2 = demoivre_de(c, p, s.N, s.bailout)
        if L2 > 6:
            continue  # chaotic orbits: rounding differences grow
        assert L == L2
        assert de == pytest.approx(de2, rel=1e-9)
        checked += 1
    assert checked > 100


def test_power2_plane_is_mandelbrot():
    s = bulb(power=2.0, N=60)
    rng = np.random.default_rng(30)
    for _ in range(400):
        c = complex(rng.uniform(-2.2, 0.8), rng.uniform(-1.3, 1.3))
        de, L = O.compute(s, (c.real, c.imag, 0.0))
        z = c
        k = s.N
        for it in range(s.N):
            z = z * z + c
            if abs(z) > 100.0:
                k = it
                break
        if 0 < k < s.N and abs(abs(z) - 100.0) < 1e-9:
            continue
        assert L == k, c
        if k == s.N:
            assert de == 0.0
    for c in (-1.0, 0.0, 0.25, complex(-0.1, 0.1), complex(-0.75, 0.05)):
        de, L = O.compute(s, (c.real, c.imag, 0.0))
        assert (L, de) == (s.N, 0.0)


def test_bulb_mirror_symmetry_exact():
    rng = np.random.default_rng(40)
    s = bulb()
    for _ in range(300):
        x, y, z = rng.uniform(-1.3, 1.3, size=3)
        a = O.compute(s, (x, y, z))
        assert O.compute(s, (x, -y, z)) == a
        assert O.compute(s, (x, y, -z)) == a


@pytest.mark.parametrize("p", [8.0, 6.0])
def test_bulb_rotational_symmetry(p):
    """f(R z) = R^p f(z) about the z axis, so rotating the point by
    2 pi k/(p-1) rotates the whole orbit (constant included)."""
    rng = np.random.default_rng(50)
    s = bulb(power=p)
    ang = 2 * math.pi / (p - 1)
    for _ in range(300):
        c = rng.uniform(-1.3, 1.3, size=3)
        de, L = O.compute(s, c)
        if L > 6:
            continue
        k = int(rng.integers(1, int(p - 1)))
        ca, sa = math.cos(k * ang), math.sin(k * ang)
        c2 = (ca * c[0] - sa * c[1], sa * c[0] + ca * c[1], c[2])
        de2, L2 = O.compute(s, c2)
        assert L2 == L
        assert de2 == pytest.approx(de, rel=1e-9, abs=1e-15)


def menger(N=40):
    s = S.defaults()
    s.formula, s.N = 7, N
    return s


def ternary_hole_level(p, levels):
    """First level at which two coordinates of p in [-1,1]^3 fall in the
    middle third of their base-3 cell (the sponge's removed crosses)."""
    t = [(v + 1.0) / 2.0 for v in p]
    for lev in range(levels):
        digits = []
        for i in range(3):
            t[i] *= 3.0
            d = int(math.floor(t[i]))
            t[i] -= d
            digits.append(d)
        if sum(1 for d in digits if d == 1) >= 2:
            return lev
    return None


def test_menger_escape_is_digit_test():
    rng = np.random.default_rng(60)
    s = menger()
    escaped = 0
    for _ in range(3000):
        p = rng.uniform(-1, 1, size=3)
        # stay away from cell boundaries, where rounding decides
        t = (p + 1) / 2 * 3 ** 12
        if np.any(np.abs(t - np.round(t)) < 1e-3):
            continue
        lev = ternary_hole_level(list(p), 12)
        de, L = O.compute(s, p)
        if lev is not None:
            assert L < s.N and L >= lev, (p, lev, L)
            assert de > 0
            escaped += 1
    assert escaped > 1000


def test_menger_points_of_the_sponge_stay_bounded():
    rng = np.random.default_rng(61)
    s = menger(N=30)
    for _ in range(200):
        # digits with at most one coordinate in the middle third per level
        t = np.zeros(3)
        scale = 1.0
        for _lev in range(33):
            scale /= 3.0
            while True:
                d = rng.integers(0, 3, size=3)
                if np.count_nonzero(d == 1) <= 1:
                    break
            t += d * scale
        p = 2 * t - 1
        de, L = O.compute(s, p)
        assert (L, de) == (s.N, 0.0)


def test_menger_octahedral_symmetry_exact():
    rng = np.random.default_rng(62)
    s = menger()
    for _ in range(300):
        p = rng.uniform(-1.5, 1.5, size=3)
        a = O.compute(s, p)
        perm = rng.permutation(3)
        sg = rng.choice([-1.0, 1.0], size=3)
        assert O.compute(s, p[perm] * sg) == a


def mandelbox(scale=2.0, N=250):
    s = S.defaults()
    s.formula, s.N, s.mb_scale = 8, N, scale
    s.mb_folding_limit, s.mb_folding_value, s.mb_min_radius, s.mb_fixed_radius = 1.0, 2.0, 0.5, 1.0
    return s


def textbook_mandelbox(c, scale, N, limit=1.0, min_r=0.5, fixed_r=1.0):
    """Tglad's Mandelbox: box fold as a clamp reflection, sphere fold,
    z = scale*z + c, running derivative dr = dr*|scale| + 1, escape 1024."""
    z = np.array(c, dtype=float)
    dr = 1.0
    for L in range(N):
        z = 2.0 * np.clip(z, -limit, limit) - z
        r2 = float(z @ z)
        if r2 < min_r ** 2:
            f = fixed_r ** 2 / min_r ** 2
            z = z * f
            dr = dr * f
        elif r2 < fixed_r ** 2:
            f = fixed_r ** 2 / r2
            z = z * f
            dr = dr * f
        z = z * scale + np.asarray(c)
        dr = dr * abs(scale) + 1.0
        r = float(np.linalg.norm(z))
        if r > 1024:
            return r / abs(dr), L
    return 0.0, N


@pytest.mark.parametrize("scale", [2.0, -1.5, 3.0])
def test_mandelbox_matches_textbook(scale):
    rng = np.random.default_rng(70)
    s = mandelbox(scale)
    checked = 0
    for _ in range(1500):
        c = rng.uniform(-6, 6, size=3)
        de, L = O.compute(s, c)
        de2, L2 = textbook_mandelbox(c, scale, s.N)
        if L2 > 20:
            continue
        assert L == L2, c
        assert de == pytest.approx(de2, rel=1e-9)
        checked += 1
    assert checked > 100


def test_mandelbox_octahedral_symmetry():
    """Sign flips mirror the orbit exactly; axis permutations reorder the
    sums of |z|, so they agree to rounding on short orbits."""
    rng = np.random.default_rng(71)
    s = mandelbox()
    for _ in range(300):
        p = rng.uniform(-3, 3, size=3)
        a = O.compute(s, p)
        sg = rng.choice([-1.0, 1.0], size=3)
        assert O.compute(s, p * sg) == a
        if a[1] <= 20:
            b = O.compute(s, p[rng.permutation(3)])
            assert b[1] == a[1]
            assert b[0] == pytest.approx(a[0], rel=1e-9)


def test_colouring_far_field_closed_form():
    """Compute<colouring> for the bulb returns 5000 x the minimum orbit
    radius once |z| > 1e15 (src/fractal.cpp:1178-1187)."""
    s = bulb()
    rng = np.random.default_rng(80)
    for _ in range(100):
        v = rng.normal(size=3)
        c = v / np.linalg.norm(v) * rng.uniform(2.0, 3.0)
        col, L = O.compute(s, c, mode=1, n_iter=s.N * 10)
        z1 = demoivre_iter(c, c, 8)
        z2 = demoivre_iter(z1, c, 8)
        r1, r2 = np.linalg.norm(z1), np.linalg.norm(z2)
        assert r2 > 1e15 > r1
        assert L == 1
        assert col == pytest.approx(5000.0 * r1, rel=1e-12)


def test_calculate_distance_rules():
    s = bulb()
    far = (5.0, 1.0, 0.5)  # escapes at L = 0
    raw, L = O.compute(s, far)
    assert L == 0
    # minN rule: L < minN and distance < detailSize -> detailSize
    d, mi, it, _ = O.calculate_distance(s, far, detail=raw * 2, min_n=1)
    assert (d, mi, it) == (raw * 2, False, 0)
    d, _, _, _ = O.calculate_distance(s, far, detail=raw * 2, min_n=0)
    assert d == raw
    # inside: L == N -> distance 0, max_iter
    d, mi, it, _ = O.calculate_distance(s, (0.0, 0.0, 0.0001), detail=1e-3)
    assert (d, mi, it) == (0.0, True, s.N)
    # iteration threshold mode: L < N and distance < detail -> 1.01 detail
    s.iter_thresh = 1
    d, _, _, _ = O.calculate_distance(s, far, detail=raw * 3, min_n=0)
    assert d == pytest.approx(raw * 3 * 1.01, rel=1e-15)
    s.iter_thresh = 0
    # interior mode: L == N -> detailSize, not max_iter
    s.interior_mode = 1
    d, mi, _, _ = O.calculate_distance(s, (0.0, 0.0, 0.0001), detail=1e-3)
    assert (d, mi) == (1e-3, False)
    # ... and in normal-calculation mode the distance is reflected
    d, _, _, _ = O.calculate_distance(s, far, detail=raw * 2, min_n=0, normal_mode=1)
    assert d == pytest.approx(raw * 2 - raw, rel=1e-15)
    s.interior_mode = 0
    # primitive-only scene: the fractal contributes 10
    s.primitive_only_plane = 1
    d, mi, _, obj = O.calculate_distance(s, far, detail=1e-3)
    assert (d, mi) == (10.0, False)
    s.plane_enabled = 1
    s.plane_centre[:] = (0, 0, 0)
    s.plane_normal[:] = (0, 0, -2)
    d, _, _, obj = O.calculate_distance(s, (0.3, 0.2, -0.7), detail=1e-3)
    assert d == pytest.approx(0.7, rel=1e-15)

"""Pins the oracle's ray march and shaders with scenes whose answers are fixed
by geometry: exact primitive SDFs (primitive-only scenes, the reference's
primitive_only_plane switch), Lambert shading, the AO of a plane, the
shadow of a sphere on a plane, the palette lerp, the background gradient and
the back-to-front step grouping of VolumetricShader on hand-made step
records.
"""
import math

import numpy as np
import pytest

from mandelbulber_amd import scene as S
from oracle import oracle as O


def prim_scene(w=64, h=48):
    s = S.defaults()
    s.primitive_only_plane = 1
    s.image_width, s.image_height = w, h
    s.vp[:] = (0.0, -3.0, 0.0)
    s.alpha = s.beta = s.gamma = 0.0
    s.persp, s.zoom = 1.0, 1e-15
    s.shadow = 0
    s.global_illumination = 0
    s.glow_intensity = 0.0
    s.fog_density = 0.0
    return s


def test_sphere_hit_lies_in_the_refinement_band():
    """RayMarching ends with the exact SDF in [thr*(1-0.01q), thr] of the
    final threshold, at the depth of the analytic ray-sphere intersection."""
    s = prim_scene()
    s.sphere_enabled = 1
    s.sphere_centre[:] = (0.1, 0.2, -0.1)
    s.sphere_radius = 1.0
    _, start, _ = O.camera(s)
    rng = np.random.default_rng(1)
    hits = 0
    for _ in range(300):
        x, y = int(rng.integers(0, s.image_width)), int(rng.integers(0, s.image_height))
        d, _ = O.view_vector(s, x, y)
        m = O.march(s, start, d, pixel=x + y * s.image_width)
        c = np.array(s.sphere_centre[:])
        # analytic intersection |start + t d - c| = R
        oc = start - c
        A, B, Cq = d @ d, 2 * d @ oc, oc @ oc - 1.0
        disc = B * B - 4 * A * Cq
        if disc < 0:
            # a miss by less than the hit threshold still counts as a hit
            # (dist < distThresh, src/Render3D.cpp:123)
            tc = -(oc @ d) / A
            gap = np.linalg.norm(oc + tc * d) - 1.0
            thr_c = tc / s.image_width * s.persp / s.quality
            if gap > 2 * thr_c:
                assert not m.found
                assert m.depth > s.view_distance_max
            continue
        if not m.found:
            # grazing rays may slip through: the analytic chord is tiny
            assert disc / (A * A) < 1e-2
            continue
        hits += 1
        p = np.array(m.point[:])
        sdf = np.linalg.norm(p - c) - 1.0
        thr = m.dist_thresh
        assert thr * (1 - 0.01 * s.quality) - 1e-15 <= sdf <= thr + 1e-15 or m.de_calls > 100
        t_hit = (-B - math.sqrt(disc)) / (2 * A)
        # the band is within ~thr of the surface along the normal
        assert abs(m.depth - t_hit) < 2 * thr * math.sqrt(A) / max(abs(-(B + 2 * A * t_hit)) / (2 * math.sqrt(A)), 1e-3)
    assert hits > 50


def test_plane_depth_is_planar():
    """The primary view vector is not normalised (src/Render3D.cpp:431-436),
    so the z-buffer of a plane facing the camera is the same for every
    pixel: the camera-forward distance minus the hit threshold."""
    s = prim_scene(48, 36)
    s.plane_enabled = 1
    s.plane_centre[:] = (0.0, 2.0, 0.0)
    s.plane_normal[:] = (0.0, -1.0, 0.0)
    s.jitter_mode = 0
    fr, st = O.render(s)
    assert st.hits == s.image_width * s.image_height
    D = 2.0 - (-3.0)
    res = 1.0 / s.image_width
    z = fr.z.astype(np.float64)
    thr = z * res * s.persp / s.quality
    # hit depth t with D - t in (0.99 thr, thr)
    assert np.all(D - z > 0.99 * thr - 1e-6) and np.all(D - z < thr + 1e-6)
    assert np.ptp(z) < 1e-5


def test_lambert_shading_on_a_plane():
    s = prim_scene()
    s.plane_enabled = 1
    s.plane_centre[:] = (0.0, 0.0, 0.5)
    s.plane_normal[:] = (0.0, 0.0, -1.0)
    s.main_light_alpha, s.main_light_beta = 0.3, 0.9
    s.shading, s.direct_light, s.specular = 0.8, 0.7, 0.0
    _, _, light = O.camera(s)
    thr = 2e-3
    p = (0.2, 0.0, 0.5 - 0.995 * thr)
    out = O.object_shader(s, p, (0.0, 1.0, 0.2), thr, 3.0, obj=1)
    n = np.array(out.normal[:])
    assert np.allclose(n, (0, 0, -1), atol=1e-9)
    expect = 0.7 * ((1 - 0.8) + 0.8 * max(0.0, -light[2]))
    assert out.shade == pytest.approx(expect, rel=1e-9)
    # plane colour 20000/65536, main light white 65535/65536
    assert out.rgb[0] == pytest.approx(65535 / 65536 * expect * 20000 / 65536, rel=1e-9)


def test_sphere_shadow_on_a_plane():
    """Light straight down (-z is up in world coordinates here): the floor
    point under the sphere is in shadow, a point far to the side is lit."""
    s = prim_scene()
    s.plane_enabled = 1
    s.plane_centre[:] = (0.0, 0.0, 1.0)
    s.plane_normal[:] = (0.0, 0.0, -1.0)
    s.sphere_enabled = 1
    s.sphere_centre[:] = (0.0, 0.0, 0.0)
    s.sphere_radius = 0.5
    s.shadow = 1
    s.penetrating_lights = 0
    # light vector (cos(a-pi/2)cos(-b), sin(a-pi/2)cos(-b), sin(-b)) = (0,0,-1)
    s.main_light_alpha, s.main_light_beta = 0.0, math.pi / 2
    _, _, light = O.camera(s)
    assert np.allclose(light, (0, 0, -1), atol=1e-12)
    thr = 2e-3
    under = O.object_shader(s, (0.0, 0.0, 1.0 - 0.995 * thr), (0, 1, 0), thr, 3.0, obj=1)
    side = O.object_shader(s, (2.5, 0.0, 1.0 - 0.995 * thr), (0, 1, 0), thr, 3.0, obj=1)
    assert under.shadow == 0.0
    assert side.shadow == 1.0


def test_ao_of_a_plane():
    """AO directions are the full sphere (src/Render3D.cpp:271-313).  From a
    point at distance h in the refinement band above a plane, the first step
    r = delta decides every direction: those with h + (v.n) delta below
    the (float) threshold are occluded with intensity r/end = resolution, all
    others climb away for ever with intensity 1."""
    s = prim_scene(100, 80)
    s.plane_enabled = 1
    s.plane_centre[:] = (0.0, 0.0, 0.5)
    s.plane_normal[:] = (0.0, 0.0, -1.0)
    s.global_illumination = 1
    s.ao_quality = 8
    thr = 3e-3
    h = 0.995 * thr
    out = O.object_shader(s, (0.1, 0.0, 0.5 - h), (0, 1, 0), thr, 3.0, obj=1)
    dirs, rgb = O.ao_table(8)
    assert len(dirs) == 412
    delta = thr * s.quality
    res = 1.0 / s.image_width
    n = np.array((0.0, 0.0, -1.0))
    occl = h + (dirs @ n) * delta < np.float32(thr)
    inten = np.where(occl, delta / (delta / res), 1.0)
    expect = (inten * rgb[:, 0]).sum() / (len(dirs) * 256.0)
    assert out.ao[0] == pytest.approx(expect, rel=1e-12)


def test_ao_table_counts():
    """Closed-form count of the AO table: one row per b in
    [-0.49 pi, 0.49 pi) stepping 1/q, per row ceil(2 pi / step_a(b)) columns
    with step_a = (2/q)/cos(b), all kept with the white fallback lightmap."""
    for q in (1, 2, 4, 8, 16):
        dirs, rgb = O.ao_table(q)
        count = 0
        b = -0.49 * math.pi
        while b < 0.49 * math.pi:
            step = (2.0 / q) / math.cos(b)
            count += math.ceil(2 * math.pi / step - 1e-12)
            b += 1.0 / q
        assert len(dirs) == count, q
        assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0)
        assert np.all(rgb == 255)
    # a black lightmap keeps no direction: the reference's fallback entry
    dirs, rgb = O.ao_table(4, lightmap=np.zeros((10, 10, 3), np.uint8))
    assert len(dirs) == 1 and np.all(dirs == 0) and np.all(rgb == 0)


def test_index_to_colour_is_palette_lerp():
    s = S.defaults()
    S.grey_palette(s)
    pal = np.array([[s.palette[i].R, s.palette[i].G, s.palette[i].B] for i in range(256)], float)
    rng = np.random.default_rng(2)
    for idx in list(rng.integers(0, 65536, size=300)) + [0, 255, 256, 65279, 65280, 65535]:
        idx = int(idx)
        k = idx % (255 * 256)
        kol, frac = k // 256, k % 256
        expect = np.trunc(pal[kol] + (pal[kol + 1] - pal[kol]) / 256.0 * frac).astype(int)
        assert O.index_to_colour(s, idx) == tuple(expect), idx
    assert O.index_to_colour(s, -5) == tuple(pal[255].astype(int))


def test_background_gradient():
    s = S.defaults()
    c1, c2, c3 = (np.array([getattr(c, k) for k in "RGB"], float) for c in
                  (s.background_color1, s.background_color2, s.background_color3))
    for vz in (-0.45, -0.1, 0.0, 0.2, 0.49):
        out = O.background(s, (0.3, 1.0, vz))
        g = 1.0 - vz
        if g < 1:
            px = np.trunc(c3 * (1 - g) + c2 * g)
        else:
            px = np.trunc(c2 * (2 - g) + c1 * (g - 1))
        assert np.allclose(out, px / 65536.0, rtol=0, atol=1e-15), vz


def vol_scene(glow=0.0, fog=0.0):
    s = S.defaults()
    s.glow_intensity = glow
    s.fog_density = fog
    s.quality = 1.0
    s.DE_factor = 1.0
    s.post_fog_enabled = 0
    return s


def test_volumetric_glow_single_group_and_index0_excluded():
    s = vol_scene(glow=2.0)
    px = (0.2, 0.4, 0.6, 1.0)
    depth = 3.0
    # two records: index 1 forms one group, index 0 is never visited
    steps = [[0.5, 0.3, 0.01], [0.25, 0.2, 0.01]]
    out, op = O.volumetric(s, depth, steps, px)
    glow = 2 * 2.0 / 512.0
    g1 = np.array([getattr(s.glow_color1, k) for k in "RGB"], float)
    g2 = np.array([getattr(s.glow_color2, k) for k in "RGB"], float)
    gc = (g1 * (1 - glow) + g2 * glow) / 65536.0
    go = glow * 0.25 / depth
    assert np.allclose(out[:3], go * gc + (1 - go) * np.array(px[:3]), rtol=1e-14)
    assert out[3] == pytest.approx(min(1.0, 1.0 + go))
    # no fog: opacity stays 0
    assert op == 0.0


def test_volumetric_groups_merge_back_to_front():
    """Steps shorter than thr*quality are merged into the next nearer record
    (src/shaders.cpp:227-232): records [a, b(small), c] act as two groups:
    c alone, then a+b at record 1's distance."""
    s = vol_scene(fog=0.5)
    px = (0.1, 0.1, 0.1, 1.0)
    thr = 0.05
    recs = [[9.0, 9.0, thr], [0.3, 0.4, thr], [0.01, 0.7, thr], [0.2, 0.9, thr]]
    out, op = O.volumetric(s, 2.0, recs, px)

    def fog_apply(col, op_, step, dist):
        F = s.fog_distance_factor
        dens = step * F / (dist * dist + F * F)
        k = min(1.0, dist / s.fog_colour1_distance)
        f1 = np.array([getattr(s.fog_colour1, c) for c in "RGB"], float)
        f2 = np.array([getattr(s.fog_colour2, c) for c in "RGB"], float)
        f3 = np.array([getattr(s.fog_colour3, c) for c in "RGB"], float)
        fc = f1 * (1 - k) + f2 * k
        k2 = min(1.0, dist / s.fog_colour2_distance * k)
        fc = fc * (1 - k2) + f3 * k2
        fd = min(1.0, 0.3 * 0.5 * dens / (1 + 0.5 * dens))
        return fd * fc / 65536.0 + (1 - fd) * col, np.float32(fd + (1 - fd) * op_)

    col = np.array(px[:3])
    op_ = np.float32(0.0)
    col, op_ = fog_apply(col, op_, 0.2, 0.9)          # record 3 alone
    col, op_ = fog_apply(col, op_, 0.01 + 0.3, 0.4)   # record 2 merged into record 1
    assert np.allclose(out[:3], col, rtol=1e-12)
    assert op == pytest.approx(float(op_), rel=1e-6)


def test_small_frame_against_brute_force_occupancy():
    """Hit/miss of a 32x24 Mandelbulb frame against dense fixed-step sampling
    of the escape-time inside test (no DE involved): every ray the oracle
    marks as a hit passes within the hit threshold of a sampled inside
    point or a point the DE places within a few thresholds, and rays that
    pass through sampled inside points are hits."""
    s = S.mandelbulb_plain(32, 24)
    s.N = 40
    s.jitter_mode = 0
    fr, _ = O.render(s)
    _, start, _ = O.camera(s)
    disagree = 0
    for y in range(s.image_height):
        for x in range(s.image_width):
            d, _ = O.view_vector(s, x, y)
            ts = np.arange(0.5, 6.0, 0.002)
            inside = False
            for t in ts:
                p = start + d * t
                if np.linalg.norm(p) > 1.3:
                    continue
                _, L = O.compute(s, p)
                if L == s.N:
                    inside = True
                    break
            hit = fr.z[y, x] < 1e19
            if inside and not hit:
                disagree += 1
    assert disagree == 0


@pytest.mark.parametrize("merge", [1, 3])
def test_volumetric_constant_steps_closed_form(merge):
    """VolumetricShader (src/shaders.cpp:185-384) over n identical step
    records: records n-1 .. 1 are visited (index 0 never), merged into
    groups of `merge` records (each record's step is 1/merge of thr *
    quality, and 0.4 of it for merge 3: groups close at 1.2 thr quality;
    the records left over at the near end form no group), and every group
    applies the same glow blend (opacity g = glow step / depth) and the same
    volumetric-fog blend (density fd), so the pixel follows the affine
    recurrence x <- A + B x m times: x_m = A (1 - B^m) / (1 - B) + B^m x_0
    with B = (1 - fd)(1 - g), A = fd F + (1 - fd) g G for the colour,
    B = 1 - fd, A = fd + (1 - fd) g for alpha, and the fog opacity is
    1 - (1 - fd)^m.  The fog colours are all F, so the fog colour does not
    depend on the distance."""
    s = vol_scene(glow=0.3, fog=0.8)
    for i, c in enumerate((s.fog_colour1, s.fog_colour2, s.fog_colour3)):
        c.R, c.G, c.B = 50000, 20000, 9000
    s.glow_color1.R, s.glow_color1.G, s.glow_color1.B = 60000, 30000, 1000
    s.glow_color2.R, s.glow_color2.G, s.glow_color2.B = 4000, 40000, 65000
    n, depth, thr, dist = 41, 4.0, 0.02, 1.7
    step = thr * s.quality * (0.4 if merge == 3 else 1.0)
    px = np.array([0.3, 0.5, 0.7])
    out, op = O.volumetric(s, depth, [[step, dist, thr]] * n, (px[0], px[1], px[2], 0.25))
    m = (n - 1) // merge
    gstep = step * merge
    glow = n * s.glow_intensity / 512.0 * s.DE_factor
    G = (np.array([s.glow_color1.R, s.glow_color1.G, s.glow_color1.B], float) * (1 - glow)
         + np.array([s.glow_color2.R, s.glow_color2.G, s.glow_color2.B], float) * glow) / 65536.0
    g = glow * gstep / depth
    dens = gstep * s.fog_distance_factor / (dist * dist + s.fog_distance_factor ** 2)
    fd = 0.3 * s.fog_density * dens / (1 + s.fog_density * dens)
    F = np.array([50000, 20000, 9000], float) / 65536.0
    B = (1 - fd) * (1 - g)
    A = fd * F + (1 - fd) * g * G
    want = A * (1 - B ** m) / (1 - B) + B ** m * px
    assert m == (40 if merge == 1 else 13)
    np.testing.assert_allclose(out[:3], want, rtol=1e-12)
    Ba, Aa = 1 - fd, fd + (1 - fd) * g
    assert out[3] == pytest.approx(Aa * (1 - Ba ** m) / (1 - Ba) + Ba ** m * 0.25, rel=1e-12)
    assert op == pytest.approx(1 - (1 - fd) ** m, rel=1e-5)  # totalOpacity is a float

"""Scene helpers: the reference's settings model (.fract keys) over mbr_scene.

``defaults()`` is the reference defaults file, ``load()`` is LoadSettings
(src/settings.cpp:611-698) with ``-o key=value`` overrides, and ``set()``
is LoadOneSetting (:700) for one key.  The workloads the benchmark and the
tests use are built here from their .fract keys.
"""
import ctypes as C
import math

import numpy as np

from . import _abi
from ._lib import check, lib


def defaults():
    s = _abi.Scene()
    check(lib().mbr_scene_defaults(C.byref(s)))
    return s


def load(path, overrides=()):
    s = _abi.Scene()
    arr = (C.c_char_p * max(1, len(overrides)))(*[o.encode() for o in overrides])
    check(lib().mbr_load_settings(path.encode(), arr, len(overrides), C.byref(s)))
    return s


def set(scene, key, value):
    """Set one .fract key (units as in the settings file: degrees for angle_*)."""
    if isinstance(value, float):
        value = repr(value)
    check(lib().mbr_set_setting(C.byref(scene), key.encode(), str(value).encode()))
    return scene


def apply(scene, keys):
    for k, v in keys.items():
        set(scene, k, v)
    return scene


def keyframe(keys, frames_per_key, frame, base):
    """The scene of animation frame `frame`: base with every double parameter
    (and N) interpolated through the keyframe scenes `keys` (Catmull-Rom,
    src/morph.cpp:49-109)."""
    arr = (_abi.Scene * len(keys))(*keys)
    out = copy(base)
    check(lib().mbr_keyframe_scene(arr, len(keys), frames_per_key, frame, C.byref(out)))
    return out


def copy(scene):
    s = _abi.Scene()
    C.memmove(C.byref(s), C.byref(scene), C.sizeof(_abi.Scene))
    return s


def grey_palette(scene):
    """An explicit palette (the reference draws a random one from glibc rand
    when a scene has none, src/settings.cpp:1185-1195)."""
    for i in range(256):
        v = (i * 37) % 256
        scene.palette[i].R, scene.palette[i].G, scene.palette[i].B = v, (v * 3) % 256, 255 - v
    return scene


# ------------------------------------------------------------- workloads
def mandelbulb_plain(width=320, height=240, bailout=100.0):
    """Mandelbulb power 8, single frame, N 250, camera at (0,-3.5,0) looking at
    the origin, fov 1, no shadows/AO/fog."""
    s = defaults()
    apply(s, {
        "formula": 2, "power": 8, "N": 250, "minN": 1, "quality": 1, "DE_factor": 1,
        "image_width": width, "image_height": height,
        "view_point_x": 0, "view_point_y": -3.5, "view_point_z": 0,
        "angle_alfa": 0, "angle_beta": 0, "angle_gamma": 0, "perspective": 1, "zoom": 1e-15,
        "shadows_enabled": 0, "ambient_occlusion_enabled": 0, "volumetric_fog_density": 0,
        "post_fog_enabled": 0, "bailout": bailout,
    })
    grey_palette(s)
    return s


def light_from_seed(seed):
    """Unit light direction u ~ Uniform(S^2) from a seed, as main_light_alfa /
    main_light_beta (the inverse of src/Render3D.cpp:253-255)."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    beta = -math.asin(u[2])
    alpha = math.atan2(u[1], u[0]) + math.pi / 2
    return alpha, beta


def mandelbulb_full(width=1920, height=1080, seed=1):
    """Mandelbulb power 8 with one hard-shadow ray per hit toward a seeded
    unit light direction, AO quality 8 (412 directions) and volumetric fog
    density 0.5."""
    s = mandelbulb_plain(width, height)
    a, b = light_from_seed(seed)
    apply(s, {
        "shadows_enabled": 1, "shadows_cone_angle": 0, "penetrating_lights": 0,
        "main_light_alfa": a, "main_light_beta": b,
        "ambient_occlusion_enabled": 1, "fast_ambient_occlusion_mode": 0, "ambient_occlusion_quality": 8,
        "volumetric_fog_density": 0.5,
    })
    return s


MANDELBOX_KEYS = {
    "formula": 8, "mandelbox_scale": 2, "mandelbox_folding_limit": 1, "mandelbox_folding_value": 2,
    "mandelbox_folding_min_radius": 0.5, "mandelbox_folding_fixed_radius": 1, "mandelbox_rotation_enabled": 0,
    "N": 250,
}


def mandelbox_view(width=320, height=240, shaded=True):
    """Mandelbox scale 2, fold limit 1 (value 2), min radius 0.5, fixed
    radius 1, N 250, seen whole from (0,-12,0), yaw 30, pitch 20; with
    shaded, hard shadow, AO quality 4 and fog 0.5."""
    s = mandelbulb_plain(width, height)
    apply(s, dict(MANDELBOX_KEYS, view_point_y=-12, angle_alfa=30, angle_beta=20))
    if shaded:
        _shade(s, seed=2, ao_quality=4)
    return s


def menger_view(width=320, height=240, shaded=True):
    """Menger sponge, N 15, seen from (0,-3.5,0), yaw 30, pitch 20; with
    shaded, hard shadow, AO quality 4 and fog 0.5."""
    s = mandelbulb_plain(width, height)
    apply(s, {"formula": 7, "N": 15, "view_point_y": -3.5, "angle_alfa": 30, "angle_beta": 20})
    if shaded:
        _shade(s, seed=3, ao_quality=4)
    return s


def _shade(s, seed, ao_quality):
    a, b = light_from_seed(seed)
    apply(s, {
        "shadows_enabled": 1, "shadows_cone_angle": 0, "penetrating_lights": 0,
        "main_light_alfa": a, "main_light_beta": b,
        "ambient_occlusion_enabled": 1, "fast_ambient_occlusion_mode": 0, "ambient_occlusion_quality": ao_quality,
        "volumetric_fog_density": 0.5,
    })


def look_at_origin(scene, vp, roll=0.0):
    """Place the camera at vp looking at the origin.  The forward vector of
    the reference camera is M (0,1,0) with M = Rz(alfa) Rx(beta) Ry(gamma)
    (src/algebra.cpp:94-143), i.e. (-sin a cos b, cos a cos b, sin b)."""
    vp = np.asarray(vp, dtype=np.float64)
    f = -vp / np.linalg.norm(vp)
    beta = math.asin(max(-1.0, min(1.0, f[2])))
    alpha = math.atan2(-f[0], f[1])
    apply(scene, {"view_point_x": float(vp[0]), "view_point_y": float(vp[1]), "view_point_z": float(vp[2]),
                  "angle_alfa": math.degrees(alpha), "angle_beta": math.degrees(beta),
                  "angle_gamma": math.degrees(roll)})
    return scene


FLIGHT_RESOLUTIONS = ((640, 480), (800, 600), (1280, 720), (1920, 1080))


def flight(n=64, seed=0, res_div=1):
    """A batched flight of n frames: formulas cycling Mandelbulb (power
    uniform in [6,10]), Mandelbox (scale 2, fold limit 1, min radius 0.5,
    fixed radius 1), Menger sponge (N 15); resolution drawn from 640x480,
    800x600, 1280x720, 1920x1080; camera on a seeded direction at distance
    uniform in [1.5,5] looking at the origin; one hard-shadow ray per hit.
    The distance is in units of the fractal's size: the scale-2 Mandelbox
    reaches about twice as far from the origin as the Mandelbulb and the
    sponge, so its cameras sit at twice the drawn distance (at the drawn
    distance half of them would start inside the solid, where every ray
    hits at depth ~0).  res_div divides every resolution (small copies for
    tests)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        w, h = FLIGHT_RESOLUTIONS[int(rng.integers(len(FLIGHT_RESOLUTIONS)))]
        s = mandelbulb_plain(w // res_div, h // res_div)
        kind = k % 3
        if kind == 0:
            apply(s, {"formula": 2, "power": float(rng.uniform(6.0, 10.0))})
        elif kind == 1:
            apply(s, MANDELBOX_KEYS)
        else:
            apply(s, {"formula": 7, "N": 15})
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        look_at_origin(s, u * rng.uniform(1.5, 5.0) * (2.0 if kind == 1 else 1.0))
        a, b = light_from_seed(int(rng.integers(1 << 30)))
        apply(s, {"shadows_enabled": 1, "shadows_cone_angle": 0, "penetrating_lights": 0,
                  "main_light_alfa": a, "main_light_beta": b, "frame_no": k})
        out.append(s)
    return out


DIFF_PARAMS = ("power", "vp_x", "vp_y", "vp_z", "alpha", "beta", "gamma", "light_alpha", "light_beta")


def get_params(scene):
    """The differentiable parameters of a scene (radians for the angles)."""
    return [scene.power, scene.vp[0], scene.vp[1], scene.vp[2], scene.alpha, scene.beta, scene.gamma,
            scene.main_light_alpha, scene.main_light_beta]


def set_params(scene, theta):
    """Write the nine differentiable parameters into a scene (in place)."""
    t = [float(v) for v in theta]
    scene.power = t[0]
    scene.vp[0], scene.vp[1], scene.vp[2] = t[1], t[2], t[3]
    scene.alpha, scene.beta, scene.gamma = t[4], t[5], t[6]
    scene.main_light_alpha, scene.main_light_beta = t[7], t[8]
    return scene


def differentiable(width=512, height=512, seed=0):
    """A Mandelbulb scene for gradient work: power uniform in [6,10], the
    camera on a seeded direction at distance uniform in [2.5,4] looking at
    the origin with a seeded roll, a seeded light direction; Lambert and
    specular shading only (no shadow, AO, glow or fog, whose contributions
    are steps in the parameters), surface colouring off (its palette index
    is an integer), no step jitter, refined hits (hit_refine)."""
    rng = np.random.default_rng(seed)
    s = mandelbulb_plain(width, height)
    apply(s, {"power": float(rng.uniform(6.0, 10.0)), "glow_intensity": 0, "fractal_color": 0,
              "jitter_mode": 0, "hit_refine": 1})
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    look_at_origin(s, u * rng.uniform(2.5, 4.0), roll=float(rng.uniform(-0.5, 0.5)))
    a, b = light_from_seed(int(rng.integers(1 << 30)))
    apply(s, {"main_light_alfa": a, "main_light_beta": b})
    return s


def _mandelbox_distance(p, n_iter=250):
    """Distance estimate of the Mandelbox of MANDELBOX_KEYS (scale 2, fold
    limit 1, fold value 2, min radius 0.5, fixed radius 1, escape 1024),
    FP64, for placing a camera on its surface; the renders use the
    library's kernels."""
    c = np.array(p, dtype=np.float64)
    z = c.copy()
    dz = 1.0
    for _ in range(n_iter):
        z = np.where(z > 1.0, 2.0 - z, np.where(z < -1.0, -2.0 - z, z))
        r2 = float(z @ z)
        if r2 < 0.25:
            z *= 4.0
            dz *= 4.0
        elif r2 < 1.0:
            z *= 1.0 / r2
            dz *= 1.0 / r2
        z = z * 2.0 + c
        dz = dz * 2.0 + 1.0
        r = float(np.sqrt(z @ z))
        if r > 1024.0:
            return r / abs(dz)
    return 0.0


def deep_zoom(width=3840, height=2160, seed=0, frame_width=1e-6):
    """Deep-zoom Mandelbox: the camera sits 10 zoom units in front of a
    seeded surface point (reached by marching a seeded direction from
    radius 8 toward the origin), looking along that direction with a seeded
    roll; zoom = frame_width / aspect, so the frame is frame_width wide at
    the view point.  Hard shadow, no AO, no fog."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    p = 8.0 * u
    for _ in range(100000):
        d = _mandelbox_distance(p)
        if d < 1e-13:
            break
        p = p - u * d
    zoom = frame_width * height / width
    s = mandelbulb_plain(width, height)
    apply(s, dict(MANDELBOX_KEYS, zoom=zoom))
    fwd = -u
    vp = p - fwd * (10.0 * zoom)
    # orientation: forward along -u (look_at_origin from a point on the ray
    # through the origin), plus a seeded roll
    look_at_origin(s, u * np.linalg.norm(vp), roll=float(rng.uniform(-np.pi, np.pi)))
    apply(s, {"view_point_x": float(vp[0]), "view_point_y": float(vp[1]), "view_point_z": float(vp[2])})
    a, b = light_from_seed(int(rng.integers(1 << 30)))
    apply(s, {"shadows_enabled": 1, "shadows_cone_angle": 0, "penetrating_lights": 0,
              "main_light_alfa": a, "main_light_beta": b})
    return s


def place_lights(scene):
    """PlaceRandomLights: aux_light_number + 4 mbr_light records (host, FP64)."""
    from . import _abi
    import ctypes as C
    n = max(scene.aux_light_number, 0) + 4
    tab = (_abi.Light * n)()
    cnt = C.c_int()
    rc = lib().mbr_place_lights(C.byref(scene), tab, n, C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"mbr_place_lights failed: {rc}")
    return list(tab)

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

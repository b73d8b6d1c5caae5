"""ctypes loader of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports this file.
"""
import ctypes as C
import os

import numpy as np

from mandelbulber_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))

JITTER_NONE, JITTER_HASH, JITTER_GLIBC = 0, 1, 2


class OracleStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "hits", "march_de_calls", "shade_de_calls",
                                          "iterations", "missed_de", "max_steps", "total_steps")]


class MarchOut(C.Structure):
    _fields_ = [("point", C.c_double * 3), ("depth", C.c_double), ("dist_thresh", C.c_double),
                ("last_dist", C.c_double), ("found", C.c_int), ("step_count", C.c_int),
                ("object", C.c_int), ("de_calls", C.c_int)]


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise OSError(f"oracle library {path} is not built (run `make -C oracle`)")
    lib = C.CDLL(path)
    P = C.POINTER
    S = P(_abi.Scene)
    D3 = C.c_double * 3
    lib.oracle_jitter_draw.argtypes = [C.c_uint32] * 4
    lib.oracle_jitter_draw.restype = C.c_uint32
    lib.oracle_compute.argtypes = [S, C.c_int, D3, C.c_int, P(C.c_int)]
    lib.oracle_compute.restype = C.c_double
    lib.oracle_calculate_distance.argtypes = [S, D3, C.c_double, C.c_int, C.c_int, C.c_int,
                                              P(C.c_int), P(C.c_int), P(C.c_int)]
    lib.oracle_calculate_distance.restype = C.c_double
    lib.oracle_camera.argtypes = [S, C.c_double * 9, D3, D3]
    lib.oracle_view_vector.argtypes = [S, C.c_int, C.c_int, D3]
    lib.oracle_view_vector.restype = C.c_int
    lib.oracle_ao_table.argtypes = [C.c_int, P(_abi.Texture), P(C.c_double), P(C.c_int), C.c_int]
    lib.oracle_ao_table.restype = C.c_int
    lib.oracle_march.argtypes = [S, D3, D3, C.c_uint32, P(MarchOut)]
    lib.oracle_march.restype = C.c_int
    lib.oracle_set_textures.argtypes = [P(_abi.Texture), P(_abi.Texture)]
    lib.oracle_env_mapping.argtypes = [D3, D3, D3]
    lib.oracle_render.argtypes = [S, P(_abi.Texture), C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                  P(_abi.Framebuffer), P(OracleStats)]
    lib.oracle_render.restype = C.c_int
    lib.oracle_rotation3.argtypes = [D3, C.c_double * 9]
    lib.oracle_volumetric.argtypes = [S, C.c_double, P(C.c_double), C.c_int, C.c_double * 4, C.c_double * 4,
                                      P(C.c_double)]
    lib.oracle_index_to_colour.argtypes = [S, C.c_int, C.c_int * 3]
    lib.oracle_background.argtypes = [S, D3, D3]
    lib.oracle_object_shader.argtypes = [S, D3, D3, C.c_double, C.c_double, C.c_double, C.c_int,
                                         P(ShadeOut)]
    lib.oracle_object_shader.restype = C.c_int
    lib.oracle_ssao.argtypes = [S, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p * 3, C.c_void_p]
    lib.oracle_ssao.restype = C.c_int
    lib.oracle_dof_order.argtypes = [S, C.c_int, C.c_int, C.c_void_p, C.c_uint32, C.c_void_p]
    lib.oracle_dof_order.restype = C.c_int
    lib.oracle_dof.argtypes = [S, C.c_int, C.c_int, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    lib.oracle_dof.restype = C.c_int
    lib.oracle_place_lights.argtypes = [S, P(_abi.Light), C.c_int]
    lib.oracle_place_lights.restype = C.c_int
    return lib


class ShadeOut(C.Structure):
    _fields_ = [("normal", C.c_double * 3), ("shade", C.c_double), ("shadow", C.c_double),
                ("specular", C.c_double), ("ao", C.c_double * 3), ("colour", C.c_double * 3),
                ("rgb", C.c_double * 3), ("spec_rgb", C.c_double * 3)]


def rotation3(angles):
    m = (C.c_double * 9)()
    lib().oracle_rotation3((C.c_double * 3)(*angles), m)
    return np.array(m[:]).reshape(3, 3)


def volumetric(scene, depth, steps, pixel):
    steps = np.ascontiguousarray(np.asarray(steps, dtype=np.float64).reshape(-1, 3))
    out = (C.c_double * 4)()
    op = C.c_double(0)
    lib().oracle_volumetric(C.byref(scene), depth, steps.ctypes.data_as(C.POINTER(C.c_double)), len(steps),
                            (C.c_double * 4)(*pixel), out, C.byref(op))
    return np.array(out[:]), op.value


def index_to_colour(scene, index):
    o = (C.c_int * 3)()
    lib().oracle_index_to_colour(C.byref(scene), index, o)
    return tuple(o)


def background(scene, view):
    o = (C.c_double * 3)()
    lib().oracle_background(C.byref(scene), (C.c_double * 3)(*view), o)
    return np.array(o[:])


def object_shader(scene, point, view, dist_thresh, depth, last_dist=0.0, obj=0):
    o = ShadeOut()
    lib().oracle_object_shader(C.byref(scene), (C.c_double * 3)(*point), (C.c_double * 3)(*view), dist_thresh,
                               depth, last_dist, obj, C.byref(o))
    return o


_libs = {}


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        _libs[name] = _load(name)
    return _libs[name]


def compute(scene, p, mode=0, n_iter=None):
    it = C.c_int(0)
    n = scene.N if n_iter is None else n_iter
    v = lib().oracle_compute(C.byref(scene), mode, (C.c_double * 3)(*p), n, C.byref(it))
    return v, it.value


def calculate_distance(scene, p, detail=0.0, n_iter=None, min_n=None, normal_mode=0):
    mi, it, ob = C.c_int(0), C.c_int(0), C.c_int(0)
    n = scene.N if n_iter is None else n_iter
    mn = scene.minN if min_n is None else min_n
    d = lib().oracle_calculate_distance(C.byref(scene), (C.c_double * 3)(*p), detail, n, mn, normal_mode,
                                        C.byref(mi), C.byref(it), C.byref(ob))
    return d, bool(mi.value), it.value, ob.value


def camera(scene):
    R, st, li = (C.c_double * 9)(), (C.c_double * 3)(), (C.c_double * 3)()
    lib().oracle_camera(C.byref(scene), R, st, li)
    return np.array(R[:]).reshape(3, 3), np.array(st[:]), np.array(li[:])


def view_vector(scene, x, y):
    d = (C.c_double * 3)()
    cut = lib().oracle_view_vector(C.byref(scene), x, y, d)
    return np.array(d[:]), bool(cut)


def ao_table(quality, lightmap=None):
    dirs = (C.c_double * 30000)()
    rgb = (C.c_int * 30000)()
    tex = None
    if lightmap is not None:
        arr = np.ascontiguousarray(lightmap, dtype=np.uint8)
        tex = _abi.Texture(arr.shape[1], arr.shape[0], arr.ctypes.data_as(C.POINTER(C.c_uint8)))
    n = lib().oracle_ao_table(quality, C.byref(tex) if tex is not None else None, dirs, rgb, 10000)
    return np.array(dirs[:3 * n]).reshape(n, 3), np.array(rgb[:3 * n]).reshape(n, 3)


def march(scene, start, direction, pixel=0):
    out = MarchOut()
    rc = lib().oracle_march(C.byref(scene), (C.c_double * 3)(*start), (C.c_double * 3)(*direction),
                            pixel, C.byref(out))
    assert rc == 0
    return out


class HostFrame:
    """Host planes in cImage layout (structure of arrays)."""

    def __init__(self, w, h):
        self.w, self.h = w, h
        self.rgb = np.zeros((3, h, w), np.float32)
        self.z = np.zeros((h, w), np.float32)
        self.colour = np.zeros((3, h, w), np.uint8)
        self.alpha = np.zeros((h, w), np.uint16)
        self.opacity = np.zeros((h, w), np.uint16)

    def fb(self):
        f = _abi.Framebuffer()
        f.width, f.height = self.w, self.h
        for c in range(3):
            f.rgb[c] = self.rgb[c].ctypes.data
            f.colour[c] = self.colour[c].ctypes.data
        f.z = self.z.ctypes.data
        f.alpha = self.alpha.ctypes.data
        f.opacity = self.opacity.ctypes.data
        return f


def _texture(arr):
    a = np.ascontiguousarray(arr, dtype=np.uint8)
    return a, _abi.Texture(a.shape[1], a.shape[0], a.ctypes.data_as(C.POINTER(C.c_uint8)))


def env_mapping(normal, view, envmap):
    """EnvMapping's texture colour for a unit normal and view vector."""
    a, t = _texture(envmap)
    out = (C.c_double * 3)()
    lib().oracle_set_textures(None, C.byref(t))
    try:
        lib().oracle_env_mapping((C.c_double * 3)(*normal), (C.c_double * 3)(*view), out)
    finally:
        lib().oracle_set_textures(None, None)
    return np.array(out[:])


def render(scene, rows=None, threads=0, jitter=None, glibc_seed=1234, fast=False, lightmap=None, background=None,
           envmap=None):
    """Render with the oracle; returns (HostFrame, OracleStats).  background,
    envmap: (h, w, 3) uint8 textures of the textured background and
    environment mapping."""
    w, h = scene.image_width, scene.image_height
    frame = HostFrame(w, h)
    fb = frame.fb()
    st = OracleStats()
    if jitter is None:
        jitter = JITTER_HASH if scene.jitter_mode == _abi.JITTER_HASH else JITTER_NONE
    r0, r1 = (0, h) if rows is None else rows
    tex = None
    if lightmap is not None:
        arr = np.ascontiguousarray(lightmap, dtype=np.uint8)
        tex = _abi.Texture(arr.shape[1], arr.shape[0], arr.ctypes.data_as(C.POINTER(C.c_uint8)))
    texs = [_texture(t) if t is not None else None for t in (background, envmap)]
    lib(fast).oracle_set_textures(*[C.byref(t[1]) if t is not None else None for t in texs])
    try:
        rc = lib(fast).oracle_render(C.byref(scene), C.byref(tex) if tex is not None else None, r0, r1, threads,
                                     jitter, glibc_seed, C.byref(fb), C.byref(st))
    finally:
        lib(fast).oracle_set_textures(None, None)
    if rc != 0:
        raise RuntimeError(f"oracle_render failed: {rc}")
    return frame, st


def ssao(scene, z, opacity, colour, image16):
    """PostRendering_SSAO of the oracle on numpy planes: z (h, w) float32,
    opacity (h, w) uint16, colour (3, h, w) uint8, image16 (h, w, 3) uint16.
    Returns the new image16 (the input is not changed)."""
    h, w = z.shape
    z = np.ascontiguousarray(z, dtype=np.float32)
    op = np.ascontiguousarray(opacity, dtype=np.uint16)
    col = [np.ascontiguousarray(colour[i], dtype=np.uint8) for i in range(3)]
    out = np.array(image16, dtype=np.uint16, order="C", copy=True)
    cp = (C.c_void_p * 3)(*[c.ctypes.data for c in col])
    rc = lib().oracle_ssao(C.byref(scene), w, h, z.ctypes.data, op.ctypes.data, cp, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle_ssao failed: {rc}")
    return out


def dof_order(scene, z, seed):
    """Pixels in the DOF processing order (sorted z, randomised with glibc
    rand after srand(seed))."""
    h, w = z.shape
    z = np.ascontiguousarray(z, dtype=np.float32)
    order = np.empty(h * w, np.int32)
    rc = lib().oracle_dof_order(C.byref(scene), w, h, z.ctypes.data, seed, order.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle_dof_order failed: {rc}")
    return order


def dof(scene, z, image16, alpha, seed):
    """PostRendering_DOF of the oracle: returns new (image16, alpha)."""
    h, w = z.shape
    z = np.ascontiguousarray(z, dtype=np.float32)
    img = np.array(image16, dtype=np.uint16, order="C", copy=True)
    a = None if alpha is None else np.array(alpha, dtype=np.uint16, order="C", copy=True)
    rc = lib().oracle_dof(C.byref(scene), w, h, z.ctypes.data, seed, img.ctypes.data,
                          None if a is None else a.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle_dof failed: {rc}")
    return img, a


def place_lights(scene):
    """PlaceRandomLights of the oracle: aux_light_number + 4 light records
    (list of (position, intensity, colour, enabled))."""
    n = max(scene.aux_light_number, 0) + 4
    tab = (_abi.Light * n)()
    rc = lib().oracle_place_lights(C.byref(scene), tab, n)
    if rc != n:
        raise RuntimeError(f"oracle_place_lights failed: {rc}")
    return [(tuple(l.position), l.intensity, tuple(l.colour), l.enabled) for l in tab]

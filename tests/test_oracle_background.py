"""The oracle's textured background (BackgroundShader, src/shaders.cpp:
118-157; cTexture::Pixel, src/texture.cpp:83-111) pinned by a texture whose
red channel is the texel column and green the texel row: the bilinear
sample then returns floor(texX) and floor(texY) (one less where the
weighted sum rounds just below an integer), so the view-vector-to-texel maps
of the equirectangular and dome layouts can be checked against their
closed forms, and directions whose coordinates fall outside the texture
read black."""
import ctypes as C
import math

import numpy as np
import pytest

from mandelbulber_amd import _abi, scene
from oracle import oracle

W, H = 200, 100


def _index_texture():
    y, x = np.mgrid[0:H, 0:W]
    t = np.stack([x, y, np.full_like(x, 77)], axis=-1).astype(np.uint8)
    return np.ascontiguousarray(t)


def _sample(s, tex, view):
    t = _abi.Texture(W, H, tex.ctypes.data_as(C.POINTER(C.c_uint8)))
    oracle.lib().oracle_set_textures(C.byref(t), None)
    try:
        return oracle.background(s, view) * 256.0
    finally:
        oracle.lib().oracle_set_textures(None, None)


def _dirs(n=400, seed=5):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _check(rgb, tx, ty):
    if 0 <= tx < W and 0 <= ty < H - 1:
        if int(tx) < W - 1:   # the last column blends with the next row's first texel
            assert int(math.floor(tx)) - 1 <= rgb[0] <= int(math.floor(tx)), (rgb, tx)
            assert int(math.floor(ty)) - 1 <= rgb[1] <= int(math.floor(ty)), (rgb, ty)
        assert rgb[2] in (76, 77) or int(tx) == W - 1   # 77 x (weights summing to 1), truncated
    else:
        assert (rgb == 0).all(), (rgb, tx, ty)


def test_equirectangular_map():
    s = scene.mandelbulb_plain(8, 8)
    s.textured_background = 1
    tex = _index_texture()
    for v in _dirs():
        a = math.fmod(math.atan2(v[1], v[0]) + 2.5 * math.pi, 2 * math.pi)
        b = math.atan2(v[2], math.hypot(v[0], v[1]))
        _check(_sample(s, tex, v), a / (2 * math.pi) * W, (b / math.pi + 0.5) * H)


def test_dome_map():
    s = scene.mandelbulb_plain(8, 8)
    s.textured_background = 1
    s.background_as_fulldome = 1
    tex = _index_texture()
    tw = W // 2
    for v in _dirs(seed=6):
        a = math.atan2(v[1], v[0])
        b = math.atan2(v[2], math.hypot(v[0], v[1]))
        off = 0
        if b < 0:   # the lower hemisphere is the right half
            b, a, off = -b, a + math.pi, tw
        rad = 1.0 - b / (0.5 * math.pi)
        _check(_sample(s, tex, v), 0.5 * tw + math.cos(a) * rad * tw * 0.5 + off, 0.5 * H + math.sin(a) * rad * H * 0.5)


def test_without_a_texture_the_gradient():
    s = scene.mandelbulb_plain(8, 8)
    g = oracle.background(s, (0.0, 0.6, 0.8))
    s.textured_background = 1
    assert np.array_equal(oracle.background(s, (0.0, 0.6, 0.8)), g)   # no texture set: the gradient


def test_environment_map_of_the_mirrored_view():
    """EnvMapping (src/shaders.cpp:785-812): with the index texture the
    sample gives the texel of the mirrored view r = v - 2 (v.n) n at
    x = (atan2(r_y, r_x) + pi) / 2pi W (+ 8.25 W, wrapped), y = (beta / pi +
    0.5) H (+ 8 H, wrapped)."""
    tex = _index_texture()
    rng = np.random.default_rng(9)
    for _ in range(300):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        r = v - 2.0 * np.dot(v, n) * n
        a = math.atan2(r[1], r[0]) + math.pi
        b = math.atan2(r[2], math.hypot(r[0], r[1]))
        tx = math.fmod(a / (2 * math.pi) * W + W * 8.25, W)
        ty = math.fmod((b / math.pi + 0.5) * H + H * 8.0, H)
        rgb = oracle.env_mapping(tuple(n), tuple(v), tex) * 256.0
        _check(rgb, tx, ty)

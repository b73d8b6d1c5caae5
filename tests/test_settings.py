"""The settings model: share/defaults values, LoadOneSetting units and
conversions, -o overrides, and the C ABI's scene layout."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from mandelbulber_amd import _abi, lib, scene as S

REF_SHARE = "/root/reference/share"


def test_scene_layout_matches_ctypes_mirror():
    offs = (C.c_size_t * 8)()
    size = lib().mbr_scene_layout(offs)
    assert (size, list(offs)) == _abi.scene_layout_py()


def test_defaults_values():
    s = S.defaults()
    assert s.abi_version == _abi.MBR_ABI_VERSION
    assert (s.image_width, s.image_height, s.tiles) == (800, 600, 1)
    assert s.formula == 2 and s.power == 9 and s.N == 250 and s.minN == 1
    assert s.alpha == pytest.approx(-20 * math.pi / 180) and s.beta == pytest.approx(30 * math.pi / 180)
    assert s.view_distance_max == 50 and s.view_distance_min == 1e-15
    assert s.bailout == 100.0 and s.jitter_mode == _abi.JITTER_HASH
    assert (s.glow_color1.R, s.glow_color1.G, s.glow_color1.B) == (40984, 44713, 49490)
    assert s.mb_scale == 2 and s.mb_min_radius == 0.5 and s.mb_fixed_radius == 1
    assert s.penetrating_lights == 1 and s.shadow == 1 and s.global_illumination == 0


@pytest.mark.skipif(not os.path.isdir(REF_SHARE), reason="reference tree absent")
def test_defaults_equal_reference_defaults_file(tmp_path):
    """Loading share/defaults itself reproduces mbr_scene_defaults."""
    a = S.defaults()
    b = S.load(os.path.join(REF_SHARE, "defaults"))
    for name, _ in _abi.Scene._fields_:
        va, vb = getattr(a, name), getattr(b, name)
        if hasattr(va, "_length_") or hasattr(va, "_fields_"):
            assert bytes(va) == bytes(vb), name
        else:
            assert va == vb, name


def test_keys_units_and_palette(tmp_path):
    f = tmp_path / "s.fract"
    pal = " ".join("%06x" % (i * 0x010203 & 0xFFFFFF) for i in range(256))
    f.write_text("Mandelbulber 1.21;\nangle_alfa 90;\nmandelbox_rotation_main_beta 45;\n"
                 "main_light_alfa 0.5;\nformula 8;\nN 17;\npalette %s;\nfish_eye 2;\n" % pal)
    s = S.load(str(f), overrides=["N=33", "zoom=1e-6"])
    assert s.alpha == pytest.approx(math.pi / 2)
    assert s.mb_rotation_main[1] == pytest.approx(math.pi / 4)
    assert s.main_light_alpha == 0.5  # radians in the file
    assert s.formula == 8 and s.N == 33 and s.zoom == 1e-6
    assert s.perspective_type == 2
    for i in (0, 1, 100, 255):
        v = i * 0x010203 & 0xFFFFFF
        assert (s.palette[i].R, s.palette[i].G, s.palette[i].B) == (v >> 16, (v >> 8) & 255, v & 255)


def test_unknown_key_rejected_by_set():
    s = S.defaults()
    with pytest.raises(Exception):
        S.set(s, "no_such_key", 1)


@pytest.mark.skipif(not os.path.isdir(REF_SHARE), reason="reference tree absent")
def test_reference_examples_load(capfd):
    """Every example scene of the reference loads, and every key in them is
    known to the loader (no "Unknown parameter" warnings)."""
    ex = os.path.join(REF_SHARE, "examples")
    n = 0
    for name in sorted(os.listdir(ex)):
        if name.endswith(".fract"):
            S.load(os.path.join(ex, name))
            n += 1
    assert n >= 70
    err = capfd.readouterr().err
    assert "Unknown parameter" not in err, err[:2000]


def test_jitter_hash_library_equals_oracle():
    from oracle import oracle as O
    rng = np.random.default_rng(3)
    for a, b, c, d in rng.integers(0, 2 ** 32, size=(500, 4), dtype=np.uint64):
        a, b, c, d = int(a), int(b), int(c), int(d)
        v = lib().mbr_jitter_draw(a, b, c, d)
        assert v == O.lib().oracle_jitter_draw(a, b, c, d)
        assert 0 <= v <= 1000

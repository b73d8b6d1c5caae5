"""Auxiliary lights on the GPU against the oracle: predefined and random
lights with their shadows (LightShading / AuxShadow, src/shaders.cpp:
997-1132) on the shaded Mandelbulb, their glow along the rays
(VolumetricShader, :241-300), volumetric lights (:314-341) and iteration
fog (:385-465), in FP64 and FP32_FIXUP; the context's light table
(placement, explicit tables, animation semantics) through the C ABI."""
import numpy as np
import pytest

from mandelbulber_amd import _abi, scene
from metrics import assert_parity, compare, assert_planes_exact

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    from mandelbulber_amd import Renderer
    r = Renderer(0)
    yield r
    r.close()


def _scene(kind):
    s = scene.mandelbulb_full(160, 120, seed=1)
    s.ao_quality = 2
    s.aux_light_intensity = 0.2
    if kind == "predefined":
        for i in range(4):
            s.aux_light_pre_enabled[i] = 1
        s.aux_light_visibility = 10.0
    elif kind == "random":
        s.aux_light_number = 12
        s.aux_light_visibility = 0.0
    elif kind == "random_glow":
        s.aux_light_number = 12
        s.aux_light_pre_enabled[1] = 1
        s.aux_light_visibility = 20.0
        s.penetrating_lights = 1
    return s


@pytest.mark.parametrize("kind", ["predefined", "random", "random_glow"])
def test_aux_lights_match_oracle(renderer, kind):
    from oracle import oracle
    sc = _scene(kind)
    ref, _ = oracle.render(sc)
    renderer.set_scene(sc)
    for prec, name in [(_abi.PREC_FP64, "fp64"), (_abi.PREC_FP32_FIXUP, "fp32")]:
        print(f"[aux lights] {kind} {name}", flush=True)
        f, st = renderer.render(precision=prec)
        m = compare(f, ref)
        print(f"  {m}", flush=True)
        assert m["hits"] > 2000
        assert_parity(m, label=f"{kind} {name}")
        if prec == _abi.PREC_FP64:
            # the FP64 frame: the same hits and depths as the oracle's
            assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
            assert_planes_exact(m)
    # the lights change the picture
    dark = scene.copy(sc)
    dark.aux_light_number = 0
    for i in range(4):
        dark.aux_light_pre_enabled[i] = 0
    renderer.set_scene(dark)
    f0, _ = renderer.render(precision=_abi.PREC_FP64)
    assert np.abs(f0.rgb - f.rgb).max() > 1e-2


def _vol_scene(kind):
    s = scene.mandelbulb_full(96, 72, seed=1)
    s.ao_quality = 2
    s.N = 40
    s.view_distance_max = 10.0
    s.fog_density = 0.0
    s.aux_light_intensity = 0.2
    for i in range(3):
        s.aux_light_pre_enabled[i] = 1
    if kind == "main":
        s.volumetric_light_enabled[0] = 1
        s.volumetric_light_intensity[0] = 0.05
    else:
        for i in (1, 2, 3):
            s.volumetric_light_enabled[i] = 1
            s.volumetric_light_intensity[i] = 0.5
        s.penetrating_lights = 1
    return s


@pytest.mark.parametrize("kind", ["main", "aux"])
def test_volumetric_lights_match_oracle(renderer, kind):
    from oracle import oracle
    sc = _vol_scene(kind)
    ref, _ = oracle.render(sc)
    renderer.set_scene(sc)
    for prec, name in [(_abi.PREC_FP64, "fp64"), (_abi.PREC_FP32_FIXUP, "fp32")]:
        print(f"[volumetric lights] {kind} {name}", flush=True)
        f, st = renderer.render(precision=prec)
        m = compare(f, ref)
        print(f"  {m}", flush=True)
        assert m["hits"] > 500
        assert_parity(m, label=f"volumetric {kind} {name}")
        if prec == _abi.PREC_FP64:
            assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
            assert_planes_exact(m)


def _fog_scene(kind):
    if kind == "box":
        s = scene.mandelbox_view(96, 72, shaded=True)
        s.N = 18
        s.iter_fog_opacity = 200.0
        s.iter_fog_opacity_trim = 10.0
        s.global_illumination = 0
    else:
        s = scene.mandelbulb_full(96 if kind == "bulb" else 32, 72 if kind == "bulb" else 24, seed=1)
        s.N = 20
        s.view_distance_max = 8.0
        s.iter_fog_opacity = 300.0
        s.iter_fog_opacity_trim = 0.0
        s.global_illumination = 1 if kind == "bulb_ao" else 0
        s.ao_quality = 2
        s.aux_light_pre_enabled[1] = 1
        s.aux_light_intensity = 0.2
    s.fog_density = 0.0
    s.iter_fog_enabled = 1
    return s


@pytest.mark.parametrize("kind", ["bulb", "bulb_ao", "box"])
def test_iteration_fog_matches_oracle(renderer, kind):
    from oracle import oracle
    sc = _fog_scene(kind)
    ref, _ = oracle.render(sc)
    renderer.set_scene(sc)
    for prec, name in [(_abi.PREC_FP64, "fp64"), (_abi.PREC_FP32_FIXUP, "fp32")]:
        print(f"[iteration fog] {kind} {name}", flush=True)
        f, st = renderer.render(precision=prec)
        m = compare(f, ref)
        print(f"  {m}", flush=True)
        assert m["hits"] > 100
        assert_parity(m, label=f"iteration fog {kind} {name}")
        if prec == _abi.PREC_FP64:
            assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
            assert_planes_exact(m)
    # the fog shows
    clear = scene.copy(sc)
    clear.iter_fog_opacity = 0.0
    renderer.set_scene(clear)
    f0, _ = renderer.render(precision=_abi.PREC_FP64)
    assert np.abs(f0.rgb - f.rgb).max() > 1e-3


def _tuples(tab):
    return [(tuple(l.position), l.intensity, tuple(l.colour), l.enabled) for l in tab]


def test_light_table_semantics(renderer):
    sc = _scene("random")
    renderer.set_lights(None)
    renderer.set_scene(sc)
    placed = _tuples(renderer.lights())
    assert placed == _tuples(scene.place_lights(sc))
    f1, _ = renderer.render(precision=_abi.PREC_FP64)
    # a fractal parameter changes (an animation frame): the random lights stay
    moved = scene.copy(sc)
    moved.power = 7.5
    renderer.set_scene(moved)
    assert _tuples(renderer.lights()) == placed
    # a new seed places them again
    reseeded = scene.copy(sc)
    reseeded.aux_light_random_seed = 77
    renderer.set_scene(reseeded)
    assert _tuples(renderer.lights()) != placed
    # an explicit table equal to the placement renders the same frame
    renderer.set_lights(scene.place_lights(sc))
    renderer.set_scene(sc)
    f2, _ = renderer.render(precision=_abi.PREC_FP64)
    assert np.array_equal(f1.rgb, f2.rgb) and np.array_equal(f1.z, f2.z)
    # a table shorter than the scene's light count is an error
    renderer.set_lights(scene.place_lights(sc)[:5])
    with pytest.raises(RuntimeError):
        renderer.set_scene(sc)
    renderer.set_lights(None)


def test_batch_rejects_aux_lights(renderer):
    import torch
    sc = _scene("predefined")
    dev = torch.device("cuda:0")
    w, h = sc.image_width, sc.image_height
    planes = {
        "rgb": torch.zeros((3, h, w), dtype=torch.float32, device=dev),
        "z": torch.zeros((h, w), dtype=torch.float32, device=dev),
        "colour": torch.zeros((3, h, w), dtype=torch.uint8, device=dev),
        "alpha": torch.zeros((h, w), dtype=torch.int16, device=dev),
        "opacity": torch.zeros((h, w), dtype=torch.int16, device=dev),
    }
    with pytest.raises(RuntimeError, match="aux lights"):
        renderer.render_batch([sc], [planes])


@pytest.mark.parametrize("kind", ["surface", "glow", "box"])
def test_fake_lights_match_oracle(renderer, kind):
    """Fake (orbit-trap) lights: FakeLights on the surface (src/shaders.cpp:
    1225-1270) and their glow in VolumetricShader (:302-312), FP64 against
    the oracle bit for bit; the FP32 modes render such scenes in FP64."""
    from oracle import oracle
    if kind == "box":
        s = scene.mandelbox_view(96, 72, shaded=True)
        s.N = 30
        s.ao_quality = 2
    else:
        s = scene.mandelbulb_full(96, 72, seed=1) if kind == "glow" else scene.mandelbulb_plain(96, 72)
        s.N = 30
        s.ao_quality = 2
    s.fake_lights_enabled = 1
    for i, v in enumerate((0.0, 0.3, 0.2)):
        s.fake_lights_orbit_trap[i] = v
    s.fake_lights_min_iter, s.fake_lights_max_iter = 0, 3
    s.fake_lights_intensity = 100.0
    s.fake_lights_visibility = 20.0 if kind != "surface" else 0.0
    s.fake_lights_visibility_size = 10.0
    ref, _ = oracle.render(s)
    renderer.set_scene(s)
    print(f"[fake lights] {kind} fp64", flush=True)
    f64, _ = renderer.render(precision=_abi.PREC_FP64)
    m = compare(f64, ref)
    print(f"  {m}", flush=True)
    assert m["hits"] > 500
    assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
    assert_parity(m, label=f"fake lights {kind}")
    assert_planes_exact(m)
    f32, _ = renderer.render(precision=_abi.PREC_FP32_FIXUP)
    assert np.array_equal(f32.rgb, f64.rgb) and np.array_equal(f32.z, f64.z)
    # the fake lights show
    off = scene.copy(s)
    off.fake_lights_enabled = 0
    renderer.set_scene(off)
    f0, _ = renderer.render(precision=_abi.PREC_FP64)
    assert np.abs(f0.rgb - f64.rgb).max() > 1e-3

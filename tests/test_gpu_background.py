"""The textured background (BackgroundShader's texture branches,
src/shaders.cpp:118-157, with cTexture's bilinear Pixel, src/texture.cpp:
83-111) on the GPU against the oracle, equirectangular and fisheye-dome
layouts, on a synthetic 8-bit texture.  Depths and hits are the FP64 path's
(bit-exact); a background texel index comes from atan2, sqrt, fmod, sin and
cos of the view vector, whose last bits may differ between the device libm
and glibc, so a pixel whose index lands on a texel edge may read its
neighbour: at most 0.1% of the background pixels may differ, by one
interpolation step (measured: none).  The FP32 modes sample the same texels
and run their volumetric pass in FP32."""
import numpy as np
import pytest

from mandelbulber_amd import _abi, scene
from metrics import compare

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


def _texture(w=96, h=48, seed=3):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    t = np.stack([x * 255 // (w - 1), y * 255 // (h - 1), (x * y) % 256], axis=-1).astype(np.int32)
    t = (t + rng.integers(0, 40, size=t.shape)) % 256
    return t.astype(np.uint8)


@pytest.mark.parametrize("dome", [0, 1])
def test_textured_background_matches_oracle(renderer, dome):
    from oracle import oracle
    s = scene.mandelbulb_plain(128, 96)
    s.textured_background = 1
    s.background_as_fulldome = dome
    tex = _texture()
    ref, _ = oracle.render(s, background=tex)
    renderer.set_textures(background=tex)
    renderer.set_scene(s)
    print(f"[background] textured, dome={dome}", flush=True)
    f, _ = renderer.render(precision=_abi.PREC_FP64)
    m = compare(f, ref)
    print(f"  {m}", flush=True)
    assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
    miss = ref.z >= 1e19
    assert miss.sum() > 1000
    d = np.abs(f.rgb[:, miss].astype(np.float64) - ref.rgb[:, miss]).max(axis=0)
    bad = int((d > 0).sum())
    print(f"  background pixels differing: {bad} of {int(miss.sum())}, largest {d.max():.3g}", flush=True)
    assert bad <= 1e-3 * miss.sum() and d.max() <= 1.0 / 256 + 1e-7
    # the texture shows: the background is not the gradient
    plain = scene.copy(s)
    plain.textured_background = 0
    renderer.set_scene(plain)
    g, _ = renderer.render(precision=_abi.PREC_FP64)
    assert np.abs(g.rgb[:, miss] - f.rgb[:, miss]).max() > 0.1
    # the FP32 modes read the same texels (their volumetric pass then runs
    # in FP32)
    renderer.set_scene(s)
    f32, _ = renderer.render(precision=_abi.PREC_FP32_FIXUP)
    assert np.abs(f32.rgb[:, miss] - f.rgb[:, miss]).max() <= 1e-5


def test_textured_background_needs_a_texture(renderer):
    from mandelbulber_amd._lib import MbrError
    s = scene.mandelbulb_plain(32, 24)
    s.textured_background = 1
    renderer.set_textures(None, None)
    with pytest.raises(MbrError) as e:
        renderer.set_scene(s)
    assert e.value.rc == -4 and "texture" in str(e.value)


@pytest.mark.parametrize("kind", ["bulb", "box"])
def test_environment_mapping_matches_oracle(renderer, kind):
    """Environment mapping (reflect > 0 without raytraced reflections,
    EnvMapping src/shaders.cpp:785-812, weighted by the AO in ObjectShader
    :73-98), FP64 against the oracle; the FP32 modes render such scenes in
    FP64."""
    from oracle import oracle
    if kind == "box":
        s = scene.mandelbox_view(96, 72, shaded=True)
        s.N = 30
    else:
        s = scene.mandelbulb_full(96, 72, seed=1)
        s.N = 30
    s.ao_quality = 2
    s.global_illumination = 1
    s.reflect = 0.6
    env = _texture(64, 32, seed=11)
    ref, _ = oracle.render(s, envmap=env)
    renderer.set_textures(envmap=env)
    renderer.set_scene(s)
    print(f"[envmap] {kind} fp64", flush=True)
    f, _ = renderer.render(precision=_abi.PREC_FP64)
    m = compare(f, ref)
    print(f"  {m}", flush=True)
    assert m["hits"] > 500
    assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
    # the texel index comes from atan2, sqrt and fmod of the mirrored view:
    # rare hits on a texel edge may read the neighbour (device libm vs glibc)
    hit = ref.z < 1e19
    d = np.abs(f.rgb[:, hit].astype(np.float64) - ref.rgb[:, hit]).max(axis=0)
    print(f"  hit pixels differing: {int((d > 0).sum())} of {int(hit.sum())}, largest {d.max():.3g}", flush=True)
    assert (d > 1e-6).sum() <= 1e-3 * hit.sum(), m
    assert m["psnr_db"] >= 60, m
    f32, _ = renderer.render(precision=_abi.PREC_FP32_FIXUP)
    assert np.array_equal(f32.rgb, f.rgb) and np.array_equal(f32.z, f.z)
    # the environment shows
    off = scene.copy(s)
    off.reflect = 0.0
    renderer.set_scene(off)
    f0, _ = renderer.render(precision=_abi.PREC_FP64)
    assert np.abs(f0.rgb - f.rgb).max() > 1e-2
    renderer.set_textures(None, None)

"""GPU parity: the HIP render core through the C ABI vs the CPU FP64 oracle.

Same scene, same jitter draws (counter hash), both precisions.  Bars (per
config): <= 0.1% hit/miss flips, depth relative error <= 1e-4 on pixels that
hit the same surface, RGB PSNR >= 45 dB.  The FP64 path is additionally held
to a far tighter bar: it is the reference arithmetic on the GPU.
"""
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


@pytest.fixture(scope="module")
def oracle_mod():
    from oracle import oracle
    return oracle


# FP64_DEMOIVRE: FP64 with the trig-free power-8 iteration (Mandelbulb
# scenes; others render as FP64), held to the bars
PRECISIONS = [(_abi.PREC_FP64, "fp64"), (_abi.PREC_FP32_FIXUP, "fp32"), (_abi.PREC_FP32_MARCH, "fp32-march"),
              (_abi.PREC_FP64_DEMOIVRE, "fp64-demoivre")]


def _check(renderer, oracle_mod, sc, label, fp64_exact=True):
    ref, ost = oracle_mod.render(sc)
    renderer.set_scene(sc)
    for prec, name in PRECISIONS:
        print(f"[gpu parity] {label} {name}", flush=True)
        frame, st = renderer.render(precision=prec)
        m = compare(frame, ref)
        print(f"  {m}", flush=True)
        assert st.rays == sc.image_width * sc.image_height
        assert st.hits == ost.hits or prec in (_abi.PREC_FP32_MARCH, _abi.PREC_FP64_DEMOIVRE)
        assert_parity(m, label=f"{label} {name}")
        if prec == _abi.PREC_FP64 and fp64_exact:
            assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12 and m["psnr_db"] >= 120, m
            assert_planes_exact(m)
        if prec == _abi.PREC_FP32_FIXUP:
            # the march is the FP64 march: same hits, depths from the same
            # arithmetic (stored as float)
            assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-7, m
            # the FP64 continuation repeats the distance estimate of the hit
            # step once per hit ray; every other call is the reference's
            assert st.march_de_calls == ost.march_de_calls + ost.hits


@pytest.mark.parametrize("bailout", [100.0, 10.0])
def test_mandelbulb_plain_320x240(renderer, oracle_mod, bailout):
    _check(renderer, oracle_mod, scene.mandelbulb_plain(320, 240, bailout=bailout), f"plain bailout {bailout}")


@pytest.mark.parametrize("seed", [1, 2])
def test_mandelbulb_shadow_ao_fog_small(renderer, oracle_mod, seed):
    sc = scene.mandelbulb_full(192, 108, seed=seed)
    _check(renderer, oracle_mod, sc, f"shadow+AO+fog seed {seed}")


@pytest.mark.parametrize("shaded", [False, True])
def test_mandelbox(renderer, oracle_mod, shaded):
    _check(renderer, oracle_mod, scene.mandelbox_view(160, 120, shaded=shaded), f"mandelbox shaded={shaded}")


@pytest.mark.parametrize("shaded", [False, True])
def test_menger_sponge(renderer, oracle_mod, shaded):
    _check(renderer, oracle_mod, scene.menger_view(160, 120, shaded=shaded), f"menger shaded={shaded}")


def test_framebuffer_planes_consistent(renderer):
    sc = scene.mandelbulb_plain(128, 96)
    renderer.set_scene(sc)
    frame, st = renderer.render(precision=_abi.PREC_FP32_FIXUP)
    hit = frame.z < 1e19
    # colour buffer is written for hits only, alpha is 1 on hits before fog
    assert np.all(frame.colour[:, ~hit] == 0)
    assert np.all(frame.alpha[hit] > 60000)
    assert int(hit.sum()) == st.hits

"""The soft-shadow branch of MainShadow (cone angle > 0, src/shaders.cpp:
503-535) on the GPU against the oracle, in FP64 and FP32_FIXUP, with hard
and penetrating lights."""
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


@pytest.mark.parametrize("penetrating", [0, 1])
def test_soft_shadow_matches_oracle(renderer, penetrating):
    from oracle import oracle
    s = scene.mandelbulb_full(128, 96, seed=3)
    s.ao_quality = 2
    s.shadow_cone_angle = 4.0
    s.penetrating_lights = penetrating
    ref, _ = oracle.render(s)
    renderer.set_scene(s)
    for prec, name in [(_abi.PREC_FP64, "fp64"), (_abi.PREC_FP32_FIXUP, "fp32")]:
        print(f"[soft shadow] penetrating={penetrating} {name}", flush=True)
        f, _ = renderer.render(precision=prec)
        m = compare(f, ref)
        print(f"  {m}", flush=True)
        assert m["hits"] > 1000
        assert_parity(m, label=f"soft shadow {name}")
        if prec == _abi.PREC_FP64:
            assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
            assert_planes_exact(m)

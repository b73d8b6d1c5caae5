"""FP32 frames of 2^18 rays or more render as bands alternating between two
band slots on two HIP streams; every band's pixels depend only on its own
rays, so such a frame must equal, plane for plane and bit for bit, the same
frame rendered as row ranges small enough to run on one stream.  Covers
band edges that fall inside a row, the FP32_FIXUP and FP32_MARCH paths, and
FP64 (which stays on one stream: two FP64 bands at once measured slower)."""
import numpy as np
import pytest

from mandelbulber_amd import _abi, scene

pytestmark = pytest.mark.gpu

SPLIT_RAYS = 1 << 18


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
    if kind == "bulb_full":
        return scene.mandelbulb_full(640, 480, seed=1)
    if kind == "bulb_ragged":
        # 641 x 409 = 262169 rays: two bands of 131085, the edge inside row 204
        s = scene.mandelbulb_full(641, 409, seed=2)
        s.ao_quality = 3
        return s
    s = scene.mandelbox_view(640, 480, shaded=True)
    s.ao_quality = 3
    return s


PLANES = ("rgb", "z", "colour", "alpha", "opacity")


@pytest.mark.parametrize("kind", ["bulb_full", "bulb_ragged", "box"])
def test_two_stream_frame_equals_row_ranges(renderer, kind):
    s = _scene(kind)
    w, h = s.image_width, s.image_height
    assert w * h >= SPLIT_RAYS
    renderer.set_scene(s)
    cut = h // 2
    assert cut * w < SPLIT_RAYS and (h - cut) * w < SPLIT_RAYS
    for prec, name in ((_abi.PREC_FP64, "fp64"), (_abi.PREC_FP32_FIXUP, "fp32"), (_abi.PREC_FP32_MARCH, "fp32-march")):
        print(f"[split] {kind} {w}x{h} {name}", flush=True)
        full, st = renderer.render(precision=prec)
        top, st_top = renderer.render(precision=prec, rows=(0, cut))
        bottom, st_bottom = renderer.render(precision=prec, rows=(cut, h))
        assert st.hits > 1000
        assert st.rays == st_top.rays + st_bottom.rays == w * h
        assert st.hits == st_top.hits + st_bottom.hits
        for p in PLANES:
            a = getattr(full, p)
            parts = (getattr(top, p)[..., :cut, :], getattr(bottom, p)[..., cut:, :])
            b = np.concatenate(parts, axis=-2)
            assert np.array_equal(a, b, equal_nan=True), f"{kind} {name}: plane {p} differs"
        # the bands overlap in time: the frame span is below the summed parts
        assert st.ms_total > 0
        print(f"  frame {st.ms_total:.2f} ms; summed parts {st.ms_march + st.ms_shade + st.ms_finish:.2f} ms",
              flush=True)

"""mbr_cancel: the reference stops a render when programClosed is set,
testing it before every pixel (src/Render3D.cpp:369-378).  Here a thread
calls mbr_cancel while another renders; the march and AO kernels see the
flag while handing out work, the render returns MBR_ERR_CANCELLED well
before the frame would have finished, and the context renders the next
frame normally (bit for bit the frame of an uncancelled render)."""
import threading
import time

import numpy as np
import pytest

from mandelbulber_amd import _abi, scene
from mandelbulber_amd._lib import MbrError

pytestmark = pytest.mark.gpu

CANCELLED = -6
PLANES = ("rgb", "z", "colour", "alpha", "opacity")


@pytest.fixture(scope="module")
def renderer():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    from mandelbulber_amd import Renderer
    r = Renderer(0)
    yield r
    r.close()


def _same(a, b):
    return all(np.array_equal(getattr(a, p), getattr(b, p), equal_nan=True) for p in PLANES)


@pytest.mark.parametrize("prec", [_abi.PREC_FP32_FIXUP, _abi.PREC_FP64])
def test_cancel_before_render_stops_only_the_next_render(renderer, prec):
    renderer.set_scene(scene.mandelbulb_full(160, 120, seed=1))
    ref, _ = renderer.render(precision=prec)
    renderer.cancel()
    with pytest.raises(MbrError) as e:
        renderer.render(precision=prec)
    assert e.value.rc == CANCELLED
    again, _ = renderer.render(precision=prec)
    assert _same(again, ref)


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_FP32_FIXUP])
def test_cancel_stops_a_frame_on_the_gpu(renderer, prec):
    # 1920x1080 power-8 bulb with AO: ~110 ms in FP64, ~13 ms in FP32
    renderer.set_scene(scene.mandelbulb_full(1920, 1080, seed=1))
    renderer.render(precision=prec)  # buffers made, kernels loaded
    t0 = time.perf_counter()
    full, _ = renderer.render(precision=prec)
    t_full = time.perf_counter() - t0
    out = {}

    def run():
        t = time.perf_counter()
        try:
            renderer.render(precision=prec)
            out["rc"] = 0
        except MbrError as e:
            out["rc"] = e.rc
        out["t"] = time.perf_counter() - t

    th = threading.Thread(target=run)
    th.start()
    time.sleep(0.2 * t_full)
    renderer.cancel()
    th.join(timeout=120)
    assert not th.is_alive()
    print(f"[cancel] prec {prec}: full frame {t_full * 1e3:.1f} ms, cancelled after {0.2 * t_full * 1e3:.1f} ms: "
          f"rc {out.get('rc')} in {out.get('t', 0) * 1e3:.1f} ms", flush=True)
    if out["rc"] == 0:
        # the frame finished before the flag was seen (only possible for a
        # very short frame): the flag then stops the next render at once
        with pytest.raises(MbrError):
            renderer.render(precision=prec)
    else:
        assert out["rc"] == CANCELLED
        if prec == _abi.PREC_FP64:
            assert out["t"] < 0.6 * t_full, (out, t_full)
    after, _ = renderer.render(precision=prec)
    assert _same(after, full)

"""mbr_render_device_noalloc (the entry point that allocates nothing: its
working buffers live in the library's static device scratch array) against
mbr_render_device, through the C ABI: every plane equal, on two calls in a
row, the second with a smaller frame; and the limits it rejects."""
import ctypes as C

import numpy as np
import pytest

from mandelbulber_amd import _abi, lib, scene
from mandelbulber_amd.render import device_framebuffer

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


def _planes(s, fill):
    import torch
    dev = torch.device("cuda:0")
    w, h = s.image_width, s.image_height
    return {
        "rgb": torch.full((3, h, w), fill, dtype=torch.float32, device=dev),
        "z": torch.full((h, w), fill, dtype=torch.float32, device=dev),
        "colour": torch.full((3, h, w), 7, dtype=torch.uint8, device=dev),
        "alpha": torch.full((h, w), 7, dtype=torch.int16, device=dev),
        "opacity": torch.full((h, w), 7, dtype=torch.int16, device=dev),
    }


def _call(renderer, name, planes, precision):
    s = renderer.scene
    fb = device_framebuffer(planes, s.image_width, s.image_height)
    o = _abi.RenderOpts()
    o.precision = precision
    st = _abi.Stats()
    rc = getattr(lib(), name)(renderer._ctx, C.byref(o), C.byref(fb), C.byref(st))
    return rc, st


def _same(a, b):
    bad = []
    for k in a:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        if not np.array_equal(x, y, equal_nan=True):
            bad.append((k, int(np.count_nonzero(x != y))))
    return bad


@pytest.mark.parametrize("case", [
    # (width, height, precision): a frame of several bands, then a smaller one
    [(640, 480, _abi.PREC_FP32_FIXUP), (160, 120, _abi.PREC_FP32_FIXUP)],
    [(400, 300, _abi.PREC_FP64), (96, 64, _abi.PREC_FP64)],
])
def test_noalloc_equals_render_device(renderer, case):
    import torch
    for w, h, prec in case:
        s = scene.mandelbulb_full(w, h, seed=2)
        renderer.set_scene(s)
        want = _planes(s, -1.0)
        rc, st_want = _call(renderer, "mbr_render_device", want, prec)
        assert rc == 0
        got = _planes(s, -2.0)
        rc, st = _call(renderer, "mbr_render_device_noalloc", got, prec)
        assert rc == 0, lib().mbr_last_error(renderer._ctx)
        torch.cuda.synchronize()
        bad = _same(got, want)
        print(f"[noalloc] {w}x{h} precision {prec}: hits {st.hits} (render_device {st_want.hits}), "
              f"launches {st.march_launches}, mismatched planes {bad}", flush=True)
        assert not bad
        assert st.hits == st_want.hits and st.rays == w * h
        assert st.march_de_calls == st_want.march_de_calls and st.shade_de_calls == st_want.shade_de_calls


def test_noalloc_rejects_what_does_not_fit(renderer):
    # AO quality 20 in FP64: 2546 directions x 8 bytes per ray leaves fewer
    # than 16384 rays per band in the scratch
    s = scene.mandelbulb_full(64, 48, seed=1)
    s.ao_quality = 20
    renderer.set_scene(s)
    rc, _ = _call(renderer, "mbr_render_device_noalloc", _planes(s, 0.0), _abi.PREC_FP64)
    msg = lib().mbr_last_error(renderer._ctx).decode()
    print(f"[noalloc] AO quality 20, FP64: rc {rc}: {msg}", flush=True)
    assert rc == -8 and "scratch" in msg  # MBR_ERR_OUT_OF_MEMORY
    # the same scene renders there in FP32 (half the bytes per ray)
    rc, _ = _call(renderer, "mbr_render_device_noalloc", _planes(s, 0.0), _abi.PREC_FP32_FIXUP)
    assert rc == 0
    # raytraced reflections need a slot per level
    r = scene.mandelbulb_plain(64, 48)
    scene.apply(r, {"raytraced_reflections": 1, "reflections_max": 1, "reflect": 0.3})
    renderer.set_scene(r)
    rc, _ = _call(renderer, "mbr_render_device_noalloc", _planes(r, 0.0), _abi.PREC_FP64)
    assert rc == -4  # MBR_ERR_UNSUPPORTED

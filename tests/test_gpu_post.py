"""Post effects on the GPU against the oracle: screen-space ambient
occlusion (PostRendering_SSAO, src/image.cpp:249-415) of a rendered,
compiled frame, for the three projections and two qualities, through the
host-plane and the device-plane calls of mbr_post_ssao.  The arithmetic is
FP64 in the reference's order on both sides, so the 16-bit results must be
equal; the only allowed difference is one unit on at most 1e-4 of the
values (a 1-ulp difference of the device atan/sin/cos at a truncation
boundary)."""
import numpy as np
import pytest

from mandelbulber_amd import _abi, scene
from mandelbulber_amd.render import compile_image

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


def _check(label, got, want):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    n_diff = int(np.count_nonzero(d))
    print(f"[post] {label}: max diff {int(d.max())}, {n_diff} of {d.size} values differ", flush=True)
    assert d.max() <= 1 and n_diff <= 1e-4 * d.size, label


@pytest.mark.parametrize("ptype,quality", [(0, 20), (0, 7), (1, 12), (2, 12)])
def test_ssao_matches_oracle(renderer, ptype, quality):
    import torch
    from oracle import oracle
    s = scene.mandelbulb_full(160, 120, seed=2)
    s.perspective_type = ptype
    if ptype != 0:
        s.persp = 0.3   # narrower angle: the bulb fills about a quarter of the frame
    s.ssao_quality = quality
    s.global_ilum = 1.3
    print(f"[ssao] perspective {ptype} quality {quality}: rendering", flush=True)
    renderer.set_scene(s)
    frame, _ = renderer.render(precision=_abi.PREC_FP64)
    img = compile_image(s, frame.rgb)
    hits = int(np.count_nonzero(frame.z < 1e19))
    assert hits > 1000
    want = oracle.ssao(s, frame.z, frame.opacity, frame.colour, img)
    assert np.count_nonzero(want != img) > hits // 2   # the effect changed most hit pixels
    got = renderer.post_ssao(s, frame.z, frame.opacity, frame.colour, img.copy())
    _check(f"host planes p{ptype} q{quality}", got, want)
    dev = torch.device("cuda:0")
    z = torch.from_numpy(frame.z).to(dev)
    op = torch.from_numpy(frame.opacity.view(np.int16)).to(dev)
    col = torch.from_numpy(frame.colour).to(dev)
    im = torch.from_numpy(img.view(np.int16)).to(dev)
    renderer.post_ssao(s, z, op, [col[0], col[1], col[2]], im)
    torch.cuda.synchronize()
    _check(f"device planes p{ptype} q{quality}", im.cpu().numpy().view(np.uint16), want)


def test_ssao_rejects_bad_quality(renderer):
    from mandelbulber_amd._lib import MbrError
    s = scene.mandelbulb_full(32, 24, seed=1)
    s.ssao_quality = 0
    z = np.ones((24, 32), np.float32)
    with pytest.raises(MbrError):
        renderer.post_ssao(s, z, np.zeros((24, 32), np.uint16), np.zeros((3, 24, 32), np.uint8),
                           np.zeros((24, 32, 3), np.uint16))


@pytest.mark.parametrize("radius,focus_quantile", [(60.0, 0.5), (400.0, 0.2)])
def test_dof_matches_oracle(renderer, radius, focus_quantile):
    """Depth of field: the GPU gather (tile lists in splat order) against the
    oracle's sequential splat loop, on a rendered frame focused inside the
    object; the radius-400 case blurs the background by 64 pixels and the
    near side up to the 100-pixel cap."""
    import math
    import torch
    from oracle import oracle
    s = scene.mandelbulb_full(160, 120, seed=2)
    print(f"[dof] radius {radius}: rendering", flush=True)
    renderer.set_scene(s)
    frame, _ = renderer.render(precision=_abi.PREC_FP64)
    hit_z = frame.z[frame.z < 1e19]
    s.dof_enabled = 1
    s.dof_focus = 10.0 * (16.0 + math.log10(float(np.quantile(hit_z, focus_quantile))))
    s.dof_radius = radius
    img = compile_image(s, frame.rgb)
    alpha = frame.alpha.copy()
    want_img, want_a = oracle.dof(s, frame.z, img, alpha, s.jitter_seed)
    assert np.count_nonzero(np.any(want_img != img, axis=2)) > img.shape[0] * img.shape[1] // 4
    got_img, got_a = renderer.post_dof(s, frame.z, img.copy(), alpha.copy())
    _check(f"dof host r{radius} rgb", got_img, want_img)
    _check(f"dof host r{radius} alpha", got_a, want_a)
    dev = torch.device("cuda:0")
    z = torch.from_numpy(frame.z).to(dev)
    im = torch.from_numpy(img.view(np.int16)).to(dev)
    a = torch.from_numpy(alpha.view(np.int16)).to(dev)
    renderer.post_dof(s, z, im, a)
    torch.cuda.synchronize()
    _check(f"dof device r{radius} rgb", im.cpu().numpy().view(np.uint16), want_img)
    _check(f"dof device r{radius} alpha", a.cpu().numpy().view(np.uint16), want_a)


@pytest.mark.parametrize("hdr", [0, 1])
def test_compile_image_device_equals_host(renderer, hdr):
    """CompileImage on the GPU against the host restatement (same table);
    with the HDR curve the device tanhf may differ in its last bit, which can
    move a value to the neighbouring gamma-table entry."""
    import torch
    s = scene.mandelbulb_full(160, 120, seed=3)
    s.brightness, s.contrast, s.image_gamma, s.hdr = 1.4, 1.2, 2.2, hdr
    renderer.set_scene(s)
    frame, _ = renderer.render(precision=_abi.PREC_FP32_FIXUP)
    want = compile_image(s, frame.rgb)
    dev = torch.device("cuda:0")
    rgb = torch.from_numpy(frame.rgb).to(dev)
    out = torch.zeros((120, 160, 3), dtype=torch.int16, device=dev)
    renderer.compile_image_device(s, rgb, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint16)
    n_diff = int(np.count_nonzero(got != want))
    print(f"[post] compile hdr {hdr}: {n_diff} of {got.size} values differ", flush=True)
    assert n_diff == 0 if not hdr else n_diff <= 1e-3 * got.size

"""Differentiable rendering: the render core as a torch.autograd.Function.

forward renders the FP64 image (3, H, W) of a scene whose nine parameters
theta = (power, vp x/y/z, alpha, beta, gamma, light alpha, light beta) come
from a tensor; backward hands the image adjoint dL/dC to the library's
adjoint kernel (mbr_adjoint) and returns dL/dtheta.  The scene must satisfy
mbr_adjoint's conditions (scene.differentiable builds such scenes).
"""
import numpy as np
import torch

from . import _abi, scene as S


class RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, renderer, base_scene):
        s = S.set_params(S.copy(base_scene), theta.detach().cpu().tolist())
        renderer.set_scene(s)
        h, w = s.image_height, s.image_width
        dev = torch.device("cuda:0")
        planes = {"rgb": torch.empty((3, h, w), dtype=torch.float32, device=dev),
                  "z": torch.empty((h, w), dtype=torch.float32, device=dev)}
        renderer.render_device(planes, precision=_abi.PREC_FP64)
        ctx.renderer, ctx.scene = renderer, s
        ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
        return planes["rgb"]

    @staticmethod
    def backward(ctx, grad_rgb):
        r = ctx.renderer
        r.set_scene(ctx.scene)
        grad, _ = r.adjoint(grad_rgb.detach().to(torch.float32).contiguous())
        return torch.tensor(grad, dtype=ctx.theta_dtype, device=ctx.theta_device), None, None


def render(theta, renderer, base_scene):
    """Image (3, H, W) float32 on the GPU, differentiable in theta (9,)."""
    return RenderFn.apply(theta, renderer, base_scene)


def render_planes(renderer, base_scene, theta):
    """FP64 image (3, H, W) as float64 and the hit mask of the scene with
    parameters theta."""
    s = S.set_params(S.copy(base_scene), theta)
    renderer.set_scene(s)
    dev = torch.device("cuda:0")
    p = {"rgb": torch.empty((3, s.image_height, s.image_width), dtype=torch.float32, device=dev),
         "z": torch.empty((s.image_height, s.image_width), dtype=torch.float32, device=dev)}
    renderer.render_device(p, precision=_abi.PREC_FP64)
    return p["rgb"].double(), p["z"] < 1e19


def check_gradient(renderer, base_scene, target_power=7.3, h=1e-4, region="hits"):
    """Adjoint against central differences (step h) of the MSE against the
    image at power target_power (same camera and light).

    The image is piecewise smooth in theta: it steps where a silhouette
    crosses a pixel centre and where an escape iteration count behind a
    pixel changes (normal and depth jump).  A central difference across a
    step measures jump / 2h, so pixels whose hit/miss status differs between
    the 19 renders, or whose one-sided differences disagree, are left out of
    both the loss and the adjoint; region selects hit or miss pixels.
    Returns (adjoint gradient, FD gradient, pixels left out, pixels in the
    region)."""
    theta0 = np.array(S.get_params(base_scene))
    tt = theta0.copy()
    tt[0] = target_power
    T, _ = render_planes(renderer, base_scene, tt)
    C0, hit0 = render_planes(renderer, base_scene, theta0)
    plus, minus = [], []
    mask = torch.ones_like(hit0)
    for k in range(9):
        e = np.zeros(9)
        e[k] = h
        cp, hp = render_planes(renderer, base_scene, theta0 + e)
        cm, hm = render_planes(renderer, base_scene, theta0 - e)
        plus.append(cp)
        minus.append(cm)
        mask &= (hp == hit0) & (hm == hit0)
        # one-sided differences (per unit theta) agree on smooth pixels to
        # O(h); the floor covers the rounding of the image / h (float32 on
        # hits, the 16-bit background on misses)
        floor = 5e-3 if region == "hits" else 3.0 / 65536.0 / h
        dp, dm = (cp - C0) / h, (C0 - cm) / h
        mask &= ((dp - dm).abs() <= 0.05 * torch.maximum(dp.abs(), dm.abs()) + floor).all(dim=0)
    mask &= hit0 if region == "hits" else ~hit0
    # grazing hits the adjoint leaves out are left out of the FD loss too
    skipped = torch.zeros(hit0.shape, dtype=torch.uint8, device=hit0.device)
    renderer.set_scene(S.set_params(S.copy(base_scene), theta0))
    renderer.adjoint(torch.ones((3,) + tuple(hit0.shape), dtype=torch.float32, device=hit0.device), skipped=skipped)
    mask &= skipped == 0
    nel = C0.numel()
    m = mask.double()
    loss = lambda C: float(((C - T) ** 2 * m).sum() / nel)
    fd = np.array([(loss(plus[k]) - loss(minus[k])) / (2 * h) for k in range(9)])
    gimg = (2.0 * (C0 - T) * m / nel).float().contiguous()
    renderer.set_scene(S.set_params(S.copy(base_scene), theta0))
    adj, _ = renderer.adjoint(gimg)
    region_px = int(hit0.sum()) if region == "hits" else int((~hit0).sum())
    return np.array(adj), fd, region_px - int(mask.sum()), region_px

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


class GradientCheck:
    """Result of check_gradient: the adjoint and the finite differences it
    is checked against, per parameter (arrays of 9)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def check_gradient(renderer, base_scene, target_power=7.3, h=1e-4, region="hits", tol=1e-3):
    """Adjoint against central differences (step h) of the MSE against the
    image at power target_power (same camera and light), parameter by
    parameter.

    The image is piecewise smooth in theta: it steps where a silhouette
    crosses a pixel centre and where the march settles on another sheet of
    the surface; between steps, fractal detail below the pixel and the Phong
    term max(0, n.h)^30 make some pixels vary so fast that a central
    difference at h is far from their derivative.  For each parameter k the
    pixels are kept whose hit/miss status is the same in all seven renders
    (theta, theta +- h, +- h/2, +- h/4), whose second differences at h/2 and
    h/4 agree (no step at theta itself), whose central differences at h,
    h/2 and h/4 agree to `tol` relative (above the float32 rounding of the
    image), and which the adjoint does not flag (grazing hits, hits whose
    refinement found no root); the others are left out of both the loss and
    the adjoint for that parameter.  region selects hit or miss pixels (for
    misses, whose background is a 16-bit staircase, the kept pixels are
    those whose one-sided differences at h agree to 5%).

    Returned per parameter (GradientCheck):
      adj     mbr_adjoint with the kept pixels' MSE adjoint 2 (C - T) / n;
      fd      central difference of the kept pixels' MSE at h;
      fd_r    the same with its h^2 term removed (Richardson: 4 fd(h/2) -
              fd(h), over 3); the MSE's third derivative holds sums of
              C' C'' over pixels, so fd(h) differs from the derivative by
              h^2 sum C' C'' / n, which on these images is 1e-4 to 1e-1 of
              the gradient;
      vjp     the central difference of the image at h contracted with the
              MSE adjoint, sum 2 (C - T) / n (C(+h) - C(-h)) / 2h: the
              vector-Jacobian product the adjoint computes, by differences;
      vjp_r   the same from the image's central differences at h and h/2
              with their h^2 term removed (4 d(h/2) - d(h), over 3);
      left    pixels of the region left out; region_px pixels in it."""
    theta0 = np.array(S.get_params(base_scene))
    tt = theta0.copy()
    tt[0] = target_power
    T, _ = render_planes(renderer, base_scene, tt)
    C0, hit0 = render_planes(renderer, base_scene, theta0)
    skipped = torch.zeros(hit0.shape, dtype=torch.uint8, device=hit0.device)
    renderer.set_scene(S.set_params(S.copy(base_scene), theta0))
    renderer.adjoint(torch.ones((3,) + tuple(hit0.shape), dtype=torch.float32, device=hit0.device), skipped=skipped)
    base = (hit0 if region == "hits" else ~hit0) & (skipped == 0)
    nel = C0.numel()
    # float32 rounding of a pixel's value, per channel
    ulp = 6e-8 * (C0.abs() + 1e-6)
    out = {key: np.zeros(9) for key in ("adj", "fd", "fd_r", "vjp", "vjp_r")}
    left = np.zeros(9, dtype=np.int64)
    for k in range(9):
        e = np.zeros(9)
        e[k] = h
        img, mask = {}, base.clone()
        for f in (1.0, -1.0, 0.5, -0.5, 0.25, -0.25):
            c, hk = render_planes(renderer, base_scene, theta0 + f * e)
            img[f] = c
            mask &= hk == hit0
        d1 = (img[1.0] - img[-1.0]) / (2 * h)
        d2 = (img[0.5] - img[-0.5]) / h
        if region == "hits":
            d4 = (img[0.25] - img[-0.25]) / (0.5 * h)
            noise = 3.0 * ulp / (0.25 * h)
            s2 = img[0.5] - 2 * C0 + img[-0.5]
            s4 = (img[0.25] - 2 * C0 + img[-0.25]) * 4.0
            ok = ((d1 - d4).abs() <= tol * d4.abs() + noise) & ((d2 - d4).abs() <= tol * d4.abs() + noise)
            ok &= (s2 - s4).abs() <= 0.25 * (s2.abs() + s4.abs()) + 24.0 * ulp
            mask &= ok.all(dim=0)
        else:
            dp, dm = (img[1.0] - C0) / h, (C0 - img[-1.0]) / h
            mask &= ((dp - dm).abs() <= 0.05 * torch.maximum(dp.abs(), dm.abs()) + 3.0 / 65536.0 / h).all(dim=0)
        m = mask.double()
        loss = lambda C: float(((C - T) ** 2 * m).sum() / nel)
        fd1 = (loss(img[1.0]) - loss(img[-1.0])) / (2 * h)
        fd2 = (loss(img[0.5]) - loss(img[-0.5])) / h
        out["fd"][k] = fd1
        out["fd_r"][k] = (4.0 * fd2 - fd1) / 3.0
        w = 2.0 * (C0 - T) * m / nel
        out["vjp"][k] = float((w * d1).sum())
        out["vjp_r"][k] = float((w * (4.0 * d2 - d1) / 3.0).sum())
        renderer.set_scene(S.set_params(S.copy(base_scene), theta0))
        out["adj"][k] = renderer.adjoint(w.float().contiguous())[0][k]
        left[k] = int(base.sum()) - int(mask.sum())
    region_px = int(hit0.sum()) if region == "hits" else int((~hit0).sum())
    return GradientCheck(left=left + (region_px - int(base.sum())), region_px=region_px, **out)


def compare(adj, ref, tol=1e-3):
    """Relative error per parameter of adj against ref: |adj - ref| / |ref|
    for components above 5% of the largest |ref|, |adj - ref| / max |ref|
    for the others; and whether all are within tol."""
    adj, ref = np.asarray(adj, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(adj - ref)
    big = np.abs(ref) > 0.05 * scale
    rel = np.where(big, err / np.maximum(np.abs(ref), 1e-300), err / scale)
    return rel, bool(np.all(rel <= tol))

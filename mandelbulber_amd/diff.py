"""Differentiable rendering: the render core as a torch.autograd.Function.

forward renders the FP64 image (3, H, W) of a scene whose nine parameters
theta = (power, vp x/y/z, alpha, beta, gamma, light alpha, light beta) come
from a tensor; backward hands the image adjoint dL/dC to the library's
adjoint kernel (mbr_adjoint) and returns dL/dtheta.  The scene must satisfy
mbr_adjoint's conditions (scene.differentiable builds such scenes).
"""
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

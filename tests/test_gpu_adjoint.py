"""Adjoint (mbr_adjoint) against central finite differences of the FP64
forward, and through torch.autograd.

Loss: MSE of the image against a target rendered at power 7.3 with the same
camera and light.  FD: central differences with h = 1e-4 of the same FP64
forward (refined hits, no jitter), with the pixels whose image steps inside
[theta - h, theta + h] left out of both sides (diff.check_gradient).  Bar:
every component within 1e-3 relative of the FD gradient where it is not
small, and within 1e-3 of the largest component everywhere."""
import numpy as np
import pytest

from mandelbulber_amd import _abi, scene

pytestmark = pytest.mark.gpu
H_FD = 1e-4


@pytest.fixture(scope="module")
def renderer():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    from mandelbulber_amd import Renderer
    r = Renderer(0)
    yield r
    r.close()


def fd_check(r, base, h=H_FD, region="hits"):
    from mandelbulber_amd import diff
    return diff.check_gradient(r, base, h=h, region=region)


def _compare(adj, fd, label, names=scene.DIFF_PARAMS):
    scale = np.abs(fd).max()
    err = np.abs(adj - fd)
    print(f"[adjoint] {label}", flush=True)
    for k, name in enumerate(names):
        print(f"  {name:12s} adjoint {adj[k]: .6e}  fd {fd[k]: .6e}  abs/max {err[k] / scale:.2e}  "
              f"rel {err[k] / max(abs(fd[k]), 1e-30):.2e}", flush=True)
    assert scale > 0
    assert np.all(err <= 1e-3 * scale), (adj, fd)
    big = np.abs(fd) > 0.05 * scale
    assert np.all(err[big] <= 1e-3 * np.abs(fd[big])), (adj, fd)


@pytest.mark.parametrize("seed", [1, 2])
def test_adjoint_matches_central_differences(renderer, seed):
    base = scene.differentiable(128, 128, seed=seed)
    adj, fd, masked, region_px = fd_check(renderer, base)
    print(f"[adjoint] seed {seed}: {masked} of {region_px} hit pixels left out", flush=True)
    # the normal is a difference of four distance estimates a fraction of a
    # pixel apart; their escape counts change often along a surface, so a
    # sizeable share of hit pixels steps within +-h: at most 60% here
    assert masked < 0.6 * region_px
    _compare(adj, fd, f"seed {seed} hits")


def test_background_gradient(renderer):
    """Misses only: the background is stored as 16-bit integers, a staircase
    of step 1/65536, and changes colour pair at the horizon (a kink); with
    h = 2e-3 the steps average out over the image, pixels whose one-sided
    differences disagree (the kink within +-h) are left out, and the central
    difference follows the slope the adjoint uses."""
    base = scene.differentiable(128, 128, seed=4)
    adj, fd, masked, region_px = fd_check(renderer, base, h=2e-3, region="misses")
    print(f"[adjoint] background: {masked} of {region_px} miss pixels left out", flush=True)
    assert region_px > 1000
    # only the camera angles move the background
    assert np.all(adj[[0, 1, 2, 3, 7, 8]] == 0) and np.all(np.abs(fd[[0, 1, 2, 3, 7, 8]]) < 1e-9)
    _compare(adj[4:7], fd[4:7], "background", names=scene.DIFF_PARAMS[4:7])


def test_autograd_function_calls_the_adjoint(renderer):
    import torch
    from mandelbulber_amd import diff
    base = scene.differentiable(96, 96, seed=3)
    theta = torch.tensor(scene.get_params(base), dtype=torch.float64, requires_grad=True)
    target_theta = torch.tensor(scene.get_params(base), dtype=torch.float64)
    target_theta[0] = 7.3
    with torch.no_grad():
        target = diff.render(target_theta, renderer, base)
    img = diff.render(theta, renderer, base)
    loss = ((img.double() - target.double()) ** 2).mean()
    loss.backward()
    # the same adjoint called directly
    gimg = (2.0 * (img.detach().double() - target.double()) / img.numel()).float().contiguous()
    renderer.set_scene(scene.set_params(scene.copy(base), theta.detach().tolist()))
    direct, _ = renderer.adjoint(gimg)
    assert theta.grad is not None and theta.grad.shape == (9,)
    assert np.allclose(theta.grad.numpy(), direct, rtol=1e-6, atol=1e-12)
    assert abs(float(theta.grad[0])) > 0

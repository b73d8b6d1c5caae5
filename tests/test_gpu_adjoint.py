"""Adjoint (mbr_adjoint, mbr_adjoint_jacobian) against central finite
differences of the FP64 forward, and through torch.autograd.

Workload: 16 Mandelbulb scenes at 512x512, power uniform in [6, 10], seeded
camera and light (scene.differentiable, seeds 100-115); loss: MSE of the
image against a target rendered at power 7.3 with the same camera and light.
FD: central differences with h = 1e-4 of the same FP64 forward (refined
hits, no jitter), parameter by parameter over the pixels where the image is
smooth within +-h (diff.check_gradient says which).

Bar: every component of the adjoint within 1e-3 relative of the loss's
central difference at h = 1e-4 with its h^2 term removed (Richardson with
h/2), and of the image's central differences contracted with the MSE
adjoint (the same), where the component is above 5% of the largest, and
within 1e-3 of the largest component everywhere.  The plain central
difference of the loss at h = 1e-4 is printed beside them: its h^2 sum of
C' C'' over the pixels is 1e-4 to 5e-2 of the gradient on these images."""
import numpy as np
import pytest

from mandelbulber_amd import _abi, scene

pytestmark = pytest.mark.gpu
H_FD = 1e-4
SCENES = range(100, 116)


@pytest.fixture(scope="module")
def renderer():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    from mandelbulber_amd import Renderer
    r = Renderer(0)
    yield r
    r.close()


def _table(chk, label, names=scene.DIFF_PARAMS, idx=range(9)):
    from mandelbulber_amd import diff
    rel_r, ok_r = diff.compare(chk.adj[list(idx)], chk.fd_r[list(idx)])
    rel_v, ok_v = diff.compare(chk.adj[list(idx)], chk.vjp_r[list(idx)])
    rel_p, _ = diff.compare(chk.adj[list(idx)], chk.fd[list(idx)])
    print(f"[adjoint] {label}: {chk.region_px} pixels in the region", flush=True)
    for j, k in enumerate(idx):
        print(f"  {names[k]:12s} adjoint {chk.adj[k]: .6e}  fd_r {chk.fd_r[k]: .6e} ({rel_r[j]:.1e})  "
              f"vjp_r {chk.vjp_r[k]: .6e} ({rel_v[j]:.1e})  fd(h) {chk.fd[k]: .6e} ({rel_p[j]:.1e})  "
              f"left out {chk.left[k] / max(chk.region_px, 1):.3f}", flush=True)
    return rel_r, rel_v, ok_r and ok_v


def test_adjoint_matches_central_differences_16_scenes_512(renderer):
    """The differentiable workload at its own size: 16 scenes, 512x512."""
    from mandelbulber_amd import diff
    worst, fails, left = 0.0, [], []
    for seed in SCENES:
        base = scene.differentiable(512, 512, seed=seed)
        chk = diff.check_gradient(renderer, base, h=H_FD)
        rel_r, rel_v, ok = _table(chk, f"scene {seed}")
        worst = max(worst, rel_r.max(), rel_v.max())
        left.append(chk.left / chk.region_px)
        if not ok:
            fails.append(seed)
    left = np.array(left)
    print(f"[adjoint] worst relative error {worst:.2e}; hit pixels left out per parameter (mean over scenes) "
          f"{np.round(left.mean(axis=0), 3).tolist()}, mean {left.mean():.3f}", flush=True)
    assert not fails, fails
    # what is left out: pixels whose image steps or varies too fast for a
    # difference at h; camera yaw and pitch move the image across the pixel
    # grid and leave out the most
    assert left.mean() < 0.25
    assert np.all(left.mean(axis=0) < 0.5)


def test_jacobian_matches_pixel_differences(renderer):
    """mbr_adjoint_jacobian pixel by pixel against the image's central
    differences (Richardson from h/2 and h/4) on the pixels where those have
    converged; and mbr_adjoint equals the Jacobian contracted with the
    image adjoint."""
    import torch
    from mandelbulber_amd import diff
    n = 256
    base = scene.differentiable(n, n, seed=101)
    th0 = np.array(scene.get_params(base))
    C0, hit0 = diff.render_planes(renderer, base, th0)
    renderer.set_scene(scene.set_params(scene.copy(base), th0))
    jac = torch.zeros((9, 3, n, n), dtype=torch.float32, device="cuda")
    sk = torch.zeros((n, n), dtype=torch.uint8, device="cuda")
    renderer.adjoint_jacobian(jac, skipped=sk)
    g = torch.randn((3, n, n), dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    g = (g * hit0).float().contiguous()
    direct, _ = renderer.adjoint(g)
    contracted = (jac.double() * g.double()).sum(dim=(1, 2, 3)).cpu().numpy()
    assert np.allclose(direct, contracted, rtol=1e-5, atol=1e-6 * np.abs(contracted).max())
    ulp = 6e-8 * (C0.abs() + 1e-6)
    for k in range(9):
        e = np.zeros(9)
        e[k] = H_FD
        img, m = {}, hit0 & (sk == 0)
        for f in (1.0, -1.0, 0.5, -0.5, 0.25, -0.25):
            c, hk = diff.render_planes(renderer, base, th0 + f * e)
            img[f] = c
            m &= hk == hit0
        d1 = (img[1.0] - img[-1.0]) / (2 * H_FD)
        d2 = (img[0.5] - img[-0.5]) / H_FD
        d4 = (img[0.25] - img[-0.25]) / (0.5 * H_FD)
        noise = 3.0 * ulp / (0.25 * H_FD)
        m &= (((d1 - d4).abs() <= 1e-3 * d4.abs() + noise) & ((d2 - d4).abs() <= 1e-3 * d4.abs() + noise)).all(0)
        ref = (4.0 * d4 - d2) / 3.0
        rel = ((jac[k].double() - ref).abs() / (ref.abs() + noise)).amax(dim=0)[m]
        p50, p99 = float(rel.median()), float(rel.quantile(0.99))
        print(f"[adjoint] pixels {scene.DIFF_PARAMS[k]:12s}: {int(m.sum())} of {int(hit0.sum())} compared, "
              f"relative error p50 {p50:.1e} p99 {p99:.1e} max {float(rel.max()):.1e}", flush=True)
        assert p50 < 1e-3 and p99 < 5e-2


def test_background_gradient(renderer):
    """Misses only: the background is stored as 16-bit integers, a staircase
    of step 1/65536, and changes colour pair at the horizon (a kink); with
    h = 2e-3 the steps average out over the image, pixels whose one-sided
    differences disagree (the kink within +-h) are left out, and the central
    difference follows the slope the adjoint uses."""
    from mandelbulber_amd import diff
    base = scene.differentiable(128, 128, seed=4)
    chk = diff.check_gradient(renderer, base, h=2e-3, region="misses")
    print(f"[adjoint] background: {chk.left.tolist()} of {chk.region_px} miss pixels left out", flush=True)
    assert chk.region_px > 1000
    # only the camera angles move the background
    assert np.all(chk.adj[[0, 1, 2, 3, 7, 8]] == 0) and np.all(np.abs(chk.fd[[0, 1, 2, 3, 7, 8]]) < 1e-9)
    rel, ok = diff.compare(chk.adj[4:7], chk.fd[4:7])
    print(f"  camera angles: adjoint {chk.adj[4:7]} fd {chk.fd[4:7]} rel {rel}", flush=True)
    assert ok


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

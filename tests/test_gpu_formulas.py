"""The generic formulas on the GPU (trig_DE, xenodreambuie, mandelbulb2/3/4:
the F = 0 kernels, FP64 in every precision mode) against the oracle, plain
and with shadow + AO + fog.  The delta-DE formulas difference orbit radii
1e-10 apart, which magnifies the last-bit differences of the device libm
(pow, atan2, acos, sin, cos) against glibc, so FP64 is held to the parity
bars rather than to bit equality (the depth bar of the two formulas with
transcendentals in every iteration on the 99.9th percentile); the other
precision modes render the same FP64 path and must equal the FP64 frame,
except for the analytic generic formulas (trig_DE, kaleidoscopic IFS,
smoothMandelbox, mandelboxVaryScale4D, generalizedFoldBox, and formulas 2, 7
and 8 under the IFS folding pre-fold): their FP32 modes march in FP64 (the
FP64 frame's depths and hits) and shade in FP32, held to the parity bars.  generalizedFoldBox (all eight fold polyhedra),
mandelboxVaryScale4D and the IFS folding pre-fold are analytic and, but for
pow and the Mandelbulb's trig, free of transcendentals: there the FP64 frame
equals the oracle's."""
import numpy as np
import pytest

from mandelbulber_amd import _abi, scene
from metrics import assert_parity, compare, assert_planes_exact

pytestmark = pytest.mark.gpu

# formulas whose FP32 modes shade in FP32 (FP32_MARCH runs as FP32_FIXUP)
FP32_SHADED = (1, 10, 17, 18, 37)

FORMULAS = [(1, 8.0, "trig_DE"), (3, 2.0, "fast_trig"), (4, 2.0, "hypercomplex"), (5, 2.0, "quaternion"),
            (6, 2.0, "minus_fast_trig"), (11, 5.0, "xenodreambuie"), (13, 2.0, "mandelbulb2"),
            (14, 2.0, "mandelbulb3"), (15, 8.0, "mandelbulb4"), (16, 2.0, "foldingIntPow2"),
            (17, 8.0, "smoothMandelbox"), (19, 2.0, "aexion"), (20, 2.0, "benesi"), (21, 2.0, "bristorbrot")]


@pytest.fixture(scope="module")
def renderer():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    from mandelbulber_amd import Renderer
    r = Renderer(0)
    yield r
    r.close()


def _scene(formula, power, shaded):
    # the binary search ends anywhere in its band [0.99 thr, thr], thr ~ depth / width: at 256
    # pixels a different (noisy) delta-DE sequence moves a hit by up to ~4e-5 of its depth
    if formula == 17:
        s = scene.mandelbox_view(192, 144, shaded=shaded)
    else:
        s = scene.mandelbulb_full(192, 144, seed=1) if shaded else scene.mandelbulb_plain(256, 192)
    s.formula = formula
    s.power = power
    s.N = 40
    if shaded:
        s.ao_quality = 2
    return s


@pytest.mark.parametrize("shaded", [False, True])
@pytest.mark.parametrize("formula,power,name", FORMULAS)
def test_generic_formula_matches_oracle(renderer, formula, power, name, shaded):
    from oracle import oracle
    sc = _scene(formula, power, shaded)
    ref, ost = oracle.render(sc)
    renderer.set_scene(sc)
    print(f"[formulas] {name} shaded={shaded} fp64", flush=True)
    f64, st = renderer.render(precision=_abi.PREC_FP64)
    m = compare(f64, ref)
    print(f"  {m}", flush=True)
    assert m["hits"] > 200
    if formula in (11, 15):
        # an iteration of xenodreambuie / mandelbulb4 calls pow, atan2, acos, sin and cos, whose last bits
        # differ between the device libm and glibc; the delta DE divides orbit-radius differences 1e-10
        # apart, so a hit can move within the binary search's acceptance band (0.01 thr, stretched by
        # 1 / cos of the incidence angle at grazing pixels): the bar is on the 99.9th percentile of the
        # depth error, with 1e-3 for the largest
        assert m["depth_rel_p999"] <= 1e-4 and m["depth_rel_max"] <= 1e-3, m
        assert_parity(m, depth_rel=1e-3, label=f"{name} shaded={shaded}")
    else:
        assert_parity(m, label=f"{name} shaded={shaded}")
    if formula in (1, 17):
        # trig_DE and smoothMandelbox are analytic: no finite differences, the FP64 path reproduces the oracle
        assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
        assert_planes_exact(m)
    _check_fp32(renderer, formula, f64, ref, f"{name} shaded={shaded}")


def _check_fp32(renderer, formula, f64, ref, label, ifs_fold=False):
    """The FP32 modes: the FP64 frame itself where the scene renders in FP64
    only; for the analytic generic formulas the FP64 march's depths and hits
    with FP32 shading, at the parity bars, and FP32_MARCH equal to
    FP32_FIXUP."""
    # under the IFS folding pre-fold formulas 2, 7 and 8 run the generic kernels too
    fp32 = formula in FP32_SHADED or (ifs_fold and formula in (2, 7, 8))
    frames = []
    for prec, pname in [(_abi.PREC_FP32_FIXUP, "fp32"), (_abi.PREC_FP32_MARCH, "fp32-march")]:
        print(f"[formulas] {label} {pname}" + ("" if fp32 else " (rendered in FP64)"), flush=True)
        f, _ = renderer.render(precision=prec)
        frames.append(f)
        if not fp32:
            assert np.array_equal(f.z, f64.z) and np.array_equal(f.rgb, f64.rgb, equal_nan=True)
    if fp32:
        f = frames[0]
        assert np.array_equal(f.z, f64.z)
        assert not np.array_equal(f.rgb, f64.rgb)   # the shading ran in FP32
        m = compare(f, ref)
        print(f"  fp32 vs oracle: {m}", flush=True)
        assert_parity(m, label=f"{label} fp32")
        assert np.array_equal(frames[1].z, f.z) and np.array_equal(frames[1].rgb, f.rgb, equal_nan=True)


def _ifs_scene(variant, shaded):
    s = scene.menger_view(192, 144, shaded=shaded)
    s.formula = 10
    s.ifs_abs[0] = s.ifs_abs[1] = s.ifs_abs[2] = 1
    s.ifs_scale = 3.0
    for c in range(3):
        s.ifs_offset[c] = 1.0
    s.ifs_menger_sponge_mode = 1
    for i, d in enumerate([(1, -1, 0), (1, 0, -1), (0, 1, -1)]):
        for c in range(3):
            s.ifs_direction[i][c] = d[c]
        s.ifs_enabled[i] = 1
    if variant == "rotated":
        # main rotation about the offset, a rotated fold and an edge fold
        s.ifs_rotation[0], s.ifs_rotation[1] = 0.3, -0.2
        s.ifs_angles[1][2] = 0.25
        s.ifs_edge[0] = 0.9
        s.ifs_menger_sponge_mode = 0
        s.ifs_scale = 2.6
    return s


@pytest.mark.parametrize("shaded", [False, True])
@pytest.mark.parametrize("variant", ["menger", "rotated"])
def test_kaleidoscopic_ifs_matches_oracle(renderer, variant, shaded):
    from oracle import oracle
    sc = _ifs_scene(variant, shaded)
    ref, _ = oracle.render(sc)
    renderer.set_scene(sc)
    print(f"[formulas] kaleidoscopic IFS {variant} shaded={shaded} fp64", flush=True)
    f64, _ = renderer.render(precision=_abi.PREC_FP64)
    m = compare(f64, ref)
    print(f"  {m}", flush=True)
    assert m["hits"] > 200
    assert_parity(m, label=f"ifs {variant} shaded={shaded}")
    _check_fp32(renderer, 10, f64, ref, f"kaleidoscopic IFS {variant} shaded={shaded}")


HYBRIDS = {
    # a Mandelbox-Mandelbulb hybrid: two box iterations (scale 2), one bulb, then bulbs
    "box+bulb": ({0: (8, 2, 2.0), 1: (2, 1, 8.0), 4: (2, 1, 8.0)}, 0, "box"),
    # cyclic building blocks: angleMultiplyZ, invertR, then one trig_optim
    "blocks": ({0: (36, 1, 2.0), 1: (25, 1, 0.5), 4: (2, 1, 8.0)}, 1, "bulb"),
}


@pytest.mark.parametrize("shaded", [False, True])
@pytest.mark.parametrize("name", sorted(HYBRIDS))
def test_hybrid_matches_oracle(renderer, name, shaded):
    """Hybrid sequences: the device picks each iteration's slot in closed
    form (hybrid_slot), the oracle from the explicit sequence the reference
    builds (CreateFormulaSequence); shading DEs run past N (normals at 5N,
    colouring at 10N) on the sequence's last entry."""
    from oracle import oracle
    slots, cyclic, cam = HYBRIDS[name]
    sc = scene.mandelbox_view(192, 144, shaded=shaded) if cam == "box" else (
        scene.mandelbulb_full(192, 144, seed=1) if shaded else scene.mandelbulb_plain(256, 192))
    if shaded:
        sc.ao_quality = 2
    sc.formula = 12
    sc.N = 30
    for k in range(5):
        f, it, pw = slots.get(k, (0, 1, 2.0))
        sc.hybrid_formula[k], sc.hybrid_iterations[k], sc.hybrid_power[k] = f, it, pw
    sc.hybrid_cyclic = cyclic
    ref, _ = oracle.render(sc)
    renderer.set_scene(sc)
    print(f"[formulas] hybrid {name} shaded={shaded} fp64", flush=True)
    f64, _ = renderer.render(precision=_abi.PREC_FP64)
    m = compare(f64, ref)
    print(f"  {m}", flush=True)
    assert m["hits"] > 200
    # trig_optim slots bring pow/atan2/sin/cos into every cycle, and the Mandelbox slots fold the
    # orbit's last-bit differences across fold edges: a few grazing hits move further within the
    # binary search's band.  Such pixels count as differing pixels: with the flips, at most 0.1%.
    both = (f64.z < 1e19) & (ref.z < 1e19)
    rel = np.abs(f64.z[both].astype(np.float64) - ref.z[both]) / ref.z[both]
    outliers = int(np.count_nonzero(rel > 1e-4))
    print(f"  depth outliers (> 1e-4): {outliers}", flush=True)
    assert m["flips"] + outliers <= 1e-3 * m["pixels"], m
    assert m["psnr_db"] >= 45.0, m


def _fold_scene(kind, shaded):
    """generalizedFoldBox (gf<type>), mandelboxVaryScale4D (vary, vary-rpow)
    and the IFS folding pre-fold on the Mandelbox (ifs-box) and on the
    Mandelbulb (ifs-bulb)."""
    if kind == "ifs-bulb":
        s = scene.mandelbulb_full(192, 144, seed=1) if shaded else scene.mandelbulb_plain(256, 192)
    else:
        s = scene.mandelbox_view(192, 144, shaded=shaded)
    if kind.startswith("gf"):
        s.formula = 37
        s.mb_fold_mode = int(kind[2:])
        s.mb_solid, s.mb_melt = 1.0, 0.05 if s.mb_fold_mode % 2 else 0.0
    elif kind.startswith("vary"):
        s.formula = 18
        s.mb_vary_scale_vary, s.mb_vary_fold, s.mb_vary_minr = 0.1, 1.0, 0.5
        s.mb_vary_rpower, s.mb_vary_wadd = (1.1 if kind == "vary-rpow" else 1.0), 0.05
    else:
        s.ifs_folding_mode = 1
        s.ifs_enabled[0] = s.ifs_enabled[1] = 1
        s.ifs_abs[0 if kind == "ifs-bulb" else 2] = 1
        for c, v in enumerate((1.0, -1.0, 0.0)):
            s.ifs_direction[0][c] = v
        for c, v in enumerate((0.0, 1.0, -1.0)):
            s.ifs_direction[1][c] = v
        s.ifs_distance[1] = 0.1
        s.ifs_angles[1][2] = 0.2
        s.ifs_rotation[0] = 0.15
        s.ifs_offset[0] = 0.2
        s.ifs_scale = 1.0
    if shaded:
        s.ao_quality = 2
    return s


# analytic and free of transcendentals (pow(x, 1) = x): the FP64 path must
# reproduce the oracle bit for bit
FOLD_EXACT = ("gf0", "gf1", "gf2", "gf3", "gf4", "gf5", "gf6", "gf7", "vary", "ifs-box")


@pytest.mark.parametrize("shaded", [False, True])
@pytest.mark.parametrize("kind", FOLD_EXACT + ("vary-rpow", "ifs-bulb"))
def test_fold_formulas_match_oracle(renderer, kind, shaded):
    from oracle import oracle
    sc = _fold_scene(kind, shaded)
    ref, _ = oracle.render(sc)
    renderer.set_scene(sc)
    print(f"[formulas] {kind} shaded={shaded} fp64", flush=True)
    f64, _ = renderer.render(precision=_abi.PREC_FP64)
    m = compare(f64, ref)
    print(f"  {m}", flush=True)
    assert m["hits"] > 200
    assert_parity(m, label=f"{kind} shaded={shaded}")
    if kind in FOLD_EXACT:
        assert m["flips"] == 0 and m["depth_rel_max"] <= 1e-12, m
        assert_planes_exact(m)
    _check_fp32(renderer, sc.formula, f64, ref, f"{kind} shaded={shaded}", ifs_fold=bool(sc.ifs_folding_mode))

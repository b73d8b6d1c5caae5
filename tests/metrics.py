"""Image comparison metrics of the parity tests.

The bars (per config): at most 0.1% of pixels may differ, the hit depth
(zBuffer) relative error on agreeing pixels must be <= 1e-4, and the RGB
PSNR of imageFloat must be >= 45 dB (peak 1.0).

A differing pixel ("flips", "flip_frac") is one of: a hit/miss flip
("hitmiss_px"); a sheet jump ("sheet_px": both renders hit, but at depths
1% or more apart, i.e. on another sheet of the surface, which happens at
grazing silhouettes when the last bits of a DE differ); a NaN colour in one
render only.  Agreeing pixels are the hits on the same sheet; "depth_rel_max"
is taken over them, and "depth_rel_max_all" over every pixel both renders
hit (sheet jumps included), so neither count hides the other.
"""
import numpy as np

MISS_Z = 1e19


def compare(a, b):
    """a, b: frames with .rgb (3,h,w) f32 and .z (h,w) f32 planes."""
    hit_a = a.z < MISS_Z
    hit_b = b.z < MISS_Z
    n = hit_a.size
    hitmiss = int(np.count_nonzero(hit_a != hit_b))
    flips = hitmiss
    sheet = 0
    depth_max_all = 0.0
    both = hit_a & hit_b
    if both.any():
        za = a.z[both].astype(np.float64)
        zb = b.z[both].astype(np.float64)
        rel = np.abs(za - zb) / np.maximum(np.abs(zb), 1e-30)
        depth_max_all = float(rel.max())
        # agreeing pixels: hit the same sheet (depths within 1%); a ray that
        # lands on another sheet is a differing pixel
        same = rel < 1e-2
        sheet = int(np.count_nonzero(~same))
        flips += sheet
        rel_same = rel[same]
        depth_max = float(rel_same.max()) if rel_same.size else 0.0
        depth_p999 = float(np.quantile(rel_same, 0.999)) if rel_same.size else 0.0
        depth_bad = int(np.count_nonzero(rel_same > 1e-4))
    else:
        depth_max = depth_p999 = 0.0
        depth_bad = 0
    ra, rb = a.rgb.astype(np.float64), b.rgb.astype(np.float64)
    d = ra - rb
    # NaN colours (the reference's NaN orbits, e.g. mandelbulb2 on an axis):
    # equal where both renders have them; a pixel with a NaN channel in one
    # render only is a differing pixel (counted with the flips) and is left
    # out of the PSNR
    na, nb = np.isnan(ra), np.isnan(rb)
    d[na & nb] = 0.0
    nan_px = np.any(na ^ nb, axis=0)
    flips += int(np.count_nonzero(nan_px & ~(hit_a != hit_b)))
    d[:, nan_px] = 0.0
    mse = float(np.mean(d * d))
    psnr = float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)
    out = {
        "pixels": n, "hits": int(hit_b.sum()), "flips": flips, "flip_frac": flips / n,
        "hitmiss_px": hitmiss, "sheet_px": sheet, "depth_rel_max_all": depth_max_all,
        "depth_rel_max": depth_max, "depth_rel_p999": depth_p999, "depth_bad_px": depth_bad, "psnr_db": psnr,
        "rgb_max_abs": float(np.abs(d).max()), "nan_mismatch_px": int(np.count_nonzero(nan_px)),
    }
    out.update(integer_planes(a, b))
    return out


def integer_planes(a, b):
    """The integer planes of the framebuffer (colourBuffer u8 x 3, alphaBuffer
    and opacityBuffer u16, src/cimage.hpp:157-162): pixels whose values
    differ and the largest difference, per plane."""
    out = {}
    for name in ("colour", "alpha", "opacity"):
        pa, pb = getattr(a, name, None), getattr(b, name, None)
        if pa is None or pb is None:
            continue
        da = np.abs(pa.astype(np.int64) - pb.astype(np.int64))
        if da.ndim == 3:
            da = da.max(axis=0)
        out[f"{name}_diff_px"] = int(np.count_nonzero(da))
        out[f"{name}_max_diff"] = int(da.max()) if da.size else 0
    return out


def assert_planes_exact(m, label=""):
    """The FP64 path's integer planes equal the oracle's bit for bit."""
    for name in ("colour", "alpha", "opacity"):
        assert m.get(f"{name}_diff_px", 0) == 0, f"{label}: {name} plane differs {m}"


def assert_planes_close(m, frac=1e-3, label=""):
    """FP32 modes: the integer planes differ on at most `frac` of the
    pixels (a colour index or an opacity sum crossing a rounding boundary)."""
    for name in ("colour", "alpha", "opacity"):
        assert m.get(f"{name}_diff_px", 0) <= frac * m["pixels"], f"{label}: {name} plane differs {m}"


def assert_parity(m, flip_frac=1e-3, depth_rel=1e-4, psnr_db=45.0, label=""):
    assert m["flip_frac"] <= flip_frac, f"{label}: hit/miss flips {m}"
    assert m["depth_rel_max"] <= depth_rel, f"{label}: depth error {m}"
    assert m["psnr_db"] >= psnr_db, f"{label}: PSNR {m}"

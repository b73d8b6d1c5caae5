#!/usr/bin/env python
"""Benchmark of the render core: Mrays/s of the 1080p Mandelbulb frame.

Workload (one step = one frame): Mandelbulb power 8, 1920x1080, N 250, camera
at (0,-3.5,0) looking at the origin, fov 1, one hard-shadow ray per hit toward
a seeded unit light direction, AO quality 8 (412 directions), volumetric fog
density 0.5.  The frame renders into device-resident planes through the C ABI
(mbr_render_device); Mrays/s = primary rays / wall time / 1e6.

Printed as one JSON line: the FP32+FP64-fixup render is `value`; the plain
FP64 GPU render, the pixel difference between the two, the roofline of the
dominant kernel and the CPU baseline (the oracle on the host cores, a bounded
sample of the same frame) ride along.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "Mrays/s at 1080p Mandelbulb: FP32+FP64-fixup vs plain FP64, ≤0.1% px diff"

# Published MI355X vector peaks (TFLOP/s).  The FP32 figure assumes packed
# FP32 issue, which this build leaves disabled; the unpacked FP32 rate equals
# the FP64 one.
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}
# pixel differences reported between two renders (tests/metrics.py): differing
# pixels = hit/miss flips + sheet jumps (+ NaN-only colours)
DIFF_KEYS = ("flip_frac", "hitmiss_px", "sheet_px", "depth_rel_max", "depth_rel_max_all", "psnr_db")

# HBM bytes per k_ao launch on this workload (FETCH_SIZE x 2 + WRITE_SIZE x 1,
# the factors that make a 1 GiB device copy read 1 GiB and write 1 GiB),
# from the rocprofv3 --pmc runs summarised in the file named here.
AO_TRAFFIC_BYTES = 447.6e6
AO_TRAFFIC_SOURCE = ("profiles/render_1080p_mandelbulb_pmc.txt (per launch, one of the frame's two bands: "
                     "FETCH_SIZE 42.2 MiB x 2 + WRITE_SIZE 342.6 MiB)")


def planes_on_gpu(torch, w, h):
    dev = torch.device("cuda:0")
    return {
        "rgb": torch.empty((3, h, w), dtype=torch.float32, device=dev),
        "z": torch.empty((h, w), dtype=torch.float32, device=dev),
        "colour": torch.empty((3, h, w), dtype=torch.uint8, device=dev),
        "alpha": torch.empty((h, w), dtype=torch.int16, device=dev),
        "opacity": torch.empty((h, w), dtype=torch.int16, device=dev),
    }


def to_frame(torch, planes):
    from mandelbulber_amd.render import Frame
    h, w = planes["z"].shape
    f = Frame(w, h)
    f.rgb[:] = planes["rgb"].cpu().numpy()
    f.z[:] = planes["z"].cpu().numpy()
    return f


def timed(torch, r, planes, precision, steps, warmup):
    for _ in range(warmup):
        r.render_device(planes, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        stats.append(r.render_device(planes, precision=precision))
    torch.cuda.synchronize()
    return time.perf_counter() - t0, stats


def host_cores():
    """(cores this process may use, what nproc reports, the cgroup CPU quota
    or None): the CPUs in the affinity mask, capped by the cgroup quota."""
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    cores = min(avail, quota) if quota else avail
    return cores, nproc, quota


def cpu_baseline(sc, seconds, threads):
    """The oracle (-O3 -ffast-math, the reference's release flags) on a
    uniform sample of the frame: bands of rows spread evenly from the top to
    the bottom, so that sky, silhouette and surface rows count in
    proportion.  A first short sample of the same shape sets the number of
    rows so that the timed sample takes about `seconds`."""
    from oracle import oracle
    w, h = sc.image_width, sc.image_height

    def bands(total):
        # bands of at least two rows per thread (the oracle hands out whole
        # rows), at most 12 of them, spread evenly
        band = min(h, total, max(2 * threads, -(-total // 12)))
        n = max(1, total // band)
        if n == 1:
            return [((h - band) // 2, (h - band) // 2 + band)]
        return [(int(round(i * (h - band) / (n - 1))), int(round(i * (h - band) / (n - 1))) + band) for i in range(n)]

    def run(total):
        t0 = time.perf_counter()
        done = 0
        for r0, r1 in bands(total):
            oracle.render(sc, rows=(r0, r1), threads=threads, fast=True)
            done += r1 - r0
        return time.perf_counter() - t0, done

    dt, done = run(min(h, 4 * 2 * threads))
    total = max(1, min(h, int(done * seconds / max(dt, 1e-3))))
    dt, done = run(total)
    rays = done * w
    sample = bands(total)
    return {
        "value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": f"{len(sample)} bands of {sample[0][1] - sample[0][0]} rows spread evenly over the {w}x{h} frame "
                  f"({done} rows, {rays} rays, {dt:.1f} s), oracle/liboracle_fast.so (-O3 -ffast-math), "
                  f"{threads} threads",
    }


def bench_flight(torch, r, _abi, scene, n_frames, replays):
    """The batched flight: n frames of mixed formulas and resolutions as one
    captured HIP graph (autotuned launch shape) against the same frames
    rendered one mbr_set_scene + mbr_render_device call after another (the
    reference renders a flight one Render() per frame)."""
    fl = scene.flight(n_frames, seed=0)
    rays = sum(s.image_width * s.image_height for s in fl)
    seq_planes = [planes_on_gpu(torch, s.image_width, s.image_height) for s in fl]
    bat_planes = [planes_on_gpu(torch, s.image_width, s.image_height) for s in fl]
    # sequential: warm-up pass, then a timed pass
    for timed_pass in (False, True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s, p in zip(fl, seq_planes):
            r.set_scene(s)
            r.render_device(p, precision=_abi.PREC_FP32_FIXUP)
        torch.cuda.synchronize()
        dt_seq = time.perf_counter() - t0
    # batch: autotune (times 9 launch shapes), then `replays` launches of the tuned graph
    _, tune = r.render_batch(fl, bat_planes, autotune=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, res = r.render_batch(fl, bat_planes, replays=replays)
    torch.cuda.synchronize()
    dt_call = time.perf_counter() - t0
    same = all(torch.equal(a[k], b[k]) for a, b in zip(seq_planes, bat_planes) for k in ("z", "colour", "alpha"))
    same = same and all(torch.equal(torch.nan_to_num(a["rgb"]), torch.nan_to_num(b["rgb"]))
                        for a, b in zip(seq_planes, bat_planes))
    return {
        "workload": f"{n_frames} frames, formulas cycling Mandelbulb (power U[6,10]) / Mandelbox / Menger (N 15), "
                    f"640x480..1920x1080, camera U[1.5,5] (x2 for the Mandelbox) looking at the origin, hard shadow",
        "rays": rays, "hits": st.hits,
        "graph_value": rays / (res.ms_per_launch * 1e-3) / 1e6, "unit": "Mrays/s",
        "graph_ms_per_launch": res.ms_per_launch, "graph_nodes": res.graph_nodes,
        "graph_build_ms": res.ms_build, "call_ms_incl_build": dt_call * 1e3,
        "sequential_value": rays / dt_seq / 1e6, "sequential_ms": dt_seq * 1e3,
        "speedup_vs_sequential": dt_seq * 1e3 / res.ms_per_launch,
        "tuned": {"slots": res.slots, "blocks_per_cu": res.blocks_per_cu,
                  "candidates_ms": {f"{a}x{b}": tune.tuned_ms[i]
                                    for i, (a, b) in enumerate((a, b) for a in (1, 2, 4) for b in (4, 8, 16))}},
        "identical_to_sequential": bool(same),
    }


def bench_adjoint(torch, r, scene, n_scenes, size):
    """Differentiable rendering: n Mandelbulb scenes (power U[6,10], seeded
    camera and light) at size x size; MSE against the image at power 7.3;
    dL/d(power, camera, light) from the adjoint kernel through the
    torch.autograd.Function, timed, and checked against central
    differences (h = 1e-4) of the FP64 forward on every scene
    (diff.check_gradient: per parameter over the pixels where the image is
    smooth within +-h; fd_r and vjp_r with the h^2 term removed)."""
    import numpy as np
    from mandelbulber_amd import diff
    fwd_ms, bwd_ms = [], []
    rel = {"fd_r": [], "vjp_r": [], "fd_h": []}
    left = []
    for i in range(n_scenes):
        base = scene.differentiable(size, size, seed=100 + i)
        theta = torch.tensor(scene.get_params(base), dtype=torch.float64, requires_grad=True)
        tt = theta.detach().clone()
        tt[0] = 7.3
        with torch.no_grad():
            target = diff.render(tt, r, base).double()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        img = diff.render(theta, r, base)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss = ((img.double() - target) ** 2).mean()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd_ms.append((t1 - t0) * 1e3)
        bwd_ms.append((t2 - t1) * 1e3)
        chk = diff.check_gradient(r, base, h=1e-4)
        rel["fd_r"].append(diff.compare(chk.adj, chk.fd_r)[0])
        rel["vjp_r"].append(diff.compare(chk.adj, chk.vjp_r)[0])
        rel["fd_h"].append(diff.compare(chk.adj, chk.fd)[0])
        left.append(chk.left / max(chk.region_px, 1))
    left = np.array(left)
    fd_check = {"h": 1e-4, "params": list(scene.DIFF_PARAMS),
                "hit_pixels_left_out_mean_per_param": [round(float(v), 4) for v in left.mean(axis=0)],
                "hit_pixels_left_out_mean": float(left.mean())}
    for key, v in rel.items():
        v = np.array(v)
        fd_check[key] = {"max_rel_err": float(v.max()), "median_rel_err": float(np.median(v)),
                         "components_above_1e-3": int((v > 1e-3).sum()), "components": int(v.size)}
    rays = size * size
    return {
        "workload": f"{n_scenes} Mandelbulb scenes {size}x{size}, power U[6,10], seeded camera and light, "
                    f"MSE vs power 7.3; gradients in power, vp (3), alfa/beta/gamma, light alfa/beta",
        "forward_ms_per_scene": sum(fwd_ms) / n_scenes, "backward_ms_per_scene": sum(bwd_ms) / n_scenes,
        "backward_Mrays_per_s": rays / (sum(bwd_ms) / n_scenes * 1e-3) / 1e6,
        "fd_check": fd_check,
    }


def bench_deep_zoom(torch, r, _abi, scene, compare, width, height):
    """Deep-zoom Mandelbox (frame width 1e-6 at a seeded surface point),
    FP32_FIXUP against the FP64 frame: time and the error of FP32 shading
    (whose distance estimates start in FP64 at the hit point)."""
    sc = scene.deep_zoom(width, height, seed=1)
    r.set_scene(sc)
    planes = planes_on_gpu(torch, width, height)
    out = {"workload": f"Mandelbox scale 2 (fold limit 1, min radius 0.5, fixed radius 1), {width}x{height}, "
                       f"frame width 1e-6 at a seeded surface point, hard shadow", "rays": width * height}
    frames = {}
    for prec, name in ((_abi.PREC_FP32_FIXUP, "fp32_fixup"), (_abi.PREC_FP64, "fp64")):
        r.render_device(planes, precision=prec)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = r.render_device(planes, precision=prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        frames[name] = to_frame(torch, planes)
        out[name] = {"value": width * height / dt / 1e6, "unit": "Mrays/s", "ms": dt * 1e3, "hits": st.hits}
    d = compare(frames["fp32_fixup"], frames["fp64"])
    out["fp32_vs_fp64"] = {k: d[k] for k in ("flip_frac", "hitmiss_px", "sheet_px", "depth_rel_max", "depth_rel_max_all", "psnr_db", "rgb_max_abs")}
    out["speedup_vs_fp64"] = out["fp32_fixup"]["value"] / out["fp64"]["value"]
    return out


def bench_post(torch, r, _abi, sc, planes, with_cpu):
    """Post effects on the 1080p frame: SSAO (quality from the scene, 20 by
    default) and DOF (focus at the median hit depth, radius 10) on device
    planes; with_cpu: the oracle's SSAO on the same frame, timed on one host
    thread and compared with the GPU result."""
    import math
    import numpy as np
    from mandelbulber_amd import scene as S
    from mandelbulber_amd.render import compile_image
    r.set_scene(sc)
    r.render_device(planes, precision=_abi.PREC_FP32_FIXUP)
    torch.cuda.synchronize()
    h, w = planes["z"].shape
    rgb = planes["rgb"].cpu().numpy()
    img = compile_image(sc, rgb)
    z_host = planes["z"].cpu().numpy()
    col = planes["colour"]
    img_dev = torch.from_numpy(img.view(np.int16)).to(planes["z"].device)
    out = {"workload": f"{w}x{h} Mandelbulb frame of the main workload, compiled to 16 bits"}

    def timed_call(fn, n):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    work = img_dev.clone()
    compile_ms = timed_call(lambda: r.compile_image_device(sc, planes["rgb"], work), 3)
    out["compile_image"] = {"ms": compile_ms, "equal_to_host": bool(np.array_equal(
        work.cpu().numpy().view(np.uint16), img))}
    ssao_ms = timed_call(lambda: r.post_ssao(sc, planes["z"], planes["opacity"], [col[0], col[1], col[2]],
                                             work.copy_(img_dev)), 3)
    out["ssao"] = {"quality": sc.ssao_quality, "ms": ssao_ms,
                   "Mpixels_per_s": w * h / ssao_ms / 1e3}
    if with_cpu:
        from oracle import oracle
        op_h = planes["opacity"].cpu().numpy().view(np.uint16)
        col_h = col.cpu().numpy()
        t0 = time.perf_counter()
        want = oracle.ssao(sc, z_host, op_h, col_h, img)
        cpu_s = time.perf_counter() - t0
        r.post_ssao(sc, planes["z"], planes["opacity"], [col[0], col[1], col[2]], work.copy_(img_dev))
        torch.cuda.synchronize()
        got = work.cpu().numpy().view(np.uint16)
        out["ssao"]["cpu_oracle_ms"] = cpu_s * 1e3
        out["ssao"]["cpu_oracle_threads"] = 1
        out["ssao"]["equal_to_oracle"] = bool(np.array_equal(got, want))
    hit_z = z_host[z_host < 1e19]
    d = S.copy(sc)
    d.dof_enabled = 1
    d.dof_focus = 10.0 * (16.0 + math.log10(float(np.median(hit_z)))) if hit_z.size else 21.7
    d.dof_radius = 10.0
    alpha = planes["alpha"].clone()
    dof_ms = timed_call(lambda: r.post_dof(d, planes["z"], work.copy_(img_dev), alpha.copy_(planes["alpha"])), 2)
    out["dof"] = {"radius": d.dof_radius, "focus_depth": float(np.median(hit_z)) if hit_z.size else None,
                  "ms": dof_ms, "note": "includes the host z sort and shuffle that fix the splat order"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flight", action="store_true")
    ap.add_argument("--flight-frames", type=int, default=64)
    ap.add_argument("--flight-replays", type=int, default=3)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-deep-zoom", action="store_true")
    ap.add_argument("--no-post", action="store_true")
    ap.add_argument("--adjoint-scenes", type=int, default=16)
    ap.add_argument("--adjoint-size", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this render core runs on one GPU; multi-GPU runs are not part of this build "
              "(frames and tiles would shard across GPUs with one framebuffer gather).", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    from mandelbulber_amd import Renderer, _abi, scene
    from mandelbulber_amd.flops import algorithmic_flops
    from metrics import compare

    sc = scene.mandelbulb_full(args.width, args.height, seed=args.seed)
    r = Renderer(0)
    r.set_scene(sc)
    planes = planes_on_gpu(torch, args.width, args.height)
    rays = args.width * args.height

    dt32, st32 = timed(torch, r, planes, _abi.PREC_FP32_FIXUP, args.steps, args.warmup)
    f32 = to_frame(torch, planes)
    dt64, st64 = timed(torch, r, planes, _abi.PREC_FP64, args.fp64_steps, 0)
    f64 = to_frame(torch, planes)
    diff = compare(f32, f64)
    dtm, stm = timed(torch, r, planes, _abi.PREC_FP32_MARCH, 1, 1)
    fm = to_frame(torch, planes)
    diffm = compare(fm, f64)
    # FP64 with the FP32 modes' trig-free power-8 iteration (march,
    # continuation, AO): splits the FP32/FP64 ratio into the iteration form
    # (fp64 / fp64_demoivre) and the arithmetic (fp64_demoivre / fp32)
    dtd, std_ = timed(torch, r, planes, _abi.PREC_FP64_DEMOIVRE, args.fp64_steps, 0)
    fd = to_frame(torch, planes)
    diffd = compare(fd, f64)

    value = rays * args.steps / dt32 / 1e6
    value64 = rays * args.fp64_steps / dt64 / 1e6

    # roofline of the dominant kernel (k_ao, the AmbientOcclusion rays):
    # algorithmic FLOPs per launch / average launch duration, the duration
    # from the library's HIP events recorded around the launch on its stream
    launches = max(1, sum(s.shade_launches for s in st32))
    ao_ms = sum(s.ms_ao for s in st32) / launches
    ao_flops = sum(algorithmic_flops(sc, s, "ao") for s in st32) / launches
    achieved = ao_flops / (ao_ms * 1e-3) / 1e12
    # the profiled traffic holds for the default frame only
    default_workload = (args.width, args.height, args.seed) == (1920, 1080, 1)
    # the frame's two bands run their AO launches concurrently on two
    # streams, so a launch's own duration includes time the other band holds
    # the CUs; the frame-level rate divides the frame's AO work by the span
    # from the first AO start to the last AO end
    ao_span = sum(s.ms_ao_span for s in st32) / len(st32)
    ao_flops_frame = sum(algorithmic_flops(sc, s, "ao") for s in st32) / len(st32)
    achieved_frame = ao_flops_frame / (ao_span * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS["f32"], "traffic": AO_TRAFFIC_BYTES if default_workload else None,
            "traffic_source": AO_TRAFFIC_SOURCE if default_workload else None, "kernel": "k_ao_f32_plain<float, 2, 8>",
            "flops_per_launch": ao_flops, "ms_per_launch": ao_ms,
            "launches_per_frame": launches / len(st32),
            "achieved_frame": achieved_frame, "frac_frame": achieved_frame / PEAK_TFLOPS["f32"],
            "ao_span_ms_per_frame": ao_span, "flops_per_frame": ao_flops_frame,
            "ao_de_calls": st32[-1].ao_de_calls, "ao_iterations": st32[-1].ao_iterations,
            "ao_steps": st32[-1].ao_steps}
    launches64 = max(1, sum(s.shade_launches for s in st64))
    ao64_ms = sum(s.ms_ao for s in st64) / launches64
    ao64_flops = sum(algorithmic_flops(sc, s, "ao") for s in st64) / launches64

    post = None if args.no_post else bench_post(torch, r, _abi, sc, planes, not args.no_cpu)
    flight = None if args.no_flight else bench_flight(torch, r, _abi, scene, args.flight_frames, args.flight_replays)
    adjoint = None if args.no_adjoint else bench_adjoint(torch, r, scene, args.adjoint_scenes, args.adjoint_size)
    deep = None if args.no_deep_zoom else bench_deep_zoom(torch, r, _abi, scene, compare, 3840, 2160)

    cpu = None
    if not args.no_cpu:
        threads, nproc, quota = host_cores()
        cpu = cpu_baseline(sc, args.cpu_seconds, threads)
        cpu["nproc"] = nproc
        cpu["cgroup_cpu_quota"] = quota

    s0 = st32[-1]
    line = {
        "metric": METRIC, "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt32 / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": f"Mandelbulb power 8, {args.width}x{args.height}, N 250, hard shadow, AO quality 8 "
                        f"({s0.shade_de_calls // max(1, s0.hits)} DE calls per hit), volumetric fog 0.5, "
                        f"light seed {args.seed}",
            "precision": "FP64 primary-ray march (origins, scan, hit decisions, binary search); FP32 DE "
                     "iteration in normals, shadow, AO, colouring; FP32 fog and composition",
            "rays": rays, "hits": s0.hits,
        },
        "fp64": {"value": value64, "unit": "Mrays/s", "ms_per_step": dt64 / args.fp64_steps * 1e3,
                 "kernel_ms": {"march": st64[-1].ms_march, "shade": st64[-1].ms_shade, "ao": st64[-1].ms_ao,
                               "finish": st64[-1].ms_finish},
                 "ao_achieved_tflops": ao64_flops / (ao64_ms * 1e-3) / 1e12,
                 "ao_frac_of_f64_peak": ao64_flops / (ao64_ms * 1e-3) / 1e12 / PEAK_TFLOPS["f64"]},
        "speedup_vs_fp64": value / value64,
        "fp64_demoivre": {"value": rays * args.fp64_steps / dtd / 1e6, "unit": "Mrays/s",
                          "ms_per_step": dtd / args.fp64_steps * 1e3,
                          "kernel_ms": {"march": std_[-1].ms_march, "shade": std_[-1].ms_shade,
                                        "ao": std_[-1].ms_ao, "finish": std_[-1].ms_finish},
                          "pixel_diff_vs_fp64": {k: diffd[k] for k in DIFF_KEYS},
                          "speedup_from_iteration_form": dt64 / dtd,
                          "speedup_from_fp32_arithmetic": (dtd / args.fp64_steps) / (dt32 / args.steps)},
        "fp32_march": {"value": rays / dtm / 1e6, "unit": "Mrays/s", "march_ms": stm[-1].ms_march,
                       "fp64_redo_rays": stm[-1].fp64_redo_rays,
                       "pixel_diff_vs_fp64": {k: diffm[k] for k in DIFF_KEYS}},
        "pixel_diff_vs_fp64": {k: diff[k] for k in DIFF_KEYS},
        "kernel_ms": {"march": s0.ms_march, "shade": s0.ms_shade, "ao": s0.ms_ao, "finish": s0.ms_finish},
        "march_counters": {"fp64_redo_rays": s0.fp64_redo_rays, "step_overflow_rays": s0.step_overflow_rays,
                           "march_de_calls": s0.march_de_calls, "shade_de_calls": s0.shade_de_calls,
                           "fp64_redo_rays_fp64_path": st64[-1].fp64_redo_rays,
                           "step_overflow_rays_fp64_path": st64[-1].step_overflow_rays},
        "roofline": roof,
        "flight": flight,
        "adjoint": adjoint,
        "deep_zoom": deep,
        "post": post,
        "cpu_baseline": cpu,
    }
    r.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()

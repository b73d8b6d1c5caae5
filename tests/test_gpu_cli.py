"""The -nogui driver on the GPU: a settings file with -o / -res overrides is
rendered, compiled and written; the PNG equals the library's render of the
same settings, compiled the same way."""
import os
import subprocess

import numpy as np
import pytest

from mandelbulber_amd import _abi, scene
from mandelbulber_amd.render import compile_image

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "mandelbulber_amd", "mbr_nogui")

SETTINGS = """formula 2;
power 8;
N 250;
view_point_x 0;
view_point_y -3.5;
view_point_z 0;
zoom 1e-15;
perspective 1;
shadows_enabled 1;
penetrating_lights 0;
main_light_alfa -45;
main_light_beta 30;
ambient_occlusion_enabled 0;
brightness 1.3;
gamma 1.8;
image_width 320;
image_height 240;
"""


def test_cli_renders_settings_file(tmp_path):
    from test_image import _read_png
    f = tmp_path / "scene.fract"
    f.write_text(SETTINGS)
    out = tmp_path / "img" / "frame.png"
    print("[cli] rendering", flush=True)
    r = subprocess.run([CLI, "-nogui", "-res", "96x72", "-o", "contrast=1.2", "-format", "png16",
                        "-out", str(out), str(f)], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "image saved" in r.stdout and out.exists()
    px, bits = _read_png(str(out))
    assert bits == 16 and px.shape == (72, 96, 3)
    # the same through the library
    s = scene.load(str(f), ["image_width=96", "image_height=72", "contrast=1.2"])
    from mandelbulber_amd import Renderer
    rr = Renderer(0)
    try:
        rr.set_scene(s)
        frame, _ = rr.render(precision=_abi.PREC_FP32_FIXUP)
        expect = compile_image(s, frame.rgb)
        # post_SSAO_enabled is on in the defaults: the driver adds it after CompileImage
        assert s.ssao_enabled == 1 and "SSAO rendered" in r.stdout
        expect = rr.post_ssao(s, frame.z, frame.opacity, frame.colour, expect)
    finally:
        rr.close()
    assert np.array_equal(px.astype(np.uint16), expect)
    assert s.brightness == 1.3 and s.contrast == 1.2 and abs(s.image_gamma - 1.8) < 1e-15


def test_cli_default_output_name(tmp_path):
    f = tmp_path / "scene.fract"
    f.write_text(SETTINGS + f"file_destination {tmp_path}/out/image;\n")
    r = subprocess.run([CLI, "-res", "32x24", "-format", "png", "-start", "7", str(f)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "out" / "image00007.png").exists()


def _library_png(s, precision=_abi.PREC_FP32_FIXUP):
    """What the driver writes for scene s: render, CompileImage, SSAO."""
    from mandelbulber_amd import Renderer
    rr = Renderer(0)
    try:
        rr.set_scene(s)
        frame, _ = rr.render(precision=precision)
        img = compile_image(s, frame.rgb)
        if s.ssao_enabled:
            img = rr.post_ssao(s, frame.z, frame.opacity, frame.colour, img)
    finally:
        rr.close()
    return img


def test_cli_keyframe_animation(tmp_path):
    from test_image import _read_png
    kdir = tmp_path / "kf"
    kdir.mkdir()
    for k, (power, vy) in enumerate([(8, -3.5), (7, -3.2), (6, -3.0)]):
        (kdir / f"key{k:05d}.fract").write_text(SETTINGS + f"power {power};\nview_point_y {vy};\n")
    main = tmp_path / "main.fract"
    main.write_text(SETTINGS + f"file_keyframes {kdir}/key;\nfile_destination {tmp_path}/anim/frame;\n"
                               "frames_per_keyframe 2;\nend_frame 1000;\n")
    print("[cli] keyframe animation", flush=True)
    r = subprocess.run([CLI, "-keyframe", "-res", "48x36", "-format", "png16", str(main)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Found 3 keyframes" in r.stdout
    frames = sorted(p.name for p in (tmp_path / "anim").iterdir())
    assert frames == [f"frame{i:05d}.png" for i in range(6)]   # end clamps to 3 keys x 2 frames
    keys = [scene.load(str(kdir / f"key{k:05d}.fract")) for k in range(3)]
    base = scene.load(str(main), ["image_width=48", "image_height=36"])
    for index in (2, 3):
        fs = scene.keyframe(keys, 2, index, base)
        fs.frame_no = index
        px, _ = _read_png(str(tmp_path / "anim" / f"frame{index:05d}.png"))
        assert np.array_equal(px.astype(np.uint16), _library_png(fs)), index
    assert scene.keyframe(keys, 2, 2, base).power == pytest.approx(7.0, rel=1e-12)   # log space: exp(log 7)


def test_cli_flight_path(tmp_path):
    from test_image import _read_png
    path = tmp_path / "path.txt"
    cams = [(0.0, -3.5, 0.0, 0.0, 0.0, 0.0), (0.1, -3.4, 0.05, 0.05, -0.02, 0.01),
            (0.2, -3.3, 0.1, 0.1, -0.04, 0.02)]
    path.write_text("".join("%.11f %.11f %.11f %f %f %f\n" % c for c in cams))
    main = tmp_path / "main.fract"
    main.write_text(SETTINGS + f"file_animation_path {path};\nfile_destination {tmp_path}/fl/frame;\n")
    print("[cli] flight path", flush=True)
    r = subprocess.run([CLI, "-flight", "-res", "48x36", "-format", "png16", "-end", "10", str(main)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    frames = sorted(p.name for p in (tmp_path / "fl").iterdir())
    assert frames == [f"frame{i:05d}.png" for i in range(3)]   # the path has three cameras
    base = scene.load(str(main), ["image_width=48", "image_height=36"])
    fs = scene.copy(base)
    fs.vp[0], fs.vp[1], fs.vp[2], fs.alpha, fs.beta, fs.gamma = cams[1]
    fs.frame_no = 1
    px, _ = _read_png(str(tmp_path / "fl" / "frame00001.png"))
    assert np.array_equal(px.astype(np.uint16), _library_png(fs))


def test_cli_batched_flight_equals_frame_by_frame(tmp_path):
    """Animations render in batches (mbr_render_batch: one graph launch for
    up to -batch frames into device framebuffers); every frame's file equals,
    byte for byte, the file of the frame rendered on its own (-batch 1)."""
    path = tmp_path / "path.txt"
    cams = [(0.02 * i, -3.5 + 0.05 * i, 0.01 * i, 0.03 * i, -0.01 * i, 0.005 * i) for i in range(5)]
    path.write_text("".join("%.11f %.11f %.11f %f %f %f\n" % c for c in cams))
    out = {}
    for batch in (2, 1):
        main = tmp_path / f"main{batch}.fract"
        main.write_text(SETTINGS + f"file_animation_path {path};\nfile_destination {tmp_path}/b{batch}/frame;\n")
        print(f"[cli] flight -batch {batch}", flush=True)
        r = subprocess.run([CLI, "-flight", "-batch", str(batch), "-res", "48x36", "-format", "png16", "-end", "10",
                            str(main)], capture_output=True, text=True, timeout=300)
        print(r.stdout[-1200:], flush=True)
        assert r.returncode == 0, r.stdout + r.stderr
        if batch == 2:
            assert r.stdout.count("batch of 2 frames") == 2   # frames 0-1, 2-3; frame 4 alone
        else:
            assert "batch of" not in r.stdout
        out[batch] = [(tmp_path / f"b{batch}" / f"frame{i:05d}.png").read_bytes() for i in range(5)]
    assert out[2] == out[1]


def test_cli_tiles_join_into_the_full_frame(tmp_path):
    """tiles 2: four 48x36 tiles rendered one by one, stored, joined into
    file_destination_fromTiles.png (96x72, 16-bit) and removed; without SSAO
    (a screen-space effect sees each tile alone) and without step jitter (its
    hash is keyed by pixel index) the joined image equals the 96x72 frame."""
    from test_image import _read_png
    f = tmp_path / "scene.fract"
    dest = tmp_path / "out" / "tiled"
    f.write_text(SETTINGS + f"file_destination {dest};\npost_SSAO_enabled 0;\ntiles 2;\njitter_mode 0;\n")
    print("[cli] tiles 2", flush=True)
    r = subprocess.run([CLI, "-res", "48x36", "-format", "png16", "-precision", "fp64", str(f)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    out = tmp_path / "out" / "tiled_fromTiles.png"
    assert out.exists()
    assert not list((tmp_path / "out").glob("*.tile"))
    px, bits = _read_png(str(out))
    assert bits == 16 and px.shape == (72, 96, 3)
    s = scene.load(str(f), ["image_width=96", "image_height=72", "tiles=1"])
    assert s.jitter_mode == 0
    expect = _library_png(s, precision=_abi.PREC_FP64)
    assert np.array_equal(px.astype(np.uint16), expect)


def test_cli_animation_resumes_skipping_existing_frames(tmp_path):
    """-nogui skips a frame whose output file exists (src/Render3D.cpp:
    2335-2350): after an interrupted flight (frame 1 missing), a second run
    renders only frame 1 and leaves the other files untouched."""
    path = tmp_path / "path.txt"
    cams = [(0.0, -3.5, 0.0, 0.0, 0.0, 0.0), (0.1, -3.4, 0.05, 0.05, -0.02, 0.01),
            (0.2, -3.3, 0.1, 0.1, -0.04, 0.02)]
    path.write_text("".join("%.11f %.11f %.11f %f %f %f\n" % c for c in cams))
    main = tmp_path / "main.fract"
    main.write_text(SETTINGS + f"file_animation_path {path};\nfile_destination {tmp_path}/fl/frame;\n")
    cmd = [CLI, "-flight", "-res", "48x36", "-format", "png16", "-end", "10", str(main)]
    print("[cli] resume: first run", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    files = [tmp_path / "fl" / f"frame{i:05d}.png" for i in range(3)]
    full = [f.read_bytes() for f in files]
    files[1].unlink()
    # mark the kept frames: a resumed run must not rewrite them
    files[0].write_bytes(b"kept")
    print("[cli] resume: second run", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("exists. Skip to next file index") == 2
    assert r.stdout.count("image saved") == 1
    assert files[0].read_bytes() == b"kept"
    assert files[1].read_bytes() == full[1] and files[2].read_bytes() == full[2]


def test_cli_tiles_resume_joins_tiles_left_by_an_interrupted_run(tmp_path):
    """A tile file already present is not rendered again: it is joined as
    it is (here a tile of zeros planted in place of tile 3), the others are
    rendered and joined around it."""
    from test_image import _read_png
    f = tmp_path / "scene.fract"
    dest = tmp_path / "out" / "tiled"
    f.write_text(SETTINGS + f"file_destination {dest};\npost_SSAO_enabled 0;\ntiles 2;\njitter_mode 0;\n")
    (tmp_path / "out").mkdir()
    planted = tmp_path / "out" / "tiled00003.tile"
    planted.write_bytes(np.zeros(3 * 48 * 36, np.uint16).tobytes())
    print("[cli] tiles resume", flush=True)
    r = subprocess.run([CLI, "-res", "48x36", "-format", "png16", "-precision", "fp64", str(f)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("exists. Skip to next file index") == 1
    assert r.stdout.count("---------- Tile:") == 3
    px, _ = _read_png(str(tmp_path / "out" / "tiled_fromTiles.png"))
    s = scene.load(str(f), ["image_width=96", "image_height=72", "tiles=1"])
    expect = _library_png(s, precision=_abi.PREC_FP64)
    # tile 3 is the bottom-right quarter
    assert not px[36:, 48:].any()
    assert np.array_equal(px[:36].astype(np.uint16), expect[:36])
    assert np.array_equal(px[36:, :48].astype(np.uint16), expect[36:, :48])

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

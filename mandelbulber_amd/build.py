"""Build the render-core library mandelbulber_amd/libmbr.so with hipcc for gfx950,
the -nogui driver mandelbulber_amd/mbr_nogui and the counter-calibration
program mandelbulber_amd/mbr_calib.

Usage: python -m mandelbulber_amd.build   (or __graft_entry__.build())
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmbr.so")
CLI = os.path.join(HERE, "mbr_nogui")
CALIB = os.path.join(HERE, "mbr_calib")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["mbr_api.cpp", "mbr_kernels.hip", "mbr_adjoint.hip", "mbr_post.hip", "mbr_post_order.cpp",
           "mbr_settings.cpp", "mbr_image.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
    # packed FP32 instructions stay disabled: the compiler forms them from
    # ordinary float code, and some of their forms are unreliable here
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 square roots and divisions use the hardware v_sqrt_f32 / v_rcp_f32
    # (1 ulp) instead of the correctly rounded sequences: the FP32 paths are
    # held to tolerances; FP64 arithmetic is not affected
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


# per-source flags
EXTRA = {}


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-x", "hip", *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    # the -nogui driver: host C++ over the C ABI, found next to the library
    tmp = CLI + ".tmp"
    # it allocates the batched animation's device framebuffers with the HIP
    # runtime API (host code only)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(rocm, "include"), os.path.join(CSRC, "mbr_nogui.cpp"),
           "-o", tmp, "-L" + HERE, "-lmbr", "-L" + os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mbr_nogui build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CLI)
    # kernels of known instruction counts, for reading the profiler's counters
    tmp = CALIB + ".tmp"
    cmd = [HIPCC, "-x", "hip", *FLAGS, os.path.join(CSRC, "mbr_calib.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mbr_calib build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CALIB)
    if verbose:
        print(f"built {OUT} and {CLI}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)

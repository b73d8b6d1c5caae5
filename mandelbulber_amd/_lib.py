"""Loader of the render-core library (mandelbulber_amd/libmbr.so).

The library is the HIP implementation behind include/mbr.h.  There is no
fallback: if it is missing or its ABI does not match, import fails loudly.
"""
import ctypes as C
import os

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmbr.so")

# Every entry point declared in include/mbr.h.
EXPORTS = (
    "mbr_version", "mbr_scene_defaults", "mbr_load_settings", "mbr_set_setting", "mbr_init",
    "mbr_destroy", "mbr_last_error", "mbr_set_scene", "mbr_render", "mbr_render_device",
    "mbr_scene_layout", "mbr_jitter_draw", "mbr_render_batch", "mbr_adjoint", "mbr_compile_image",
    "mbr_write_png", "mbr_post_ssao", "mbr_post_dof", "mbr_compile_image_device", "mbr_keyframe_scene",
    "mbr_place_lights", "mbr_set_lights", "mbr_get_lights", "mbr_adjoint_jacobian",
    "mbr_render_device_noalloc", "mbr_cancel", "mbr_set_textures",
)

_lib = None


class MbrError(RuntimeError):
    def __init__(self, rc, msg=""):
        super().__init__(f"{_abi.STATUS.get(rc, rc)}: {msg}")
        self.rc = rc


def _declare(lib):
    P = C.POINTER
    lib.mbr_version.restype = C.c_int
    lib.mbr_scene_defaults.argtypes = [P(_abi.Scene)]
    lib.mbr_load_settings.argtypes = [C.c_char_p, P(C.c_char_p), C.c_int, P(_abi.Scene)]
    lib.mbr_set_setting.argtypes = [P(_abi.Scene), C.c_char_p, C.c_char_p]
    lib.mbr_init.argtypes = [P(_abi.DeviceOpts), P(C.c_void_p)]
    lib.mbr_destroy.argtypes = [C.c_void_p]
    lib.mbr_destroy.restype = None
    lib.mbr_last_error.argtypes = [C.c_void_p]
    lib.mbr_last_error.restype = C.c_char_p
    lib.mbr_set_scene.argtypes = [C.c_void_p, P(_abi.Scene), P(_abi.Texture)]
    lib.mbr_render.argtypes = [C.c_void_p, P(_abi.RenderOpts), P(_abi.Framebuffer), P(_abi.Stats)]
    lib.mbr_render_device.argtypes = [C.c_void_p, P(_abi.RenderOpts), P(_abi.Framebuffer), P(_abi.Stats)]
    lib.mbr_render_device_noalloc.argtypes = [C.c_void_p, P(_abi.RenderOpts), P(_abi.Framebuffer), P(_abi.Stats)]
    lib.mbr_scene_layout.argtypes = [C.c_size_t * 8]
    lib.mbr_scene_layout.restype = C.c_size_t
    lib.mbr_jitter_draw.argtypes = [C.c_uint32] * 4
    lib.mbr_jitter_draw.restype = C.c_uint32
    lib.mbr_adjoint.argtypes = [C.c_void_p, P(_abi.RenderOpts), C.c_void_p * 3, C.c_double * 9, C.c_void_p,
                                P(_abi.Stats)]
    lib.mbr_adjoint_jacobian.argtypes = [C.c_void_p, P(_abi.RenderOpts), C.c_void_p, C.c_void_p, P(_abi.Stats)]
    lib.mbr_compile_image.argtypes = [P(_abi.Scene), C.c_int, C.c_int, C.c_void_p * 3, C.c_void_p]
    lib.mbr_write_png.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    lib.mbr_post_ssao.argtypes = [C.c_void_p, P(_abi.Scene), P(_abi.PostPlanes), C.c_int, C.c_void_p]
    lib.mbr_compile_image_device.argtypes = [C.c_void_p, P(_abi.Scene), C.c_int, C.c_int, C.c_void_p * 3, C.c_void_p,
                                             C.c_void_p]
    lib.mbr_keyframe_scene.argtypes = [P(_abi.Scene), C.c_int, C.c_int, C.c_int, P(_abi.Scene)]
    lib.mbr_place_lights.argtypes = [P(_abi.Scene), P(_abi.Light), C.c_int, P(C.c_int)]
    lib.mbr_set_lights.argtypes = [C.c_void_p, P(_abi.Light), C.c_int]
    lib.mbr_get_lights.argtypes = [C.c_void_p, P(_abi.Light), C.c_int, P(C.c_int)]
    lib.mbr_post_dof.argtypes = [C.c_void_p, P(_abi.Scene), P(_abi.PostPlanes), C.c_int, C.c_void_p]
    lib.mbr_cancel.argtypes = [C.c_void_p]
    lib.mbr_set_textures.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mbr_render_batch.argtypes = [C.c_void_p, P(_abi.Scene), C.c_int, P(_abi.Texture), P(_abi.Framebuffer),
                                     P(_abi.BatchOpts), P(_abi.Stats), P(_abi.BatchResult)]


def lib():
    """The loaded library (raises OSError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        l = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            if not hasattr(l, name):
                raise OSError(f"{LIB_PATH} does not export {name}")
        _declare(l)
        if l.mbr_version() != _abi.MBR_ABI_VERSION:
            raise OSError("libmbr.so ABI version mismatch")
        offs = (C.c_size_t * 8)()
        size = l.mbr_scene_layout(offs)
        py_size, py_offs = _abi.scene_layout_py()
        if size != py_size or list(offs) != py_offs:
            raise OSError(f"mbr_scene layout mismatch: C {size} {list(offs)} vs ctypes {py_size} {py_offs}")
        _lib = l
    return _lib


def check(rc, ctx=None):
    if rc != _abi.MBR_OK:
        msg = ""
        if ctx:
            m = lib().mbr_last_error(ctx)
            msg = m.decode() if m else ""
        raise MbrError(rc, msg)

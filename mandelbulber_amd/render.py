"""Renderer: the binding of mbr_init / mbr_set_scene / mbr_render[_device]."""
import ctypes as C

import numpy as np

from . import _abi
from ._lib import check, lib


class Frame:
    """Host planes in cImage layout (src/cimage.hpp:157-162), one array per plane."""

    def __init__(self, w, h):
        self.w, self.h = w, h
        self.rgb = np.zeros((3, h, w), np.float32)
        self.z = np.zeros((h, w), np.float32)
        self.colour = np.zeros((3, h, w), np.uint8)
        self.alpha = np.zeros((h, w), np.uint16)
        self.opacity = np.zeros((h, w), np.uint16)

    def fb(self):
        f = _abi.Framebuffer()
        f.width, f.height = self.w, self.h
        for c in range(3):
            f.rgb[c] = self.rgb[c].ctypes.data
            f.colour[c] = self.colour[c].ctypes.data
        f.z = self.z.ctypes.data
        f.alpha = self.alpha.ctypes.data
        f.opacity = self.opacity.ctypes.data
        return f


def device_framebuffer(planes, w, h):
    """mbr_framebuffer over device tensors: dict with keys rgb (3,h,w) f32,
    z (h,w) f32, colour (3,h,w) u8, alpha/opacity (h,w) int16 storage."""
    f = _abi.Framebuffer()
    f.width, f.height = w, h
    for c in range(3):
        f.rgb[c] = planes["rgb"][c].data_ptr() if "rgb" in planes else None
        f.colour[c] = planes["colour"][c].data_ptr() if "colour" in planes else None
    f.z = planes["z"].data_ptr() if "z" in planes else None
    f.alpha = planes["alpha"].data_ptr() if "alpha" in planes else None
    f.opacity = planes["opacity"].data_ptr() if "opacity" in planes else None
    return f


class Renderer:
    """One render context (mbr_ctx) on one HIP device."""

    def __init__(self, device=0):
        self._ctx = C.c_void_p()
        opts = _abi.DeviceOpts(device)
        check(lib().mbr_init(C.byref(opts), C.byref(self._ctx)))
        self.scene = None

    def close(self):
        if self._ctx:
            lib().mbr_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scene(self, scene, lightmap=None):
        tex = None
        if lightmap is not None:
            self._lm = np.ascontiguousarray(lightmap, dtype=np.uint8)
            tex = _abi.Texture(self._lm.shape[1], self._lm.shape[0],
                               self._lm.ctypes.data_as(C.POINTER(C.c_uint8)))
        check(lib().mbr_set_scene(self._ctx, C.byref(scene), C.byref(tex) if tex is not None else None),
              self._ctx)
        self.scene = scene

    def lights(self):
        """The context's aux light table (list of mbr_light records)."""
        n = C.c_int()
        check(lib().mbr_get_lights(self._ctx, None, 0, C.byref(n)), self._ctx)
        tab = (_abi.Light * max(n.value, 1))()
        check(lib().mbr_get_lights(self._ctx, tab, n.value, C.byref(n)), self._ctx)
        return list(tab)[:n.value]

    def set_lights(self, lights):
        """Use `lights` (mbr_light records, or None for automatic placement)."""
        if not lights:
            check(lib().mbr_set_lights(self._ctx, None, 0), self._ctx)
            return
        tab = (_abi.Light * len(lights))(*lights)
        check(lib().mbr_set_lights(self._ctx, tab, len(lights)), self._ctx)

    def _opts(self, precision, rows, stream):
        o = _abi.RenderOpts()
        o.precision = precision
        o.stream = stream
        o.row_begin, o.row_end = rows if rows is not None else (0, 0)
        return o

    def set_textures(self, background=None, envmap=None):
        """mbr_set_textures: the background and environment-map textures,
        each an (h, w, 3) uint8 array or None (before set_scene)."""
        keep, ptrs = [], []
        for t in (background, envmap):
            if t is None:
                ptrs.append(None)
                continue
            a = np.ascontiguousarray(t, dtype=np.uint8)
            tex = _abi.Texture(a.shape[1], a.shape[0], a.ctypes.data_as(C.POINTER(C.c_uint8)))
            keep.append((a, tex))
            ptrs.append(C.byref(tex))
        check(lib().mbr_set_textures(self._ctx, ptrs[0], ptrs[1]), self._ctx)

    def cancel(self):
        """mbr_cancel: stop the render in progress on this context (from
        another thread); it then raises MbrError(MBR_ERR_CANCELLED)."""
        check(lib().mbr_cancel(self._ctx), self._ctx)

    def render(self, precision=_abi.PREC_FP32_FIXUP, rows=None):
        """Render into host planes; returns (Frame, Stats)."""
        s = self.scene
        frame = Frame(s.image_width, s.image_height)
        fb = frame.fb()
        st = _abi.Stats()
        check(lib().mbr_render(self._ctx, C.byref(self._opts(precision, rows, None)), C.byref(fb), C.byref(st)),
              self._ctx)
        return frame, st

    def render_batch(self, scenes, planes_list, precision=_abi.PREC_FP32_FIXUP, slots=0, blocks_per_cu=0,
                     autotune=False, replays=1, stream=None, lightmap=None):
        """Render many scenes into device planes (one dict per scene, see
        device_framebuffer) as one captured HIP graph; returns
        (Stats summed over the frames, BatchResult)."""
        n = len(scenes)
        arr = (_abi.Scene * n)()
        for i, s in enumerate(scenes):
            C.memmove(C.byref(arr[i]), C.byref(s), C.sizeof(_abi.Scene))
        fbs = (_abi.Framebuffer * n)()
        for i, (s, p) in enumerate(zip(scenes, planes_list)):
            fbs[i] = device_framebuffer(p, s.image_width, s.image_height)
        tex = None
        if lightmap is not None:
            self._lm = np.ascontiguousarray(lightmap, dtype=np.uint8)
            tex = _abi.Texture(self._lm.shape[1], self._lm.shape[0],
                               self._lm.ctypes.data_as(C.POINTER(C.c_uint8)))
        o = _abi.BatchOpts(precision, stream, slots, blocks_per_cu, 1 if autotune else 0, replays)
        st = _abi.Stats()
        res = _abi.BatchResult()
        check(lib().mbr_render_batch(self._ctx, arr, n, C.byref(tex) if tex is not None else None, fbs, C.byref(o),
                                     C.byref(st), C.byref(res)), self._ctx)
        return st, res

    def adjoint(self, grad_rgb, stream=None, skipped=None):
        """dL/dtheta (9 floats: power, vp x/y/z, alpha, beta, gamma, light
        alpha, light beta) of the current scene for the image adjoint
        grad_rgb: a contiguous (3, H, W) float32 device tensor dL/dC.
        skipped: optional (H, W) uint8 device tensor, set to 1 on the
        grazing hit pixels the adjoint leaves out."""
        s = self.scene
        assert tuple(grad_rgb.shape) == (3, s.image_height, s.image_width) and grad_rgb.is_contiguous()
        planes = (C.c_void_p * 3)(*[grad_rgb[c].data_ptr() for c in range(3)])
        out = (C.c_double * 9)()
        st = _abi.Stats()
        sk = None
        if skipped is not None:
            assert tuple(skipped.shape) == (s.image_height, s.image_width) and skipped.is_contiguous()
            sk = skipped.data_ptr()
        check(lib().mbr_adjoint(self._ctx, C.byref(self._opts(_abi.PREC_FP64, None, stream)), planes, out, sk,
                                C.byref(st)), self._ctx)
        return list(out), st

    def adjoint_jacobian(self, jac, stream=None, skipped=None):
        """Per-pixel tangents dC/dtheta of the current scene into jac, a
        contiguous (9, 3, H, W) float32 device tensor (theta as for
        adjoint); skipped as for adjoint."""
        s = self.scene
        assert tuple(jac.shape) == (9, 3, s.image_height, s.image_width) and jac.is_contiguous()
        assert str(jac.dtype) == "torch.float32"
        st = _abi.Stats()
        sk = None
        if skipped is not None:
            assert tuple(skipped.shape) == (s.image_height, s.image_width) and skipped.is_contiguous()
            sk = skipped.data_ptr()
        check(lib().mbr_adjoint_jacobian(self._ctx, C.byref(self._opts(_abi.PREC_FP64, None, stream)),
                                         jac.data_ptr(), sk, C.byref(st)), self._ctx)
        return jac, st

    def post_ssao(self, scene, z, opacity, colour, image16, stream=None):
        """PostRendering_SSAO on one frame.  numpy planes: host call, image16
        (H, W, 3) uint16 is updated in place and returned; torch CUDA tensors:
        queued on `stream` (default: torch's current stream)."""
        h, w = z.shape
        on_device = hasattr(z, "data_ptr")
        for a in (z, opacity, image16) + tuple(colour[i] for i in range(3)):
            assert (a.is_contiguous() if on_device else a.flags["C_CONTIGUOUS"])
        assert tuple(image16.shape) == (h, w, 3) and tuple(opacity.shape) == (h, w)
        if on_device and stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        p = post_planes(w, h, z, opacity, colour, image16)
        check(lib().mbr_post_ssao(self._ctx, C.byref(scene), C.byref(p), 1 if on_device else 0,
                                  stream if on_device else None), self._ctx)
        return image16

    def compile_image_device(self, scene, rgb, image16, stream=None):
        """CompileImage on the GPU: rgb (3, H, W) float32 and image16 (H, W, 3)
        int16/uint16 CUDA tensors; queued on `stream` (default: torch's)."""
        _, h, w = rgb.shape
        assert rgb.is_contiguous() and image16.is_contiguous() and tuple(image16.shape) == (h, w, 3)
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        planes = (C.c_void_p * 3)(*[rgb[c].data_ptr() for c in range(3)])
        check(lib().mbr_compile_image_device(self._ctx, C.byref(scene), w, h, planes, image16.data_ptr(), stream),
              self._ctx)
        return image16

    def post_dof(self, scene, z, image16, alpha=None, stream=None):
        """PostRendering_DOF on one frame (numpy: host call, image16 and alpha
        updated in place; torch CUDA tensors: queued on `stream`)."""
        h, w = z.shape
        on_device = hasattr(z, "data_ptr")
        assert tuple(image16.shape) == (h, w, 3)
        if on_device and stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        p = post_planes(w, h, z, None, [None, None, None], image16, alpha)
        check(lib().mbr_post_dof(self._ctx, C.byref(scene), C.byref(p), 1 if on_device else 0,
                                 stream if on_device else None), self._ctx)
        return image16, alpha

    def render_device(self, planes, precision=_abi.PREC_FP32_FIXUP, rows=None, stream=None):
        """Render into device planes (see device_framebuffer); returns Stats."""
        s = self.scene
        fb = device_framebuffer(planes, s.image_width, s.image_height)
        st = _abi.Stats()
        check(lib().mbr_render_device(self._ctx, C.byref(self._opts(precision, rows, stream)), C.byref(fb),
                                      C.byref(st)), self._ctx)
        return st

    def render_device_noalloc(self, planes, precision=_abi.PREC_FP32_FIXUP, rows=None, stream=None):
        """render_device through mbr_render_device_noalloc (no device
        allocation: the library's static scratch array); returns Stats."""
        s = self.scene
        fb = device_framebuffer(planes, s.image_width, s.image_height)
        st = _abi.Stats()
        check(lib().mbr_render_device_noalloc(self._ctx, C.byref(self._opts(precision, rows, stream)), C.byref(fb),
                                              C.byref(st)), self._ctx)
        return st


def post_planes(w, h, z, opacity, colour, image16, alpha=None):
    """mbr_post_planes over numpy arrays (host) or torch tensors (device):
    z (H, W) float32, opacity (H, W) uint16/int16, colour (3, H, W) uint8,
    image16 (H, W, 3) uint16/int16, alpha (H, W) or None."""
    def ptr(a):
        if a is None:
            return None
        return a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
    p = _abi.PostPlanes()
    p.width, p.height = w, h
    p.z, p.opacity, p.image16, p.alpha = ptr(z), ptr(opacity), ptr(image16), ptr(alpha)
    for i in range(3):
        p.colour[i] = ptr(colour[i])
    return p


def compile_image(scene, rgb):
    """CompileImage: tone-map float planes (3, H, W) with the scene's
    brightness, contrast, hdr and gamma into 16-bit RGB (H, W, 3)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    _, h, w = rgb.shape
    out = np.empty((h, w, 3), np.uint16)
    planes = (C.c_void_p * 3)(*[rgb[c].ctypes.data for c in range(3)])
    check(lib().mbr_compile_image(C.byref(scene), w, h, planes, out.ctypes.data))
    return out


def write_png(path, image16, alpha=None, bits=16):
    """Write (H, W, 3) uint16 (and an optional (H, W) uint16 alpha) as PNG."""
    image16 = np.ascontiguousarray(image16, dtype=np.uint16)
    h, w, _ = image16.shape
    a = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.uint16)
    check(lib().mbr_write_png(path.encode(), w, h, image16.ctypes.data, None if a is None else a.ctypes.data, bits))

"""The C ABI library loads, exports every entry point include/mbr.h
declares, and reports errors without a device (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

from mandelbulber_amd import _abi, _lib, lib, scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mbr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mbr_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 12
    l = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(l, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_version_and_layout():
    assert lib().mbr_version() == _abi.MBR_ABI_VERSION


def test_init_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    ctx = C.c_void_p()
    rc = lib().mbr_init(C.byref(_abi.DeviceOpts(0)), C.byref(ctx))
    assert rc == -2  # MBR_ERR_NO_DEVICE: the caller keeps its CPU path
    assert not ctx.value


def test_null_arguments_rejected():
    assert lib().mbr_scene_defaults(None) == -1
    assert lib().mbr_set_scene(None, None, None) == -1
    assert lib().mbr_render(None, None, None, None) == -1
    assert lib().mbr_cancel(None) == -1

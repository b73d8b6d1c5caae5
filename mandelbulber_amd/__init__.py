"""MI355X-native render core for Mandelbulber 1.21 scenes.

The render path (distance estimation, ray marching, shading) runs in the HIP
library libmbr.so behind the C ABI of include/mbr.h; this package is the
Python binding of that ABI.  Importing it loads the library and fails loudly
when it is missing.
"""
from . import _abi
from ._lib import MbrError, lib
from . import scene
from .render import Renderer, Frame

__all__ = ["_abi", "lib", "MbrError", "scene", "Renderer", "Frame"]

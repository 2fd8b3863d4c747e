"""MI355X-native wavelet potential recovery (HIP kernels for gfx950 behind a C ABI).

The compute path is libwhhd.so (see include/whhd.h); this package loads it with
ctypes and mirrors the reference's C++ classes in :mod:`.api`.
"""
from ._lib import WhhdError, lib  # noqa: F401
from .api import *  # noqa: F401,F403

__version__ = "0.1.0"

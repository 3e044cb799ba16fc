"""bgb -- MI355X-native B2SR bit-packed binary SpMV (host package).

The compute lives in libbgb.so (HIP kernels for gfx950 behind the C ABI in
include/bgb.h); this package loads it with ctypes and offers numpy / torch
conveniences around it.
"""
from ._lib import BgbError, LIB_PATH, HEADER_PATH, call, lib  # noqa: F401
from . import graph  # noqa: F401

__all__ = ["BgbError", "LIB_PATH", "HEADER_PATH", "call", "lib", "graph"]


def __getattr__(name):
    # the device module imports torch; load it lazily
    if name == "b2sr":
        import importlib
        return importlib.import_module(".b2sr", __name__)
    raise AttributeError(name)

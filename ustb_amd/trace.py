"""roctx markers and ranges for the serving loop (rocprofv3 --marker-trace).

The ROCm profiler's roctx library (librocprofiler-sdk-roctx) is loaded on
first use; where it is absent every call is a no-op, so the markers cost
nothing outside a profiled run beyond a ctypes call.
"""
from __future__ import annotations

import ctypes
import os

_lib = None
_tried = False


def _roctx():
    global _lib, _tried
    if not _tried:
        _tried = True
        for name in (os.environ.get("USTBX_ROCTX_LIB"), "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1",
                     "librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/libroctx64.so.4"):
            if not name:
                continue
            try:
                lib = ctypes.CDLL(name)
                lib.roctxMarkA.argtypes = [ctypes.c_char_p]
                lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                lib.roctxRangePushA.restype = ctypes.c_int
                lib.roctxRangePop.restype = ctypes.c_int
                _lib = lib
                break
            except (OSError, AttributeError):
                continue
    return _lib


def mark(msg: str) -> None:
    lib = _roctx()
    if lib is not None:
        lib.roctxMarkA(msg.encode())


class range_:
    """with range_("name"): ... -- a roctx range around the block."""

    def __init__(self, msg: str):
        self.msg = msg

    def __enter__(self):
        lib = _roctx()
        if lib is not None:
            lib.roctxRangePushA(self.msg.encode())
        return self

    def __exit__(self, *exc):
        lib = _roctx()
        if lib is not None:
            lib.roctxRangePop()
        return False

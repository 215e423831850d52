"""MI355X-native implementation of Stream's PFEM mesh-and-solve path.

The compute lives in one HIP library, ``lib/libstream_amd.so`` (built by
``python stream_amd/build.py``).  The reference's three library names
(``libgeometry``, ``libmesh``, ``libnumerics``) are aliases of it, loaded
here the way the reference's ``python/__init__.py:26-42`` loads its own,
with ``numpy.ctypeslib.load_library``.

Importing this package without the built library raises: there is no CPU
fallback for any operation.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIBDIR = Path(__file__).resolve().parent / "lib"

if not (LIBDIR / "libstream_amd.so").exists():
    raise ImportError(
        f"stream_amd native library missing under {LIBDIR}; build it with 'python stream_amd/build.py'"
    )

libgeo = np.ctypeslib.load_library("libgeometry", str(LIBDIR))
libnum = np.ctypeslib.load_library("libnumerics", str(LIBDIR))
libmesh = np.ctypeslib.load_library("libmesh", str(LIBDIR))
lib = np.ctypeslib.load_library("libstream_amd", str(LIBDIR))

lib.Stream_last_error.restype = ctypes.c_char_p
lib.Stream_last_error.argtypes = []
lib.Stream_device_count.restype = ctypes.c_int32
lib.Stream_device_count.argtypes = []
lib.Stream_synchronize.restype = ctypes.c_int32
lib.Stream_synchronize.argtypes = []
lib.Stream_tune.restype = ctypes.c_int32
lib.Stream_tune.argtypes = [ctypes.c_char_p, ctypes.c_double]
lib.Stream_device_bytes.restype = ctypes.c_int32
lib.Stream_device_bytes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                    ctypes.c_void_p]


class StreamError(RuntimeError):
    """A native entry point returned a non-zero status."""


def check(status: int) -> None:
    """Raise StreamError carrying the library's last message on failure."""
    if status != 0:
        msg = lib.Stream_last_error()
        raise StreamError(f"status {status}: {msg.decode() if msg else ''}")


def device_count() -> int:
    return int(lib.Stream_device_count())


def synchronize() -> None:
    check(lib.Stream_synchronize())


def tune(key: str, value: float) -> None:
    """Set a solver knob: 'replace_factor', 'batch', 'cg_blocks', 'graph',
    'xcd_map', 'fold_group', 'sell_chunk' or 'llt_threads'."""
    check(lib.Stream_tune(key.encode(), float(value)))


def device_bytes(reset_peak: bool = False) -> dict:
    """Device bytes held by the library's buffers (now and high-water mark)
    and the device's free / total memory."""
    import ctypes

    v = [ctypes.c_uint64(0) for _ in range(4)]
    check(lib.Stream_device_bytes(ctypes.byref(v[0]), ctypes.byref(v[1]), int(reset_peak), ctypes.byref(v[2]),
                                  ctypes.byref(v[3])))
    return dict(current=v[0].value, peak=v[1].value, free=v[2].value, total=v[3].value)


from . import geometry, mesh, numerics, pfem  # noqa: E402,F401

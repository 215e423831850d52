"""ctypes helpers shared by the binding modules."""
from __future__ import annotations

import ctypes

import numpy as np

u32 = ctypes.c_uint32
i32 = ctypes.c_int32
u64 = ctypes.c_uint64
f64 = ctypes.c_double
vp = ctypes.c_void_p


def arr(dtype, ndim=None, writeable=False):
    flags = ["C", "ALIGNED"] + (["WRITEABLE"] if writeable else [])
    return np.ctypeslib.ndpointer(dtype=dtype, ndim=ndim, flags=flags)


def opt(dtype, writeable=True):
    """Nullable array argument: pass None or a contiguous array."""

    class _Opt:
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            a = obj
            if not isinstance(a, np.ndarray) or not a.flags.c_contiguous:
                raise TypeError("expected a C-contiguous numpy array")
            if a.dtype != np.dtype(dtype):
                raise TypeError(f"expected {np.dtype(dtype)}, got {a.dtype}")
            if writeable and not a.flags.writeable:
                raise TypeError("array must be writeable")
            return a.ctypes.data_as(ctypes.c_void_p)

    return _Opt


def bind(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn

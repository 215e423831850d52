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


def node_arrays(nodes, dim, dtype, *per_node):
    """C-contiguous copies of an (n, dim) node array and of per-node arrays
    of length n, each checked for its shape before any of them reaches the
    native library (which reads n rows of each)."""
    nodes = np.ascontiguousarray(nodes, dtype=dtype)
    if nodes.ndim != 2 or nodes.shape[1] != dim:
        raise ValueError(f"nodes must have shape (n, {dim}), got {nodes.shape}")
    n = nodes.shape[0]
    out = [nodes]
    for a in per_node:
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.shape != (n,):
            raise ValueError(f"per-node array must have shape ({n},), got {a.shape}")
        out.append(a)
    return out


def permute_field(fn, handle, field, n):
    """Call a native permute_forward / permute_backward entry point on a
    field of n nodes (shape (n, ...), any dtype) and return the permuted
    copy; each node's element is field.itemsize * prod(shape[1:]) bytes."""
    f = np.ascontiguousarray(field)
    if f.ndim < 1 or f.shape[0] != n:
        raise ValueError(f"field must have {n} rows, got shape {f.shape}")
    out = np.empty_like(f)
    eb = f.itemsize * int(np.prod(f.shape[1:], dtype=np.int64))
    if n and eb:
        rc = fn(handle, f.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), eb)
        from . import check

        check(rc)
    return out

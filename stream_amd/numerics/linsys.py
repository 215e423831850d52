"""LinSys, the reference's python/numerics/linsys.py surface (float32 b),
plus float64 right-hand sides."""
from __future__ import annotations

import ctypes

import numpy as np

from .. import check, libnum
from .._ct import arr, bind, i32
from .csr import DeviceCSR, d_csr_ptr

linsys_ptr = ctypes.c_void_p
linsys_create = bind(libnum, "LinSys_create", linsys_ptr, [])
linsys_destroy = bind(libnum, "LinSys_destroy", None, [linsys_ptr])
linsys_set = bind(libnum, "LinSys_set", None, [linsys_ptr, d_csr_ptr, arr(np.float32, 1)])
linsys_set_f64 = bind(libnum, "LinSys_set_f64", i32, [linsys_ptr, d_csr_ptr, arr(np.float64, 1)])


class LinSys:
    def __init__(self, A: DeviceCSR, b):
        self.sys = linsys_create()
        self.A = A  # keeps the shared CSR alive
        b = np.asarray(b)
        if b.dtype == np.float64:
            check(linsys_set_f64(self.sys, A.dcsr, np.ascontiguousarray(b)))
        else:
            linsys_set(self.sys, A.dcsr, np.ascontiguousarray(b, np.float32))

    def __del__(self):
        if getattr(self, "sys", None):
            linsys_destroy(self.sys)
            self.sys = None

"""HostCSR / DeviceCSR, the reference's python/numerics/csr.py:60-95 surface.

Extension: float64 values are kept as float64 (the reference casts to
float32)."""
from __future__ import annotations

import ctypes

import numpy as np
import numpy.typing as npt

from .. import StreamError, lib, libnum
from .._ct import arr, bind, u32

d_csr_ptr = ctypes.c_void_p
h_csr_ptr = ctypes.c_void_p

d_csr_create = bind(libnum, "d_csr_create", d_csr_ptr, [])
d_csr_destroy = bind(libnum, "d_csr_destroy", None, [d_csr_ptr])
h_csr_create = bind(libnum, "h_csr_create", h_csr_ptr,
                    [u32, arr(np.uint32, 1), arr(np.uint32, 1), arr(np.float32, 1)])
h_csr_create_f64 = bind(libnum, "h_csr_create_f64", h_csr_ptr,
                        [u32, arr(np.uint32, 1), arr(np.uint32, 1), arr(np.float64, 1)])
h_csr_destroy = bind(libnum, "h_csr_destroy", None, [h_csr_ptr])
d_csr_h2d = bind(libnum, "d_csr_h2d", None, [d_csr_ptr, h_csr_ptr])


class HostCSR:
    def __init__(self, row: npt.NDArray[np.uint32], col: npt.NDArray[np.uint32], val: npt.NDArray):
        self.n = len(row) - 1
        self.nnz = row[self.n]
        self.dtype = np.float64 if np.asarray(val).dtype == np.float64 else np.float32
        r = np.ascontiguousarray(row, np.uint32)
        c = np.ascontiguousarray(col, np.uint32)
        if self.dtype == np.float64:
            self.hcsr = h_csr_create_f64(self.n, r, c, np.ascontiguousarray(val, np.float64))
        else:
            self.hcsr = h_csr_create(self.n, r, c, np.ascontiguousarray(val, np.float32))

    def __del__(self):
        if getattr(self, "hcsr", None):
            h_csr_destroy(self.hcsr)
            self.hcsr = None


class DeviceCSR:
    def __init__(self, host_csr: HostCSR):
        self.n = host_csr.n
        self.dcsr = d_csr_create()
        if not self.dcsr:
            raise StreamError(lib.Stream_last_error().decode())
        self.set_from_host(host_csr)

    def __del__(self):
        if getattr(self, "dcsr", None):
            d_csr_destroy(self.dcsr)
            self.dcsr = None

    def set_from_host(self, host_csr: HostCSR):
        d_csr_h2d(self.dcsr, host_csr.hcsr)

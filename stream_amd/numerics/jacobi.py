"""PrecJacobi, the reference's python/numerics/jacobi.py surface."""
from __future__ import annotations

import ctypes

from .. import StreamError, lib, libnum
from .._ct import bind, u32
from .csr import DeviceCSR, d_csr_ptr

prec_jacobi_ptr = ctypes.c_void_p
prec_jacobi_create = bind(libnum, "PrecJacobi_create", prec_jacobi_ptr, [d_csr_ptr])
prec_jacobi_destroy = bind(libnum, "PrecJacobi_destroy", None, [prec_jacobi_ptr])
prec_jacobi_missing = bind(libnum, "PrecJacobi_missing_diagonals", u32, [prec_jacobi_ptr])


class PrecJacobi:
    def __init__(self, A: DeviceCSR):
        self.A = A
        self.prec = prec_jacobi_create(A.dcsr)
        if not self.prec:
            raise StreamError(lib.Stream_last_error().decode())

    def __del__(self):
        if getattr(self, "prec", None):
            prec_jacobi_destroy(self.prec)
            self.prec = None

    def missing_diagonals(self) -> int:
        """Rows without a usable diagonal (applied as identity)."""
        return int(prec_jacobi_missing(self.prec))

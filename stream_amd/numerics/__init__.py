"""Numerics bindings: the reference's ``stream.numerics`` (python/numerics/__init__.py:20-30)."""
from . import cg, csr, jacobi, linsys
from .cg import SolverCG
from .csr import DeviceCSR, HostCSR
from .jacobi import PrecJacobi
from .linsys import LinSys

__all__ = ["HostCSR", "DeviceCSR", "LinSys", "PrecJacobi", "SolverCG", "csr", "linsys", "jacobi", "cg"]

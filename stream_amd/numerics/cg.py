"""SolverCG, the reference's python/numerics/cg.py:73-99 surface.

``solve_jacobi`` / ``solve`` keep the reference semantics (float32, stop on
sqrt(r.z) or sqrt(r.r) <= 1e-6, at most 10000 iterations, the reference's
iteration count).  Extensions: ``solve_rel`` (FP64 Jacobi-PCG to a true
relative residual) and ``solve_mixed`` (FP32 inner iterations with FP64
residual replacement); both return (iterations, true relative residual).
"""
from __future__ import annotations

import ctypes

import numpy as np
import numpy.typing as npt

from .. import StreamError, check, lib, libnum
from .._ct import arr, bind, f64, i32, u32
from .jacobi import PrecJacobi, prec_jacobi_ptr
from .linsys import LinSys, linsys_ptr

solver_cg_ptr = ctypes.c_void_p
solver_cg_create = bind(libnum, "SolverCG_create", solver_cg_ptr, [])
solver_cg_destroy = bind(libnum, "SolverCG_destroy", None, [solver_cg_ptr])
solver_cg_jacobi_solve = bind(libnum, "SolverCG_jacobi_solve", u32,
                              [solver_cg_ptr, linsys_ptr, prec_jacobi_ptr, arr(np.float32, 1, writeable=True)])
solver_cg_solve = bind(libnum, "SolverCG_solve", u32, [solver_cg_ptr, linsys_ptr, arr(np.float32, 1, writeable=True)])
solver_cg_solve_rel_f64 = bind(libnum, "SolverCG_solve_rel_f64", i32,
                               [solver_cg_ptr, linsys_ptr, prec_jacobi_ptr, arr(np.float64, 1, writeable=True), f64,
                                u32, arr(np.float64, 1, writeable=True)])
solver_cg_solve_mixed = bind(libnum, "SolverCG_solve_mixed", i32,
                             [solver_cg_ptr, linsys_ptr, prec_jacobi_ptr, arr(np.float64, 1, writeable=True), f64,
                              u32, arr(np.float64, 1, writeable=True)])


class SolverCG:
    def __init__(self):
        self.solver = solver_cg_create()
        if not self.solver:
            raise StreamError(lib.Stream_last_error().decode())

    def __del__(self):
        if getattr(self, "solver", None):
            solver_cg_destroy(self.solver)
            self.solver = None

    def solve_jacobi(self, sys: LinSys, prec: PrecJacobi, x: npt.NDArray[np.float32]):
        if type(prec) is not PrecJacobi:
            raise TypeError(f"SolverCG.solve_jacobi expects a 'PrecJacobi' preconditioner, not a '{type(prec)}'")
        sol = np.copy(x).astype(np.float32)
        niter = solver_cg_jacobi_solve(self.solver, sys.sys, prec.prec, sol)
        x[:] = sol[:]
        return niter

    def solve(self, sys: LinSys, x: npt.NDArray[np.float32]):
        sol = np.copy(x).astype(np.float32)
        niter = solver_cg_solve(self.solver, sys.sys, sol)
        x[:] = sol[:]
        return niter

    # ---- extensions ----
    def _run(self, fn, sys, prec, x, rtol, maxit):
        sol = np.ascontiguousarray(np.array(x, np.float64))
        stats = np.zeros(4)
        st = fn(self.solver, sys.sys, prec.prec, sol, rtol, maxit, stats)
        x[:] = sol[:]
        self.last_stats = dict(iters=int(stats[0]), true_rel=float(stats[1]), ms=float(stats[2]),
                               replacements=int(stats[3]))
        check(st)
        return int(stats[0]), float(stats[1])

    def solve_rel(self, sys: LinSys, prec: PrecJacobi, x, rtol=1e-8, maxit=200000):
        return self._run(solver_cg_solve_rel_f64, sys, prec, x, rtol, maxit)

    def solve_mixed(self, sys: LinSys, prec: PrecJacobi, x, rtol=1e-8, maxit=200000):
        return self._run(solver_cg_solve_mixed, sys, prec, x, rtol, maxit)

"""The PFEM mesh-and-solve step on device-resident points: LBVH, 2D
alpha-shape, P1 pressure assembly and CG to a true FP64 relative residual.
No reference counterpart (the reference never connects mesh and solve)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import StreamError, check, lib
from ._ct import arr, bind, f64, i32, node_arrays, opt, permute_field, u32

_p = ctypes.c_void_p
PfemStep2D_create = bind(lib, "PfemStep2D_create", _p, [])
PfemStep2D_destroy = bind(lib, "PfemStep2D_destroy", None, [_p])
PfemStep2D_set_points = bind(lib, "PfemStep2D_set_points", i32,
                             [_p, u32, arr(np.float64, 2), arr(np.float64, 1), arr(np.float64, 1)])
PfemStep2D_set_morton_bits = bind(lib, "PfemStep2D_set_morton_bits", i32, [_p, i32])
PfemStep2D_run = bind(lib, "PfemStep2D_run", i32, [_p, i32, f64, u32, arr(np.float64, 1, writeable=True)])
PfemStep2D_get_solution = bind(lib, "PfemStep2D_get_solution", i32, [_p, arr(np.float64, 1, writeable=True)])
PfemStep2D_get_solution_morton = bind(lib, "PfemStep2D_get_solution_morton", i32,
                                      [_p, arr(np.float64, 1, writeable=True)])
PfemStep2D_nnz = bind(lib, "PfemStep2D_nnz", u32, [_p])
PfemStep2D_get_system = bind(lib, "PfemStep2D_get_system", i32,
                             [_p, opt(np.uint32), opt(np.uint32), opt(np.float64), opt(np.float64)])
PfemStep2D_get_mesh = bind(lib, "PfemStep2D_get_mesh", i32,
                           [_p, opt(np.uint32), opt(np.uint8), opt(np.float64), opt(np.uint32)])

PfemStep2D_assemble = bind(lib, "PfemStep2D_assemble", i32, [_p, opt(np.float64)])
PfemStep2D_set_source_modes = bind(lib, "PfemStep2D_set_source_modes", i32, [_p, u32, opt(np.float64, False)])
PfemStep2D_advect = bind(lib, "PfemStep2D_advect", i32, [_p, u32, opt(np.float64, False), f64, u32])
PfemStep2D_get_points = bind(lib, "PfemStep2D_get_points", i32, [_p, arr(np.float64, 2, writeable=True)])
PfemStep2D_set_learned = bind(lib, "PfemStep2D_set_learned", i32, [_p, _p])

PfemStep2D_time_dir_kernel = bind(lib, "PfemStep2D_time_dir_kernel", i32,
                                  [_p, i32, u32, ctypes.POINTER(ctypes.c_double)])
PfemStep2D_permute_forward = bind(lib, "PfemStep2D_permute_forward", i32, [_p, _p, _p, u32])
PfemStep2D_permute_backward = bind(lib, "PfemStep2D_permute_backward", i32, [_p, _p, _p, u32])
PfemStep3D_permute_forward = bind(lib, "PfemStep3D_permute_forward", i32, [_p, _p, _p, u32])
PfemStep3D_permute_backward = bind(lib, "PfemStep3D_permute_backward", i32, [_p, _p, _p, u32])
PfemStep2D_set_velocity = bind(lib, "PfemStep2D_set_velocity", i32, [_p, arr(np.float64, 2)])
PfemStep2D_get_velocity = bind(lib, "PfemStep2D_get_velocity", i32, [_p, arr(np.float64, 2, True)])
PfemStep2D_projection_step = bind(lib, "PfemStep2D_projection_step", i32,
                                  [_p, i32, f64, f64, f64, arr(np.float64, 1), f64, u32, i32,
                                   arr(np.float64, 1, True)])
PfemStep2D_get_flow_fields = bind(lib, "PfemStep2D_get_flow_fields", i32,
                                  [_p, opt(np.float64), opt(np.float64), opt(np.float64), opt(np.float64)])
PfemStep2D_time_sections = bind(lib, "PfemStep2D_time_sections", i32, [_p, u32, arr(np.float64, 1, True)])

# stage names of PfemStep2D.time_sections, in launch order
SECTION_NAMES = ("morton", "sort", "gather", "karras", "refit", "fill", "delaunay", "scans", "compress", "p1",
                 "mesh_and_assembly")

SOLVER_FP64_JACOBI = 0
SOLVER_MIXED_JACOBI = 1
SOLVER_MIXED_FSAI = 2
SOLVER_MIXED_LEARNED = 3
SOLVER_NAMES = {"fp64": SOLVER_FP64_JACOBI, "mixed": SOLVER_MIXED_JACOBI, "fsai": SOLVER_MIXED_FSAI,
                "learned": SOLVER_MIXED_LEARNED}


class PfemStep2D:
    def __init__(self, xy, alpha, f, morton_bits: int = 32):
        self.h = PfemStep2D_create()
        if not self.h:
            raise StreamError(lib.Stream_last_error().decode())
        xy, alpha, f = node_arrays(xy, 2, np.float64, alpha, f)
        self.n = xy.shape[0]
        check(PfemStep2D_set_morton_bits(self.h, morton_bits))
        check(PfemStep2D_set_points(self.h, self.n, xy, alpha, f))

    def __del__(self):
        if getattr(self, "h", None):
            PfemStep2D_destroy(self.h)
            self.h = None

    def run(self, solver="fp64", rtol=1e-8, maxit=200000):
        st = np.zeros(10)
        kind = SOLVER_NAMES[solver] if isinstance(solver, str) else int(solver)
        check(PfemStep2D_run(self.h, kind, rtol, maxit, st))
        return dict(mesh_ms=st[0], assemble_ms=st[1], solve_ms=st[2], iters=int(st[3]), true_rel=st[4],
                    replacements=int(st[5]), nnz=int(st[6]), n_elems=int(st[7]), prep_ms=st[8],
                    fallback=bool(st[9]))

    def set_points(self, xy, alpha, f):
        """Replace the points (same count or a new one), alpha and host source."""
        xy, alpha, f = node_arrays(xy, 2, np.float64, alpha, f)
        self.n = xy.shape[0]
        check(PfemStep2D_set_source_modes(self.h, 0, None))
        check(PfemStep2D_set_points(self.h, self.n, xy, alpha, f))

    def assemble(self):
        """Mesh and assemble without solving; returns (mesh ms, assembly ms)."""
        st = np.zeros(2)
        check(PfemStep2D_assemble(self.h, st))
        return float(st[0]), float(st[1])

    def set_source_modes(self, modes):
        """Device source f from Fourier modes rows (kx, ky, phi, ax, ay)."""
        modes = np.ascontiguousarray(modes, np.float64).reshape(-1, 5)
        check(PfemStep2D_set_source_modes(self.h, modes.shape[0], modes.ravel() if modes.size else None))

    def advect(self, modes, dt, steps=1):
        """Forward-Euler advection in the stream-function field of modes rows (kx, ky, phi, a)."""
        modes = np.ascontiguousarray(modes, np.float64).reshape(-1, 4)
        check(PfemStep2D_advect(self.h, modes.shape[0], modes.ravel(), dt, steps))

    def points(self):
        xy = np.empty((self.n, 2))
        check(PfemStep2D_get_points(self.h, xy))
        return xy

    def set_learned(self, net):
        """Network used by solver "learned" (kept alive by this object)."""
        self._learned = net
        check(PfemStep2D_set_learned(self.h, net.h if net is not None else None))

    def nnz(self):
        return int(PfemStep2D_nnz(self.h))

    def time_dir_kernel(self, fp32: bool, reps: int = 200, llt: bool = False) -> float:
        """Average ms of a CG direction kernel on the current system: FP64
        Jacobi path, FP32 row-sum Jacobi path, or (llt) the FP32 L L^T path's
        kernel on a ready-made direction."""
        if llt and not fp32:
            raise ValueError("the L L^T direction kernel is FP32")
        ms = ctypes.c_double(0.0)
        check(PfemStep2D_time_dir_kernel(self.h, 2 if llt else int(fp32), reps, ctypes.byref(ms)))
        return ms.value

    def time_update_kernel(self, reps: int = 200, lower: bool = False) -> float:
        """Average ms of the L L^T update kernel (x += alpha d, r' = r - alpha
        q, w = L^T r') or, with lower, of the lower apply (z = L w, d_new =
        z + beta d_old), on the learned factor (FSAI without a network)."""
        ms = ctypes.c_double(0.0)
        check(PfemStep2D_time_dir_kernel(self.h, 4 if lower else 3, reps, ctypes.byref(ms)))
        return ms.value

    def set_velocity(self, uv):
        """Nodal velocity (n, 2) in input order for projection_step."""
        uv = np.ascontiguousarray(uv, np.float64)
        if uv.shape != (self.n, 2):
            raise ValueError(f"velocity must have shape ({self.n}, 2), got {uv.shape}")
        check(PfemStep2D_set_velocity(self.h, uv))

    def velocity(self):
        uv = np.empty((self.n, 2))
        check(PfemStep2D_get_velocity(self.h, uv))
        return uv

    def projection_step(self, solver="fp64", dt=1e-3, rho=1.0, nu=0.0, g=(0.0, 0.0), rtol=1e-8, maxit=200000,
                        move=True):
        """One fractional step of the PFEM velocity (predictor, pressure
        solve with p = 0 on the free surface, correction; then, with move,
        the points move by dt u).  Returns run()'s fields plus flow_ms."""
        st = np.zeros(11)
        kind = SOLVER_NAMES[solver] if isinstance(solver, str) else int(solver)
        check(PfemStep2D_projection_step(self.h, kind, dt, rho, nu, np.asarray(g, np.float64), rtol, maxit,
                                         int(bool(move)), st))
        return dict(mesh_ms=st[0], assemble_ms=st[1], solve_ms=st[2], iters=int(st[3]), true_rel=st[4],
                    replacements=int(st[5]), nnz=int(st[6]), n_elems=int(st[7]), prep_ms=st[8],
                    fallback=bool(st[9]), flow_ms=st[10])

    def flow_fields(self):
        """Last projection step, Morton order: (u* (n, 2), new velocity (n, 2),
        lumped mass (n,), D(new velocity) (n,))."""
        us = np.empty((self.n, 2))
        un = np.empty((self.n, 2))
        m = np.empty(self.n)
        d = np.empty(self.n)
        check(PfemStep2D_get_flow_fields(self.h, us, un, m, d))
        return us, un, m, d

    def time_sections(self, reps: int = 5) -> dict:
        """GPU ms of the mesh-and-assembly stages on the current points
        (averaged over reps runs), keyed by SECTION_NAMES."""
        ms = np.zeros(len(SECTION_NAMES))
        check(PfemStep2D_time_sections(self.h, reps, ms))
        return dict(zip(SECTION_NAMES, ms.tolist()))

    def permute_forward(self, field):
        """Nodal field (shape (n, ...)) from input order to the Morton order
        of the last mesh, system and solution: out[i] = field[perm[i]]."""
        return permute_field(PfemStep2D_permute_forward, self.h, field, self.n)

    def permute_backward(self, field):
        """Nodal field from the Morton order of the last mesh back to input
        order: out[perm[i]] = field[i]."""
        return permute_field(PfemStep2D_permute_backward, self.h, field, self.n)

    def solution(self, morton=False):
        p = np.empty(self.n)
        check((PfemStep2D_get_solution_morton if morton else PfemStep2D_get_solution)(self.h, p))
        return p

    def system(self):
        nnz = PfemStep2D_nnz(self.h)
        row = np.empty(self.n + 1, np.uint32)
        col = np.empty(nnz, np.uint32)
        val = np.empty(nnz)
        b = np.empty(self.n)
        check(PfemStep2D_get_system(self.h, row, col, val, b))
        return row, col, val, b

    def mesh(self, n_elems):
        elem = np.empty((n_elems, 3), np.uint32)
        bnd = np.empty(self.n, np.uint8)
        xy = np.empty((self.n, 2))
        perm = np.empty(self.n, np.uint32)
        check(PfemStep2D_get_mesh(self.h, elem, bnd, xy, perm))
        return elem, bnd, xy, perm


PfemStep3D_create = bind(lib, "PfemStep3D_create", _p, [])
PfemStep3D_destroy = bind(lib, "PfemStep3D_destroy", None, [_p])
PfemStep3D_set_points = bind(lib, "PfemStep3D_set_points", i32,
                             [_p, u32, arr(np.float64, 2), arr(np.float64, 1), arr(np.float64, 1)])
PfemStep3D_set_morton_bits = bind(lib, "PfemStep3D_set_morton_bits", i32, [_p, i32])
PfemStep3D_run = bind(lib, "PfemStep3D_run", i32, [_p, i32, f64, u32, arr(np.float64, 1, writeable=True)])
PfemStep3D_assemble = bind(lib, "PfemStep3D_assemble", i32, [_p, opt(np.float64)])
PfemStep3D_get_solution = bind(lib, "PfemStep3D_get_solution", i32, [_p, arr(np.float64, 1, writeable=True)])
PfemStep3D_nnz = bind(lib, "PfemStep3D_nnz", u32, [_p])
PfemStep3D_get_nelem = bind(lib, "PfemStep3D_get_nelem", u32, [_p])
PfemStep3D_get_system = bind(lib, "PfemStep3D_get_system", i32,
                             [_p, opt(np.uint32), opt(np.uint32), opt(np.float64), opt(np.float64)])
PfemStep3D_advect = bind(lib, "PfemStep3D_advect", i32, [_p, u32, arr(np.float64, 2), f64, u32])
PfemStep3D_set_learned = bind(lib, "PfemStep3D_set_learned", i32, [_p, _p])
PfemStep3D_get_points = bind(lib, "PfemStep3D_get_points", i32, [_p, arr(np.float64, 2, writeable=True)])
PfemStep3D_get_mesh = bind(lib, "PfemStep3D_get_mesh", i32,
                           [_p, opt(np.uint32), opt(np.uint8), opt(np.float64), opt(np.uint32)])


class PfemStep3D:
    """The 3D step: LBVH + 3D alpha-shape + P1 on tetrahedra + CG (solvers
    "fp64", "mixed", "fsai", "learned")."""

    def __init__(self, xyz, alpha, f, morton_bits: int = 21):
        self.h = PfemStep3D_create()
        if not self.h:
            raise StreamError(lib.Stream_last_error().decode())
        xyz, alpha, f = node_arrays(xyz, 3, np.float64, alpha, f)
        self.n = xyz.shape[0]
        check(PfemStep3D_set_morton_bits(self.h, morton_bits))
        check(PfemStep3D_set_points(self.h, self.n, xyz, alpha, f))

    def __del__(self):
        if getattr(self, "h", None):
            PfemStep3D_destroy(self.h)
            self.h = None

    def run(self, solver="mixed", rtol=1e-8, maxit=200000):
        st = np.zeros(10)
        kind = SOLVER_NAMES[solver] if isinstance(solver, str) else int(solver)
        check(PfemStep3D_run(self.h, kind, rtol, maxit, st))
        return dict(mesh_ms=st[0], assemble_ms=st[1], solve_ms=st[2], iters=int(st[3]), true_rel=st[4],
                    replacements=int(st[5]), nnz=int(st[6]), n_elems=int(st[7]), prep_ms=st[8],
                    fallback=bool(st[9]))

    def set_learned(self, net):
        """Network used by solver "learned" (kept alive by this object)."""
        self._learned = net
        check(PfemStep3D_set_learned(self.h, net.h if net is not None else None))

    def nnz(self):
        return int(PfemStep3D_nnz(self.h))

    def assemble(self):
        st = np.zeros(2)
        check(PfemStep3D_assemble(self.h, st))
        return float(st[0]), float(st[1])

    def set_points(self, xyz, alpha, f):
        xyz, alpha, f = node_arrays(xyz, 3, np.float64, alpha, f)
        self.n = xyz.shape[0]
        check(PfemStep3D_set_points(self.h, self.n, xyz, alpha, f))

    def advect(self, modes, dt: float, steps: int = 1):
        """Advect the points on the device in the divergence-free field of
        clouds.potential_modes rows (kx, ky, kz, phi, bx, by, bz)."""
        m = np.ascontiguousarray(modes, np.float64).reshape(-1, 7)
        check(PfemStep3D_advect(self.h, m.shape[0], m, float(dt), int(steps)))

    def points(self):
        xyz = np.empty((self.n, 3))
        check(PfemStep3D_get_points(self.h, xyz))
        return xyz

    def permute_forward(self, field):
        """Nodal field (shape (n, ...)) from input order to the Morton order
        of the last mesh, system and solution: out[i] = field[perm[i]]."""
        return permute_field(PfemStep3D_permute_forward, self.h, field, self.n)

    def permute_backward(self, field):
        """Nodal field from the Morton order of the last mesh back to input
        order: out[perm[i]] = field[i]."""
        return permute_field(PfemStep3D_permute_backward, self.h, field, self.n)

    def solution(self):
        p = np.empty(self.n)
        check(PfemStep3D_get_solution(self.h, p))
        return p

    def system(self):
        nnz = PfemStep3D_nnz(self.h)
        row = np.empty(self.n + 1, np.uint32)
        col = np.empty(nnz, np.uint32)
        val = np.empty(nnz)
        b = np.empty(self.n)
        check(PfemStep3D_get_system(self.h, row, col, val, b))
        return row, col, val, b

    def mesh(self):
        ne = PfemStep3D_get_nelem(self.h)
        elem = np.empty((ne, 4), np.uint32)
        bnd = np.empty(self.n, np.uint8)
        xyz = np.empty((self.n, 3))
        perm = np.empty(self.n, np.uint32)
        check(PfemStep3D_get_mesh(self.h, elem, bnd, xyz, perm))
        return elem, bnd, xyz, perm

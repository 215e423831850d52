"""AlphaShape3D binding, same class surface as the reference's
``stream.mesh.AlphaShape3D`` (python/mesh/ashape3D.py:75-103).

Reference behaviour kept: ``set_nodes`` casts to float32 and calls
``init``; ``compute`` runs compute + compress; ``get_elem`` returns
``uint32[M, 4]`` (tetrahedra) in Morton index space; ``get_ordered_nodes`` returns the
Morton-ordered coordinates.  Extensions: ``dtype=np.float64`` for FP64
nodes and exact FP64 predicates, ``morton_bits``, and the edge / boundary /
permutation queries of the reference's C++ class (getEdge,
getBoundaryNodes, getPermutation; src/mesh/DirectAlphaShape3D.cpp:75-112).
"""
from __future__ import annotations

import ctypes

import numpy as np
import numpy.typing as npt

from .. import StreamError, check, lib, libmesh
from .._ct import arr, bind, i32, node_arrays, opt, permute_field, u32, vp

AlphaShape3D_ptr = ctypes.c_void_p

AlphaShape3D_create = bind(libmesh, "AlphaShape3D_create", AlphaShape3D_ptr, [])
AlphaShape3D_destroy = bind(libmesh, "AlphaShape3D_destroy", None, [AlphaShape3D_ptr])
AlphaShape3D_set_nodes = bind(libmesh, "AlphaShape3D_set_nodes", None,
                              [AlphaShape3D_ptr, u32, arr(np.float32, 2), arr(np.float32, 1)])
AlphaShape3D_set_nodes_f64 = bind(libmesh, "AlphaShape3D_set_nodes_f64", i32,
                                  [AlphaShape3D_ptr, u32, arr(np.float64, 2), arr(np.float64, 1)])
AlphaShape3D_set_nodes_device_f64 = bind(libmesh, "AlphaShape3D_set_nodes_device_f64", i32,
                                         [AlphaShape3D_ptr, u32, vp])
AlphaShape3D_set_morton_bits = bind(libmesh, "AlphaShape3D_set_morton_bits", i32, [AlphaShape3D_ptr, i32])
AlphaShape3D_init = bind(libmesh, "AlphaShape3D_init", None, [AlphaShape3D_ptr])
AlphaShape3D_compute = bind(libmesh, "AlphaShape3D_compute", None, [AlphaShape3D_ptr])
AlphaShape3D_get_nelem = bind(libmesh, "AlphaShape3D_get_nelem", u32, [AlphaShape3D_ptr])
AlphaShape3D_get_elem = bind(libmesh, "AlphaShape3D_get_elem", None,
                             [AlphaShape3D_ptr, arr(np.uint32, 2, writeable=True)])
AlphaShape3D_get_ordered_nodes = bind(libmesh, "AlphaShape3D_get_ordered_nodes", None,
                                      [AlphaShape3D_ptr, arr(np.float32, 2, writeable=True)])
AlphaShape3D_get_ordered_nodes_f64 = bind(libmesh, "AlphaShape3D_get_ordered_nodes_f64", i32,
                                          [AlphaShape3D_ptr, arr(np.float64, 2, writeable=True)])
AlphaShape3D_get_nedge = bind(libmesh, "AlphaShape3D_get_nedge", u32, [AlphaShape3D_ptr])
AlphaShape3D_get_edges = bind(libmesh, "AlphaShape3D_get_edges", i32,
                              [AlphaShape3D_ptr, opt(np.uint32), opt(np.uint32), opt(np.uint8), opt(np.uint8)])
AlphaShape3D_get_permutation = bind(libmesh, "AlphaShape3D_get_permutation", i32,
                                    [AlphaShape3D_ptr, arr(np.uint32, 1, writeable=True)])
AlphaShape3D_get_nexact = bind(libmesh, "AlphaShape3D_get_nexact", u32, [AlphaShape3D_ptr])
AlphaShape3D_get_nclamped = bind(libmesh, "AlphaShape3D_get_nclamped", u32, [AlphaShape3D_ptr])
AlphaShape3D_last_status = bind(libmesh, "AlphaShape3D_last_status", i32, [AlphaShape3D_ptr])
AlphaShape3D_timed_run = bind(libmesh, "AlphaShape3D_timed_run", i32, [AlphaShape3D_ptr, arr(np.float64, 1, writeable=True)])
AlphaShape3D_permute_forward = bind(libmesh, "AlphaShape3D_permute_forward", i32, [AlphaShape3D_ptr, vp, vp, u32])
AlphaShape3D_permute_backward = bind(libmesh, "AlphaShape3D_permute_backward", i32, [AlphaShape3D_ptr, vp, vp, u32])


class AlphaShape3D:
    def __init__(self, dtype=np.float32, morton_bits: int = 21):
        self.mesh = AlphaShape3D_create()
        if not self.mesh:
            raise StreamError(lib.Stream_last_error().decode())
        self.nnodes = 0
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise TypeError("dtype must be float32 or float64")
        check(AlphaShape3D_set_morton_bits(self.mesh, morton_bits))

    def __del__(self):
        if getattr(self, "mesh", None):
            AlphaShape3D_destroy(self.mesh)
            self.mesh = None

    def _status(self):
        check(AlphaShape3D_last_status(self.mesh))

    def set_nodes(self, nodes: npt.NDArray, alpha: npt.NDArray) -> None:
        nodes, alpha = node_arrays(nodes, 3, self.dtype, alpha)
        self.nnodes = nodes.shape[0]
        if self.dtype == np.float32:
            AlphaShape3D_set_nodes(self.mesh, self.nnodes, nodes, alpha)
            self._status()
        else:
            check(AlphaShape3D_set_nodes_f64(self.mesh, self.nnodes, nodes, alpha))
        self.init()

    def init(self) -> None:
        AlphaShape3D_init(self.mesh)
        self._status()

    def compute(self) -> None:
        AlphaShape3D_compute(self.mesh)
        self._status()

    def get_elem(self) -> npt.NDArray[np.uint32]:
        nelems = AlphaShape3D_get_nelem(self.mesh)
        self._status()
        elems = np.empty((nelems, 4), dtype=np.uint32)
        AlphaShape3D_get_elem(self.mesh, elems)
        self._status()
        return elems

    def get_ordered_nodes(self) -> npt.NDArray:
        nodes = np.empty((self.nnodes, 3), dtype=self.dtype)
        if self.dtype == np.float32:
            AlphaShape3D_get_ordered_nodes(self.mesh, nodes)
            self._status()
        else:
            check(AlphaShape3D_get_ordered_nodes_f64(self.mesh, nodes))
        return nodes

    # ---- extensions (the reference's C++-only queries) ----
    def get_edges(self):
        """(row u32[N+1], neig u32[E], edge_is_bnd u8[E]) in Morton ids."""
        ne = AlphaShape3D_get_nedge(self.mesh)
        self._status()
        row = np.empty(self.nnodes + 1, np.uint32)
        neig = np.empty(ne, np.uint32)
        eb = np.empty(ne, np.uint8)
        check(AlphaShape3D_get_edges(self.mesh, row, neig, eb, None))
        return row, neig, eb

    def get_boundary_nodes(self) -> npt.NDArray[np.uint8]:
        nb = np.empty(self.nnodes, np.uint8)
        check(AlphaShape3D_get_edges(self.mesh, None, None, None, nb))
        return nb

    def timed_run(self) -> dict:
        """init + compute, timed as the reference's IMR26 benchmark loop:
        GPU ms of the LBVH build (tlbvh), the candidate search, local
        Delaunay, filter and scans (tashape) and the compaction (tcompress),
        the host wall ms of both calls and the GPU ms from the build's start
        to the compaction's end; within tashape the candidate search (fill),
        the local Delaunay of every node and the exact redo of the nodes
        whose filters could not decide a sign."""
        ms = np.zeros(8)
        check(AlphaShape3D_timed_run(self.mesh, ms))
        return dict(tlbvh=ms[0], tashape=ms[1], tcompress=ms[2], wall=ms[3], gpu_total=ms[4], fill=ms[5],
                    delaunay=ms[6], exact_redo=ms[7])

    def permute_forward(self, field: npt.NDArray) -> npt.NDArray:
        """Nodal field (shape (n, ...)) from input order to the Morton order
        of the outputs: out[i] = field[perm[i]] (permuteForward)."""
        return permute_field(AlphaShape3D_permute_forward, self.mesh, field, self.nnodes)

    def permute_backward(self, field: npt.NDArray) -> npt.NDArray:
        """Nodal field from Morton order back to input order:
        out[perm[i]] = field[i] (permuteBackward)."""
        return permute_field(AlphaShape3D_permute_backward, self.mesh, field, self.nnodes)

    def get_permutation(self) -> npt.NDArray[np.uint32]:
        """perm[i] = input index of Morton slot i."""
        p = np.empty(self.nnodes, np.uint32)
        check(AlphaShape3D_get_permutation(self.mesh, p))
        return p

    def get_nexact(self) -> int:
        """Nodes whose local tetrahedralisation needed the exact predicate stage."""
        r = AlphaShape3D_get_nexact(self.mesh)
        self._status()
        return int(r)

    def get_nclamped(self) -> int:
        """Nodes whose candidate list was clamped to 124 (the reference's 128
        entries including the four infinity nodes)."""
        r = AlphaShape3D_get_nclamped(self.mesh)
        self._status()
        return int(r)

"""Mesh2D binding, same class surface as the reference's
``stream.mesh.Mesh2D`` (python/mesh/mesh.py:86-128): full 2D Delaunay
triangulation by per-node local stars on the GPU.

Reference behaviour kept: ``set_nodes`` casts to float32 and calls
``init``; ``get_elem`` compresses the stars and returns ``uint32[M, 3]``
triangles in Morton index space (``remove_super_nodes`` first);
``get_ordered_nodes`` returns the Morton-ordered coordinates.  Extensions:
``get_permutation``, ``get_stats`` (overflow count, rounds, exact
evaluations, per-phase GPU milliseconds), ``get_stars`` and ``get_cells``
for tests.
"""
from __future__ import annotations

import ctypes

import numpy as np
import numpy.typing as npt

from .. import StreamError, check, lib, libmesh
from .._ct import arr, bind, i32, node_arrays, opt, u32

mesh2D_ptr = ctypes.c_void_p

mesh2D_create = bind(libmesh, "Mesh2D_create", mesh2D_ptr, [])
mesh2D_destroy = bind(libmesh, "Mesh2D_destroy", None, [mesh2D_ptr])
mesh2D_set_nodes = bind(libmesh, "Mesh2D_set_nodes", None, [mesh2D_ptr, u32, arr(np.float32, 2)])
mesh2D_init = bind(libmesh, "Mesh2D_init", None, [mesh2D_ptr])
mesh2D_insert_morton_neighbors = bind(libmesh, "Mesh2D_insert_morton_neighbors", None, [mesh2D_ptr])
mesh2D_insert_BVH_neighbors = bind(libmesh, "Mesh2D_insert_BVH_neighbors", None, [mesh2D_ptr])
mesh2D_insert_iterative = bind(libmesh, "Mesh2D_insert_iterative", None, [mesh2D_ptr])
mesh2D_remove_super_nodes = bind(libmesh, "Mesh2D_remove_super_nodes", None, [mesh2D_ptr])
mesh2D_get_nelem = bind(libmesh, "Mesh2D_get_nelem", u32, [mesh2D_ptr])
mesh2D_get_elem = bind(libmesh, "Mesh2D_get_elem", None, [mesh2D_ptr, arr(np.uint32, 2, writeable=True)])
mesh2D_get_ordered_nodes = bind(libmesh, "Mesh2D_get_ordered_nodes", None,
                                [mesh2D_ptr, arr(np.float32, 2, writeable=True)])
mesh2D_last_status = bind(libmesh, "Mesh2D_last_status", i32, [mesh2D_ptr])
mesh2D_get_stats = bind(libmesh, "Mesh2D_get_stats", i32, [mesh2D_ptr, opt(np.uint32), opt(np.float32)])
mesh2D_get_stars = bind(libmesh, "Mesh2D_get_stars", i32,
                        [mesh2D_ptr, opt(np.uint16), opt(np.uint32), opt(np.uint32), opt(np.uint32), opt(np.uint32)])
mesh2D_get_cells = bind(libmesh, "Mesh2D_get_cells", i32, [mesh2D_ptr, opt(np.float32)])

STAR_CAP = 64


class Mesh2D:
    def __init__(self):
        self.mesh = mesh2D_create()
        if not self.mesh:
            raise StreamError(lib.Stream_last_error().decode())
        self.nnodes = 0

    def __del__(self):
        if getattr(self, "mesh", None):
            mesh2D_destroy(self.mesh)
            self.mesh = None

    def _status(self):
        check(mesh2D_last_status(self.mesh))

    def init(self) -> None:
        mesh2D_init(self.mesh)
        self._status()

    def set_nodes(self, nodes: npt.NDArray[np.float32]) -> None:
        (nodes,) = node_arrays(nodes, 2, np.float32)
        self.nnodes = nodes.shape[0]
        mesh2D_set_nodes(self.mesh, self.nnodes, nodes)
        self._status()
        self.init()

    def insert_morton_neighbors(self) -> None:
        mesh2D_insert_morton_neighbors(self.mesh)
        self._status()

    def insert_BVH_neighbors(self) -> None:
        mesh2D_insert_BVH_neighbors(self.mesh)
        self._status()

    def insert_iterative(self) -> None:
        mesh2D_insert_iterative(self.mesh)
        self._status()

    def remove_super_nodes(self) -> None:
        mesh2D_remove_super_nodes(self.mesh)
        self._status()

    def get_elem(self) -> npt.NDArray[np.uint32]:
        nelems = mesh2D_get_nelem(self.mesh)
        self._status()
        elems = np.empty((nelems, 3), dtype=np.uint32)
        mesh2D_get_elem(self.mesh, elems)
        self._status()
        return elems

    def get_ordered_nodes(self) -> npt.NDArray[np.float32]:
        nodes = np.empty((self.nnodes, 2), dtype=np.float32)
        mesh2D_get_ordered_nodes(self.mesh, nodes)
        self._status()
        return nodes

    # ---- extensions ----
    def get_permutation(self) -> npt.NDArray[np.uint32]:
        """perm[i] = input index of Morton slot i."""
        p = np.empty(self.nnodes, np.uint32)
        check(mesh2D_get_stars(self.mesh, None, None, None, None, p))
        return p

    def get_stats(self) -> dict:
        c = np.zeros(3, np.uint32)
        ms = np.zeros(5, np.float32)
        check(mesh2D_get_stats(self.mesh, c, ms))
        return dict(overflow=int(c[0]), rounds=int(c[1]), exact=int(c[2]),
                    ms=dict(zip(("init", "morton", "bvh", "iterative", "remove"), map(float, ms))))

    def get_stars(self) -> dict:
        n = self.nnodes
        r = dict(elem=np.empty((n, STAR_CAP), np.uint16), nelem=np.empty(n, np.uint32),
                 neig=np.empty((n, STAR_CAP), np.uint32), nlocal=np.empty(n, np.uint32))
        check(mesh2D_get_stars(self.mesh, r["elem"], r["nelem"], r["neig"], r["nlocal"], None))
        return r

    def get_cells(self) -> npt.NDArray[np.float32]:
        c = np.empty((2 * self.nnodes - 1, 4), np.float32)
        check(mesh2D_get_cells(self.mesh, c))
        return c

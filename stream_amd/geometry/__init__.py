"""LBVH bindings (include/geometry/lbvh.hpp:61-362).

The reference exposes its LBVH only as C++ templates; these functions are
the test and tooling surface of the GPU build: Morton codes, the stable sort
permutation, the Karras topology and the refitted boxes, returned as numpy
arrays.
"""
from __future__ import annotations

import numpy as np

from .. import check, libgeo
from .._ct import bind, i32, opt, u32, vp

_build = {}
for dim, tag, dt in ((2, "f32", np.float32), (2, "f64", np.float64), (3, "f64", np.float64)):
    _build[(dim, dt)] = bind(
        libgeo, f"LBVH{dim}D_build_{tag}", i32,
        [u32, opt(dt, False), i32, i32, opt(np.uint64), opt(np.uint32), opt(np.int32), opt(np.uint8),
         opt(np.uint32), opt(np.uint32), opt(dt), opt(dt)],
    )
_morton = {
    (2, np.float32): bind(libgeo, "LBVH2D_morton_f32", i32, [u32, opt(np.float32, False), i32, opt(np.uint64)]),
    (2, np.float64): bind(libgeo, "LBVH2D_morton_f64", i32, [u32, opt(np.float64, False), i32, opt(np.uint64)]),
    (3, np.float64): bind(libgeo, "LBVH3D_morton_f64", i32, [u32, opt(np.float64, False), i32, opt(np.uint64)]),
}
_collide = {
    np.float32: bind(libgeo, "LBVH2D_collisions_f32", i32, [u32, opt(np.float32, False), i32, vp]),
    np.float64: bind(libgeo, "LBVH2D_collisions_f64", i32, [u32, opt(np.float64, False), i32, vp]),
}

LEAF_CENTRE = 0  # alpha-shape node functor: box of the centre
LEAF_SPHERE = 1  # ComputeBB_Functor: box of the sphere


def _spheres(centres, radii, dtype):
    c = np.asarray(centres, dtype=dtype)
    n, dim = c.shape
    out = np.empty((n, dim + 1), dtype=dtype)
    out[:, :dim] = c
    out[:, dim] = radii
    return np.ascontiguousarray(out)


def morton_codes(centres, radii, bits=None, dtype=np.float64):
    """Unsorted Morton codes of the sphere centres in the global sphere box."""
    xyr = _spheres(centres, radii, dtype)
    n, dim = xyr.shape[0], xyr.shape[1] - 1
    bits = 64 // dim if bits is None else bits
    out = np.empty(n, np.uint64)
    check(_morton[(dim, np.dtype(dtype).type)](n, xyr, bits, out))
    return out


def build(centres, radii, bits=None, leaf_mode=LEAF_CENTRE, dtype=np.float64):
    """Build the LBVH; returns a dict of host arrays."""
    xyr = _spheres(centres, radii, dtype)
    n, dim = xyr.shape[0], xyr.shape[1] - 1
    bits = 64 // dim if bits is None else bits
    r = dict(
        codes_sorted=np.empty(n, np.uint64), perm=np.empty(n, np.uint32), sep=np.empty(n - 1, np.int32),
        child_is_leaf=np.empty(n - 1, np.uint8), leaf_parent=np.empty(n, np.uint32),
        internal_parent=np.empty(n - 1, np.uint32), internal_box=np.empty((n - 1, 2 * dim), dtype),
        global_box=np.empty(2 * dim, dtype),
    )
    check(_build[(dim, np.dtype(dtype).type)](
        n, xyr, bits, leaf_mode, r["codes_sorted"], r["perm"], r["sep"], r["child_is_leaf"], r["leaf_parent"],
        r["internal_parent"], r["internal_box"], r["global_box"]))
    return r


def collisions(centres, radii, bits=None, dtype=np.float64):
    """Total sphere-sphere overlaps (self excluded) found through the LBVH."""
    import ctypes

    xyr = _spheres(centres, radii, dtype)
    n = xyr.shape[0]
    bits = 32 if bits is None else bits
    tot = ctypes.c_uint64(0)
    check(_collide[np.dtype(dtype).type](n, xyr, bits, ctypes.byref(tot)))
    return int(tot.value)


_abuild = {
    (2, np.float32): bind(libgeo, "LBVHa2D_build_f32", i32,
                          [u32, opt(np.float32, False), i32, i32] + [opt(np.uint32)] * 5 + [opt(np.float32),
                                                                                            opt(np.uint32)]),
    (2, np.float64): bind(libgeo, "LBVHa2D_build_f64", i32,
                          [u32, opt(np.float64, False), i32, i32] + [opt(np.uint32)] * 5 + [opt(np.float64),
                                                                                            opt(np.uint32)]),
    (3, np.float64): bind(libgeo, "LBVHa3D_build_f64", i32,
                          [u32, opt(np.float64, False), i32, i32] + [opt(np.uint32)] * 5 + [opt(np.float64),
                                                                                            opt(np.uint32)]),
}
_acollide = {
    (2, np.float32): bind(libgeo, "LBVHa2D_collisions_f32", i32, [u32, opt(np.float32, False), i32, u32, vp]),
    (2, np.float64): bind(libgeo, "LBVHa2D_collisions_f64", i32, [u32, opt(np.float64, False), i32, u32, vp]),
    (3, np.float64): bind(libgeo, "LBVHa3D_collisions_f64", i32, [u32, opt(np.float64, False), i32, u32, vp]),
}


def build_agglomerative(centres, radii, bits=None, leaf_mode=LEAF_SPHERE, dtype=np.float64):
    """Agglomerative LBVH (the reference's LBVHa); returns a dict of host arrays."""
    xyr = _spheres(centres, radii, dtype)
    n, dim = xyr.shape[0], xyr.shape[1] - 1
    bits = 64 // dim if bits is None else bits
    r = dict(perm=np.empty(n, np.uint32), children=np.empty((n - 1, 2), np.uint32),
             range=np.empty((n - 1, 2), np.uint32), parent=np.empty(2 * n - 1, np.uint32),
             root=np.zeros(1, np.uint32), boxes=np.empty((2 * n - 1, 2 * dim), dtype),
             split=np.empty(n - 1, np.uint32))
    check(_abuild[(dim, np.dtype(dtype).type)](n, xyr, bits, leaf_mode, r["perm"], r["children"], r["range"],
                                               r["parent"], r["root"], r["boxes"], r["split"]))
    return r


def collisions_agglomerative(centres, radii, leaf_size=1, bits=None, dtype=np.float64):
    """Overlaps (self excluded) through the agglomerative LBVH with leaf buckets."""
    import ctypes

    xyr = _spheres(centres, radii, dtype)
    n, dim = xyr.shape[0], xyr.shape[1] - 1
    bits = 64 // dim if bits is None else bits
    tot = ctypes.c_uint64(0)
    check(_acollide[(dim, np.dtype(dtype).type)](n, xyr, bits, leaf_size, ctypes.byref(tot)))
    return int(tot.value)


_tbuild = {
    2: bind(libgeo, "TRBVH2D_build_f32", i32,
            [u32, opt(np.float32, False), i32] + [opt(np.uint32)] * 5 + [opt(np.float32)] * 2
            + [opt(np.uint32), opt(np.float64)]),
    3: bind(libgeo, "TRBVH3D_build_f32", i32,
            [u32, opt(np.float32, False), i32] + [opt(np.uint32)] * 5 + [opt(np.float32)] * 2
            + [opt(np.uint32), opt(np.float64)]),
}
_tcollide = {
    2: bind(libgeo, "TRBVH2D_collisions_f32", i32, [u32, opt(np.float32, False), i32, vp]),
    3: bind(libgeo, "TRBVH3D_collisions_f32", i32, [u32, opt(np.float32, False), i32, vp]),
}


def build_trbvh(centres, radii, rounds=3):
    """TRBVH (the reference's treelet-restructured BVH, float); returns a
    dict of host arrays: perm, children, range, parent, root, boxes, sah,
    subtree and ms = [build, rounds] GPU milliseconds."""
    xyr = _spheres(centres, radii, np.float32)
    n, dim = xyr.shape[0], xyr.shape[1] - 1
    r = dict(perm=np.empty(n, np.uint32), children=np.empty((n - 1, 2), np.uint32),
             range=np.empty((n - 1, 2), np.uint32), parent=np.empty(2 * n - 1, np.uint32),
             root=np.zeros(1, np.uint32), boxes=np.empty((2 * n - 1, 2 * dim), np.float32),
             sah=np.empty(2 * n - 1, np.float32), subtree=np.empty(2 * n - 1, np.uint32),
             ms=np.zeros(2, np.float64))
    check(_tbuild[dim](n, xyr, rounds, r["perm"], r["children"], r["range"], r["parent"], r["root"], r["boxes"],
                       r["sah"], r["subtree"], r["ms"]))
    return r


def collisions_trbvh(centres, radii, rounds=3):
    """Overlaps (self excluded) through the TRBVH."""
    import ctypes

    xyr = _spheres(centres, radii, np.float32)
    n, dim = xyr.shape[0], xyr.shape[1] - 1
    tot = ctypes.c_uint64(0)
    check(_tcollide[dim](n, xyr, rounds, ctypes.byref(tot)))
    return int(tot.value)

"""ctypes loader of the CPU oracle (oracle/lib/liboracle.so) and of the
reference-header harness (oracle/_ref/libref_geom.so).  Test
infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "lib", "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref_geom.so")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        _lib = ctypes.CDLL(ORACLE_SO)
        _declare(_lib)
    return _lib


def _bind_ref(lib):
    P = ctypes.POINTER
    f = ctypes.c_float
    for name in ("ref_orient2d", "ref_orient2dexact"):
        getattr(lib, name).restype = f
        getattr(lib, name).argtypes = [P(f)] * 3
    lib.ref_incircle.restype = f
    lib.ref_incircle.argtypes = [P(f)] * 5
    lib.ref_incircleexact.restype = f
    lib.ref_incircleexact.argtypes = [P(f)] * 4
    for name in ("ref_orient3d", "ref_orient3dexact"):
        getattr(lib, name).restype = f
        getattr(lib, name).argtypes = [P(f)] * 4
    lib.ref_insphere.restype = f
    lib.ref_insphere.argtypes = [P(f)] * 6
    lib.ref_insphereexact.restype = f
    lib.ref_insphereexact.argtypes = [P(f)] * 5
    return lib


_ref_fma = None


def ref_fma():
    """The same reference headers compiled with fma contraction (-mfma
    -ffp-contract=fast), or None if absent."""
    global _ref_fma
    so = REF_SO.replace("libref_geom.so", "libref_geom_fma.so")
    if _ref_fma is None and os.path.exists(so):
        _ref_fma = _bind_ref(ctypes.CDLL(so))
    return _ref_fma


def ref():
    """The reference's own predicates/Morton headers compiled (None if absent)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO)
        P = ctypes.POINTER
        f = ctypes.c_float
        for name in ("ref_orient2d", "ref_orient2dexact"):
            getattr(_ref, name).restype = f
            getattr(_ref, name).argtypes = [P(f)] * 3
        _ref.ref_incircle.restype = f
        _ref.ref_incircle.argtypes = [P(f)] * 5
        _ref.ref_incircleexact.restype = f
        _ref.ref_incircleexact.argtypes = [P(f)] * 4
        for name in ("ref_orient3d", "ref_orient3dexact"):
            getattr(_ref, name).restype = f
            getattr(_ref, name).argtypes = [P(f)] * 4
        _ref.ref_insphere.restype = f
        _ref.ref_insphere.argtypes = [P(f)] * 6
        _ref.ref_insphereexact.restype = f
        _ref.ref_insphereexact.argtypes = [P(f)] * 5
        _ref.ref_encode2d.restype = ctypes.c_uint64
        _ref.ref_encode2d.argtypes = [ctypes.c_uint32] * 2
        _ref.ref_encode3d.restype = ctypes.c_uint64
        _ref.ref_encode3d.argtypes = [ctypes.c_uint32] * 3
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _declare(L):
    vp, u32, u64, i32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double
    for tag in ("f32", "f64"):
        getattr(L, f"orc_lbvh2d_build_{tag}").restype = i32
        getattr(L, f"orc_ashape2d_run_{tag}").restype = vp
        getattr(L, f"orc_ashape2d_run_{tag}").argtypes = [u32, vp, i32]
        getattr(L, f"orc_lbvh2d_collisions_{tag}").restype = u64
        getattr(L, f"orc_lbvh2d_collisions_{tag}").argtypes = [u32, vp, i32]
    L.orc_lbvh3d_build_f64.restype = i32
    L.orc_interleave.restype = u64
    L.orc_interleave.argtypes = [vp, i32, i32]
    L.orc_orient2d_f32.restype = ctypes.c_float
    L.orc_orient2d_f64.restype = f64
    L.orc_orient2d_exact_f32.restype = i32
    L.orc_orient2d_exact_f64.restype = i32
    L.orc_incircle_filter_f32.restype = ctypes.c_float
    L.orc_incircle_filter_f64.restype = f64
    L.orc_incircle_exact_f32.restype = i32
    L.orc_incircle_exact_f64.restype = i32
    L.orc_ashape2d_sizes.argtypes = [vp, i32, vp]
    L.orc_ashape2d_free.argtypes = [vp, i32]
    L.orc_assemble_p1.restype = ctypes.c_int64
    L.orc_assemble_p1.argtypes = [u32, vp, u32, vp, vp, vp, u64, vp, vp, vp, vp]
    L.orc_step_solve_f64.restype = None
    L.orc_step_solve_f64.argtypes = [u32, vp, vp, f64, f64, u32, vp, vp, vp, vp]
    L.orc_jacobi_pcg.restype = u32
    L.orc_jacobi_pcg.argtypes = [u32, vp, vp, vp, vp, vp, f64, u32, vp]
    for tag, ft in (("f32", ctypes.c_float), ("f64", f64)):
        getattr(L, f"orc_orient3d_{tag}").restype = ft
        getattr(L, f"orc_orient3d_{tag}").argtypes = [vp] * 4
        getattr(L, f"orc_orient3d_exact_{tag}").restype = i32
        getattr(L, f"orc_orient3d_exact_{tag}").argtypes = [vp] * 4
        getattr(L, f"orc_insphere_filter_{tag}").restype = ft
        getattr(L, f"orc_insphere_filter_{tag}").argtypes = [vp] * 6
        getattr(L, f"orc_insphere_exact_{tag}").restype = i32
        getattr(L, f"orc_insphere_exact_{tag}").argtypes = [vp] * 5
        getattr(L, f"orc_ashape3d_run_{tag}").restype = vp
        getattr(L, f"orc_ashape3d_run_{tag}").argtypes = [u32, vp, i32]
        getattr(L, f"orc_ashape3d_get_{tag}").argtypes = [vp] * 9
    L.orc_ashape3d_run_contract_f32.restype = vp
    L.orc_ashape3d_run_contract_f32.argtypes = [u32, vp, i32, i32]
    L.orc_ashape3d_capped_f32.argtypes = [vp, vp]
    L.orc_ashape3d_run_probe_f32.restype = vp
    L.orc_ashape3d_run_probe_f32.argtypes = [u32, vp, i32, i32, i32]
    L.orc_ashape3d_exact_points_f32.restype = u64
    L.orc_ashape3d_exact_points_f32.argtypes = [vp, vp]
    L.orc_ashape2d_run_contract_f32.restype = vp
    L.orc_ashape2d_run_contract_f32.argtypes = [u32, vp, i32, i32]
    L.orc_ashape3d_sizes.argtypes = [vp, i32, vp]
    L.orc_lbvha2d_build_f64.restype = i32
    L.orc_lbvha2d_build_f64.argtypes = [u32, vp, i32, i32] + [vp] * 7
    for name in ("orc_lbvha2d_collisions_f64", "orc_lbvha2d_collisions_f32", "orc_lbvha3d_collisions_f64"):
        getattr(L, name).restype = u64
        getattr(L, name).argtypes = [u32, vp, i32, u32]
    for d in (2, 3):
        getattr(L, f"orc_trbvh{d}d_build_f32").restype = i32
        getattr(L, f"orc_trbvh{d}d_build_f32").argtypes = [u32, vp, i32] + [vp] * 8
        getattr(L, f"orc_trbvh{d}d_collisions_f32").restype = u64
        getattr(L, f"orc_trbvh{d}d_collisions_f32").argtypes = [u32, vp, i32]
    L.orc_mesh2d_create.restype = vp
    L.orc_mesh2d_create.argtypes = [u32, vp]
    for name in ("orc_mesh2d_free", "orc_mesh2d_insert_morton", "orc_mesh2d_insert_bvh", "orc_mesh2d_remove_super"):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = [vp]
    L.orc_mesh2d_insert_iterative.restype = u32
    L.orc_mesh2d_insert_iterative.argtypes = [vp]
    L.orc_mesh2d_nelem.restype = u32
    L.orc_mesh2d_nelem.argtypes = [vp]
    L.orc_mesh2d_elem.restype = None
    L.orc_mesh2d_elem.argtypes = [vp, vp]
    L.orc_mesh2d_overflow.restype = u64
    L.orc_mesh2d_overflow.argtypes = [vp]
    L.orc_mesh2d_stars.restype = None
    L.orc_mesh2d_stars.argtypes = [vp] * 6
    L.orc_mesh2d_cells.restype = None
    L.orc_mesh2d_cells.argtypes = [vp, vp]
    L.orc_treelet_dp2d.restype = None
    L.orc_treelet_dp2d.argtypes = [vp] * 6
    L.orc_assemble_p1_tet.restype = ctypes.c_int64
    L.orc_assemble_p1_tet.argtypes = [u32, vp, u32, vp, vp, vp, u64, vp, vp, vp, vp]
    L.orc_ashape3d_free.argtypes = [vp, i32]
    L.orc_step_f64.restype = u32
    L.orc_step_f64.argtypes = [u32, vp, vp, u32, vp, vp]


def spheres(coords, alpha, dtype):
    c = np.asarray(coords, dtype)
    out = np.empty((c.shape[0], c.shape[1] + 1), dtype)
    out[:, :-1] = c
    out[:, -1] = alpha
    return np.ascontiguousarray(out)


def lbvh2d(coords, radii, bits=32, leaf_mode=0, dtype=np.float64):
    L = lib()
    xyr = spheres(coords, radii, dtype)
    n = xyr.shape[0]
    r = dict(codes=np.empty(n, np.uint64), codes_sorted=np.empty(n, np.uint64), perm=np.empty(n, np.uint32),
             sep=np.empty(n - 1, np.int32), child_is_leaf=np.empty(n - 1, np.uint8),
             leaf_parent=np.empty(n, np.uint32), internal_parent=np.empty(n - 1, np.uint32),
             internal_box=np.empty((n - 1, 4), dtype), global_box=np.empty(4, dtype))
    fn = L.orc_lbvh2d_build_f64 if dtype == np.float64 else L.orc_lbvh2d_build_f32
    rc = fn(ctypes.c_uint32(n), _p(xyr), ctypes.c_int32(bits), ctypes.c_int32(leaf_mode), _p(r["codes"]),
            _p(r["codes_sorted"]), _p(r["perm"]), _p(r["sep"]), _p(r["child_is_leaf"]), _p(r["leaf_parent"]),
            _p(r["internal_parent"]), _p(r["internal_box"]), _p(r["global_box"]))
    assert rc == 0
    return r


def lbvh2d_collisions(coords, radii, bits=32, dtype=np.float64):
    L = lib()
    xyr = spheres(coords, radii, dtype)
    fn = L.orc_lbvh2d_collisions_f64 if dtype == np.float64 else L.orc_lbvh2d_collisions_f32
    return int(fn(xyr.shape[0], _p(xyr), bits))


def ashape2d(coords, alpha, bits=32, dtype=np.float64, contract=0):
    """Run the oracle alpha-shape; returns a dict of outputs.  `contract`
    (FP32 only): fmac mode bits of oracle.hpp for the inexact tests."""
    L = lib()
    xyr = spheres(coords, alpha, dtype)
    n = xyr.shape[0]
    is64 = dtype == np.float64
    if contract:
        assert not is64
        h = L.orc_ashape2d_run_contract_f32(n, _p(xyr), bits, contract)
    else:
        h = (L.orc_ashape2d_run_f64 if is64 else L.orc_ashape2d_run_f32)(n, _p(xyr), bits)
    sz = np.zeros(4, np.uint64)
    L.orc_ashape2d_sizes(h, int(is64), _p(sz))
    ne, nedge = int(sz[1]), int(sz[2])
    r = dict(elem=np.empty((ne, 3), np.uint32), row=np.empty(n + 1, np.uint32), neig=np.empty(nedge, np.uint32),
             edge_is_bnd=np.empty(nedge, np.uint8), node_is_bnd=np.empty(n, np.uint8), perm=np.empty(n, np.uint32),
             ordered=np.empty((n, 2), dtype), nineig=np.empty(n, np.uint32), n_exact_calls=int(sz[3]))
    (L.orc_ashape2d_get_f64 if is64 else L.orc_ashape2d_get_f32)(
        ctypes.c_void_p(h), _p(r["elem"]), _p(r["row"]), _p(r["neig"]), _p(r["edge_is_bnd"]), _p(r["node_is_bnd"]),
        _p(r["perm"]), _p(r["ordered"]), _p(r["nineig"]))
    L.orc_ashape2d_free(h, int(is64))
    return r


def lbvha2d(coords, radii, bits=32, leaf_mode=1):
    L = lib()
    xyr = spheres(coords, radii, np.float64)
    n = xyr.shape[0]
    r = dict(perm=np.empty(n, np.uint32), children=np.empty((n - 1, 2), np.uint32),
             range=np.empty((n - 1, 2), np.uint32), parent=np.empty(2 * n - 1, np.uint32),
             root=np.zeros(1, np.uint32), boxes=np.empty((2 * n - 1, 4)), split=np.empty(n - 1, np.uint32))
    rc = L.orc_lbvha2d_build_f64(n, _p(xyr), bits, leaf_mode, _p(r["perm"]), _p(r["children"]), _p(r["range"]),
                                 _p(r["parent"]), _p(r["root"]), _p(r["boxes"]), _p(r["split"]))
    assert rc == 0
    return r


def lbvha_collisions(coords, radii, leaf_size, bits=32, dtype=np.float64):
    L = lib()
    xyr = spheres(coords, radii, dtype)
    d = xyr.shape[1] - 1
    if d == 3:
        return int(L.orc_lbvha3d_collisions_f64(xyr.shape[0], _p(xyr), bits, leaf_size))
    fn = L.orc_lbvha2d_collisions_f64 if dtype == np.float64 else L.orc_lbvha2d_collisions_f32
    return int(fn(xyr.shape[0], _p(xyr), bits, leaf_size))


def trbvh(coords, radii, rounds=3):
    """Oracle TRBVH (float): dict of perm, children, range, parent, root,
    boxes, sah, subtree."""
    L = lib()
    xyr = spheres(coords, radii, np.float32)
    n, d = xyr.shape[0], xyr.shape[1] - 1
    r = dict(perm=np.empty(n, np.uint32), children=np.empty((n - 1, 2), np.uint32),
             range=np.empty((n - 1, 2), np.uint32), parent=np.empty(2 * n - 1, np.uint32),
             root=np.zeros(1, np.uint32), boxes=np.empty((2 * n - 1, 2 * d), np.float32),
             sah=np.empty(2 * n - 1, np.float32), subtree=np.empty(2 * n - 1, np.uint32))
    fn = getattr(L, f"orc_trbvh{d}d_build_f32")
    rc = fn(n, _p(xyr), rounds, _p(r["perm"]), _p(r["children"]), _p(r["range"]), _p(r["parent"]), _p(r["root"]),
            _p(r["boxes"]), _p(r["sah"]), _p(r["subtree"]))
    assert rc == 0
    return r


def trbvh_collisions(coords, radii, rounds=3):
    L = lib()
    xyr = spheres(coords, radii, np.float32)
    d = xyr.shape[1] - 1
    return int(getattr(L, f"orc_trbvh{d}d_collisions_f32")(xyr.shape[0], _p(xyr), rounds))


def treelet_dp(boxes, cost, size):
    """The treelet optimiser's subset programme on 5 leaves (2D boxes as rows
    lo0, lo1, hi0, hi1): returns (copt[32], popt[32], size[32])."""
    b = np.ascontiguousarray(boxes, np.float32)
    c = np.ascontiguousarray(cost, np.float32)
    z = np.ascontiguousarray(size, np.uint32)
    copt = np.zeros(32, np.float32)
    popt = np.zeros(32, np.uint32)
    sz = np.zeros(32, np.uint32)
    lib().orc_treelet_dp2d(_p(b), _p(c), _p(z), _p(copt), _p(popt), _p(sz))
    return copt, popt, sz


def ashape3d(coords, alpha, bits=21, dtype=np.float64, contract=0, probe=None):
    """Run the 3D oracle alpha-shape; returns a dict of outputs.  `contract`
    (FP32 only) evaluates the inexact search and filter tests fused (fmac
    mode bits of oracle.hpp).  `probe` (FP32 only): {"cap": cavity-walk step
    cap, "record": 1 to return the points of every exact insphere call as
    r["exact_points"], float32 [m, 5, 3]}."""
    L = lib()
    xyzr = spheres(coords, alpha, dtype)
    n = xyzr.shape[0]
    is64 = dtype == np.float64
    tag = "f64" if is64 else "f32"
    if probe is not None:
        assert not is64
        h = L.orc_ashape3d_run_probe_f32(n, _p(xyzr), bits, int(probe.get("cap", 100)), int(probe.get("record", 0)))
    elif contract:
        assert not is64
        h = L.orc_ashape3d_run_contract_f32(n, _p(xyzr), bits, contract)
    else:
        h = getattr(L, f"orc_ashape3d_run_{tag}")(n, _p(xyzr), bits)
    sz = np.zeros(8, np.uint64)
    L.orc_ashape3d_sizes(h, int(is64), _p(sz))
    ne, nedge = int(sz[1]), int(sz[2])
    r = dict(elem=np.empty((ne, 4), np.uint32), row=np.empty(n + 1, np.uint32), neig=np.empty(nedge, np.uint32),
             edge_is_bnd=np.empty(nedge, np.uint8), node_is_bnd=np.empty(n, np.uint8), perm=np.empty(n, np.uint32),
             ordered=np.empty((n, 3), dtype), nineig=np.empty(n, np.uint32), n_exact_calls=int(sz[3]),
             over_cap=int(sz[4]), broken_cycles=int(sz[5]), capped_walks=int(sz[6]))
    if not is64:
        r["node_capped"] = np.empty(n, np.uint8)
        L.orc_ashape3d_capped_f32(ctypes.c_void_p(h), _p(r["node_capped"]))
        m = L.orc_ashape3d_exact_points_f32(ctypes.c_void_p(h), None)
        r["exact_points"] = np.empty((m, 5, 3), np.float32)
        if m:
            L.orc_ashape3d_exact_points_f32(ctypes.c_void_p(h), _p(r["exact_points"]))
    getattr(L, f"orc_ashape3d_get_{tag}")(
        ctypes.c_void_p(h), _p(r["elem"]), _p(r["row"]), _p(r["neig"]), _p(r["edge_is_bnd"]), _p(r["node_is_bnd"]),
        _p(r["perm"]), _p(r["ordered"]), _p(r["nineig"]))
    L.orc_ashape3d_free(h, int(is64))
    return r


def assemble_p1(xy, elem, bnd, f):
    L = lib()
    n = xy.shape[0]
    xy = np.ascontiguousarray(xy, np.float64)
    elem = np.ascontiguousarray(elem, np.uint32)
    bnd = np.ascontiguousarray(bnd, np.uint8)
    f = np.ascontiguousarray(f, np.float64)
    cap = n + 6 * elem.shape[0] + 16
    row = np.empty(n + 1, np.uint32)
    col = np.empty(cap, np.uint32)
    val = np.empty(cap, np.float64)
    b = np.empty(n, np.float64)
    nnz = L.orc_assemble_p1(n, _p(xy), elem.shape[0], _p(elem), _p(bnd), _p(f), cap, _p(row), _p(col), _p(val), _p(b))
    assert nnz >= 0
    return row, col[:nnz].copy(), val[:nnz].copy(), b


def assemble_p1_tet(xyz, elem, bnd, f):
    L = lib()
    n = xyz.shape[0]
    xyz = np.ascontiguousarray(xyz, np.float64)
    elem = np.ascontiguousarray(elem, np.uint32)
    bnd = np.ascontiguousarray(bnd, np.uint8)
    f = np.ascontiguousarray(f, np.float64)
    cap = n + 12 * elem.shape[0] + 16
    row = np.empty(n + 1, np.uint32)
    col = np.empty(cap, np.uint32)
    val = np.empty(cap, np.float64)
    b = np.empty(n, np.float64)
    nnz = L.orc_assemble_p1_tet(n, _p(xyz), elem.shape[0], _p(elem), _p(bnd), _p(f), cap, _p(row), _p(col),
                                _p(val), _p(b))
    assert nnz >= 0
    return row, col[:nnz].copy(), val[:nnz].copy(), b


def jacobi_pcg(row, col, val, b, x0=None, rtol=1e-8, maxit=100000):
    L = lib()
    n = b.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    rel = np.zeros(1)
    row = np.ascontiguousarray(row, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    it = L.orc_jacobi_pcg(n, _p(row), _p(col), _p(val), _p(b), _p(x), rtol, maxit, _p(rel))
    return x, int(it), float(rel[0])


def glibc_uniform(n, dim, seed=42):
    """Points as testcases/IMR26/uniform2D.cpp:130-141 draws them:
    srand(seed); (float)rand()/RAND_MAX per coordinate, x then y."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    rnd = libc.rand
    r = np.fromiter((rnd() for _ in range(n * dim)), dtype=np.int64, count=n * dim)
    return (r.astype(np.float32) / np.float32(2147483647)).reshape(n, dim)


def trbvh_test_spheres(n):
    """Spheres as tests/geometry/trbvh_random_collision.cpp:26-34 draws them
    (glibc rand() from its default seed): per sphere x, y = rand()/RAND_MAX,
    r = 0.25 / sqrt(n) * (1 + rand()/RAND_MAX), all in float."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(1))
    r = np.fromiter((libc.rand() for _ in range(3 * n)), dtype=np.int64, count=3 * n).reshape(n, 3)
    u = r.astype(np.float32) / np.float32(2147483647)
    min_r = np.float32(1.0) / np.sqrt(np.float32(n))
    rad = (np.float32(0.25) * min_r) * (np.float32(1.0) + u[:, 2])
    return np.ascontiguousarray(u[:, :2]), rad.astype(np.float32)


class Mesh2D:
    """Oracle Mesh2D: same phases as the reference's Mesh2D."""

    def __init__(self, xy):
        self.L = lib()
        self.xy = np.ascontiguousarray(xy, np.float32)
        self.n = self.xy.shape[0]
        self.h = self.L.orc_mesh2d_create(self.n, _p(self.xy))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_mesh2d_free(self.h)
            self.h = None

    def insert_morton_neighbors(self):
        self.L.orc_mesh2d_insert_morton(self.h)

    def insert_BVH_neighbors(self):
        self.L.orc_mesh2d_insert_bvh(self.h)

    def insert_iterative(self):
        return int(self.L.orc_mesh2d_insert_iterative(self.h))

    def remove_super_nodes(self):
        self.L.orc_mesh2d_remove_super(self.h)

    def get_elem(self):
        k = self.L.orc_mesh2d_nelem(self.h)
        e = np.empty((k, 3), np.uint32)
        self.L.orc_mesh2d_elem(self.h, _p(e))
        return e

    def overflow(self):
        return int(self.L.orc_mesh2d_overflow(self.h))

    def stars(self):
        n = self.n
        r = dict(elem=np.zeros((n, 64), np.uint16), nelem=np.zeros(n, np.uint32),
                 neig=np.zeros((n, 64), np.uint32), nlocal=np.zeros(n, np.uint32), xy=np.zeros((n, 2), np.float32))
        self.L.orc_mesh2d_stars(self.h, _p(r["elem"]), _p(r["nelem"]), _p(r["neig"]), _p(r["nlocal"]), _p(r["xy"]))
        return r

    def cells(self):
        c = np.zeros((2 * self.n - 1, 4), np.float32)
        self.L.orc_mesh2d_cells(self.h, _p(c))
        return c


def super_nodes(xym):
    """The three super-triangle vertices of Mesh::init (mesh.cpp:109-126)."""
    f = np.float32
    lo, hi = xym.min(axis=0), xym.max(axis=0)
    a0 = f(0.5) * (hi[0] + lo[0])
    a1 = f(0.5) * (hi[1] + lo[1])
    mr = f(max(hi[0] - lo[0], hi[1] - lo[1])) + f(10)
    return np.array([[a0 - mr, a1 - mr], [a0 + mr, a1 - mr], [a0, a1 + mr]], np.float32)


def step_solve(coords, alpha, f, rtol=1e-8, ref_rtol=1e-11, maxit=200000):
    """The oracle's timed step (mesh + P1 + FP64 Jacobi-PCG to rtol from 0)
    and an untimed reference solve continued from it to ref_rtol.  Returns
    dict(times_ms=[mesh, assembly, cg], iters, rel, ref_iters, ref_rel,
    triangles, nnz, x_ref (Morton order))."""
    L = lib()
    n = coords.shape[0]
    xyr = spheres(coords, alpha, np.float64)
    fa = np.ascontiguousarray(f, np.float64)
    times = np.zeros(3)
    sizes = np.zeros(2, np.uint64)
    x = np.zeros(n)
    info = np.zeros(4)
    L.orc_step_solve_f64(n, _p(xyr), _p(fa), rtol, ref_rtol, maxit, _p(times), _p(sizes), _p(x), _p(info))
    return dict(times_ms=times.tolist(), iters=int(info[0]), rel=float(info[1]), ref_iters=int(info[2]),
                ref_rel=float(info[3]), triangles=int(sizes[0]), nnz=int(sizes[1]), x_ref=x)

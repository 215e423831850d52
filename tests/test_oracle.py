"""Pins of the CPU oracle (runs without a GPU).

Each test checks the oracle against something that fixes the answer
independently: the reference's published element counts, the reference's
own predicates and Morton encoders compiled from its headers
(oracle/_ref), exact rational arithmetic, an independent recursive
construction of the Karras tree, Qhull's Delaunay triangulation, dense
finite-element assembly from basis-function gradients, and closed-form
solutions.
"""
from __future__ import annotations

import ctypes
import random
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O
import oracle_lib as ol
from stream_amd import clouds

P = ctypes.POINTER


def fptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# Reference golden vector: testcases/IMR26/README.md:49-54 (2D, N=1e6,
# alpha=0.0012, seed 42, FP32) -> 1,877,465 triangles.
# ---------------------------------------------------------------------------
def test_readme_kat_uniform2d_1m():
    pts = O.glibc_uniform(1_000_000, 2, 42)
    r = O.ashape2d(pts, np.full(pts.shape[0], 0.0012, np.float32), bits=32, dtype=np.float32)
    assert r["elem"].shape[0] == 1_877_465


# ---------------------------------------------------------------------------
# Morton: oracle bit loop vs the reference's magic-bit encoders and vs an
# independent Python bit loop.
# ---------------------------------------------------------------------------
def _py_interleave(q, bits):
    code = 0
    for k in range(bits):
        for a, v in enumerate(q):
            code |= ((v >> k) & 1) << (k * len(q) + a)
    return code


def test_morton_interleave_against_reference_and_bitloop():
    L = O.lib()
    R = O.ref()
    rng = random.Random(5)
    for _ in range(500):
        q2 = np.array([rng.getrandbits(32), rng.getrandbits(32)], np.uint32)
        q3 = np.array([rng.getrandbits(21), rng.getrandbits(21), rng.getrandbits(21)], np.uint32)
        c2 = L.orc_interleave(fptr(q2), 2, 32)
        c3 = L.orc_interleave(fptr(q3), 3, 21)
        assert c2 == _py_interleave([int(v) for v in q2], 32)
        assert c3 == _py_interleave([int(v) for v in q3], 21)
        if R is not None:
            assert c2 == R.ref_encode2d(int(q2[0]), int(q2[1]))
            assert c3 == R.ref_encode3d(int(q3[0]), int(q3[1]), int(q3[2]))


def test_morton_quantisation_closed_form():
    # codes of the oracle LBVH equal the closed-form quantisation of
    # lbvh.hpp:186-196 evaluated here in numpy
    pts, alpha = clouds.jittered_square(32, seed=3)
    r = O.lbvh2d(pts, alpha, bits=15)
    lo = (pts - alpha[:, None]).min(axis=0)
    hi = (pts + alpha[:, None]).max(axis=0)
    imax = float(2 ** 15)
    codes = np.zeros(pts.shape[0], np.uint64)
    q = []
    for a in range(2):
        rng_ = hi[a] - lo[a]
        f = imax / rng_
        if rng_ * f >= imax:
            f = np.nextafter(f, 0.0)
        q.append(((pts[:, a] - lo[a]) * f).astype(np.uint32))
    for i in range(pts.shape[0]):
        codes[i] = _py_interleave([int(q[0][i]), int(q[1][i])], 15)
    assert np.array_equal(codes, r["codes"])
    order = np.argsort(codes, kind="stable")
    assert np.array_equal(order.astype(np.uint32), r["perm"])
    assert np.array_equal(codes[order], r["codes_sorted"])


# ---------------------------------------------------------------------------
# Predicates: exact rationals, and the reference's own exact stage.
# ---------------------------------------------------------------------------
def _orient_q(a, b, c):
    ax, ay, bx, by, cx, cy = (Fraction(float(v)) for v in (*a, *b, *c))
    d = (ax - cx) * (by - cy) - (ay - cy) * (bx - cx)
    return (d > 0) - (d < 0)


def _incircle_q(a, b, c, d):
    rows = []
    for p in (a, b, c):
        x, y = Fraction(float(p[0])) - Fraction(float(d[0])), Fraction(float(p[1])) - Fraction(float(d[1]))
        rows.append((x, y, x * x + y * y))
    (a1, a2, a3), (b1, b2, b3), (c1, c2, c3) = rows
    det = a1 * (b2 * c3 - b3 * c2) - a2 * (b1 * c3 - b3 * c1) + a3 * (b1 * c2 - b2 * c1)
    return (det > 0) - (det < 0)


def _near_degenerate_cases(dtype, rng, n):
    """Random quadruples plus nearly cocircular / collinear ones."""
    out = []
    for k in range(n):
        kind = k % 4
        if kind == 0:
            pts = rng.uniform(0, 1, (4, 2))
        elif kind == 1:  # cocircular up to rounding
            th = rng.uniform(0, 2 * np.pi, 4)
            c = rng.uniform(0, 1, 2)
            pts = c + 0.3 * np.stack([np.cos(th), np.sin(th)], axis=1)
        elif kind == 2:  # grid points: exactly cocircular squares
            pts = rng.integers(0, 4, (4, 2)).astype(np.float64) * 0.125
        else:  # nearly collinear
            t = rng.uniform(0, 1, 4)
            pts = np.stack([t, 0.5 * t + 1e-9 * rng.standard_normal(4)], axis=1)
        out.append(np.ascontiguousarray(pts.astype(dtype)))
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_predicates_exact_against_rationals(dtype):
    L = O.lib()
    rng = np.random.default_rng(7)
    tag = "f32" if dtype == np.float32 else "f64"
    o_ex = getattr(L, f"orc_orient2d_exact_{tag}")
    i_ex = getattr(L, f"orc_incircle_exact_{tag}")
    o_fl = getattr(L, f"orc_orient2d_{tag}")
    for pts in _near_degenerate_cases(dtype, rng, 400):
        a, b, c, d = (np.ascontiguousarray(p) for p in pts)
        so = _orient_q(a, b, c)
        si = _incircle_q(a, b, c, d)
        assert o_ex(fptr(a), fptr(b), fptr(c)) == so
        assert np.sign(o_fl(fptr(a), fptr(b), fptr(c))) == so
        assert i_ex(fptr(a), fptr(b), fptr(c), fptr(d)) == si


def test_incircle_sign_convention_unit_circle():
    L = O.lib()
    a, b, c = (np.array(p, np.float64) for p in ((1, 0), (0, 1), (-1, 0)))  # counter-clockwise
    inside = np.array([0.0, 0.0])
    outside = np.array([2.0, 0.0])
    assert L.orc_incircle_exact_f64(fptr(a), fptr(b), fptr(c), fptr(inside)) == 1
    assert L.orc_incircle_exact_f64(fptr(a), fptr(b), fptr(c), fptr(outside)) == -1


def test_predicates_against_reference_fp32():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    L = O.lib()
    rng = np.random.default_rng(11)
    fp = lambda a: a.ctypes.data_as(P(ctypes.c_float))  # noqa: E731
    for pts in _near_degenerate_cases(np.float32, rng, 400):
        a, b, c, d = (np.ascontiguousarray(p) for p in pts)
        assert np.sign(R.ref_orient2d(fp(a), fp(b), fp(c))) == np.sign(L.orc_orient2d_f32(fptr(a), fptr(b), fptr(c)))
        assert np.sign(R.ref_incircleexact(fp(a), fp(b), fp(c), fp(d))) == L.orc_incircle_exact_f32(
            fptr(a), fptr(b), fptr(c), fptr(d))
        e1 = np.zeros(1, np.float32)
        e2 = np.zeros(1, np.float32)
        d1 = R.ref_incircle(fp(a), fp(b), fp(c), fp(d), fp(e1))
        d2 = L.orc_incircle_filter_f32(fptr(a), fptr(b), fptr(c), fptr(d), fptr(e2))
        assert np.float32(d1) == np.float32(d2) and e1[0] == e2[0]


# ---------------------------------------------------------------------------
# Karras topology and refit against an independent recursive construction.
# ---------------------------------------------------------------------------
def _clz64(x):
    return 64 - int(x).bit_length()


def _recursive_karras(codes):
    n = len(codes)
    sep = np.zeros(n - 1, np.int64)
    cl = np.zeros(n - 1, np.int64)
    lp = np.zeros(n, np.int64)
    ip = np.zeros(n - 1, np.int64)
    ip[0] = n
    rng_ = {}
    stack = [(0, n - 1, 0)]
    while stack:
        first, last, idx = stack.pop()
        rng_[idx] = (first, last)
        pref = _clz64(int(codes[first]) ^ int(codes[last]))
        g = first
        for k in range(first, last):
            if _clz64(int(codes[first]) ^ int(codes[k])) > pref:
                g = k
        sep[idx] = g
        if g == first:
            cl[idx] |= 1
            lp[g] = idx
        else:
            ip[g] = idx
            stack.append((first, g, g))
        if g + 1 == last:
            cl[idx] |= 2
            lp[g + 1] = idx | 0x80000000
        else:
            ip[g + 1] = idx | 0x80000000
            stack.append((g + 1, last, g + 1))
    return sep, cl, lp, ip, rng_


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (17, 2), (300, 3), (1000, 4)])
def test_karras_and_refit_against_recursion(n, seed):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, (n, 2))
    r = O.lbvh2d(pts, np.full(n, 0.01), bits=32, leaf_mode=0)
    assert len(np.unique(r["codes_sorted"])) == n
    sep, cl, lp, ip, ranges = _recursive_karras(r["codes_sorted"])
    assert np.array_equal(sep, r["sep"])
    assert np.array_equal(cl, r["child_is_leaf"])
    assert np.array_equal(lp, r["leaf_parent"])
    assert np.array_equal(ip, r["internal_parent"])
    ordered = pts[r["perm"]]
    for idx, (f, l) in ranges.items():
        seg = ordered[f:l + 1]
        assert np.array_equal(r["internal_box"][idx], np.concatenate([seg.min(0), seg.max(0)]))


def test_lbvh_line_of_spheres_kat():
    # tests/geometry/lbvh_simple_collision.cpp:49-57,83-86,162-164
    for dtype in (np.float32, np.float64):
        n = 100
        pts = np.stack([np.arange(n), np.zeros(n)], axis=1).astype(dtype)
        rad = np.full(n, 0.55, dtype)
        r = O.lbvh2d(pts, rad, bits=32, leaf_mode=1, dtype=dtype)
        root = r["internal_box"][0]
        assert root[0] == dtype(-0.55) and root[2] == dtype((n - 1) * 1.0) + dtype(0.55)
        assert root[1] == dtype(-0.55) and root[3] == dtype(0.55)
        assert O.lbvh2d_collisions(pts, rad, dtype=dtype) == 2 * (n - 2) + 2


# ---------------------------------------------------------------------------
# Alpha-shape against Qhull + the alpha filter, and mesh invariants.
# ---------------------------------------------------------------------------
def _canon(tri):
    t = np.sort(np.asarray(tri, np.int64), axis=1)
    return set(map(tuple, t))


def _alpha_filter_np(P, R, tris):
    """Keep Delaunay triangles with R^2 <= min(alpha)^2, evaluated with the
    operation order of DirectAlphaShape2D.cpp:553-566."""
    keep = []
    for t in tris:
        i1, i2, i3 = sorted(int(v) for v in t)
        p1, p2, p3 = P[i1], P[i2], P[i3]
        det = (p3[0] - p2[0]) * (p1[1] - p2[1]) - (p3[1] - p2[1]) * (p1[0] - p2[0])
        n12 = (p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1])
        n23 = (p2[0] - p3[0]) * (p2[0] - p3[0]) + (p2[1] - p3[1]) * (p2[1] - p3[1])
        n31 = (p3[0] - p1[0]) * (p3[0] - p1[0]) + (p3[1] - p1[1]) * (p3[1] - p1[1])
        rsq = ((n12 * n23) * n31) / ((4.0 * det) * det)
        a = min(R[i3], min(R[i1], R[i2]))
        if rsq <= a * a:
            keep.append((i1, i2, i3))
    return keep


@pytest.mark.parametrize("seed,alpha", [(1, 0.03), (2, 0.05), (3, 0.02)])
def test_alpha_shape_against_qhull(seed, alpha):
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, (3000, 2))
    al = np.full(pts.shape[0], alpha)
    r = O.ashape2d(pts, al, bits=32)
    P = r["ordered"]
    dt = Delaunay(P)
    expect = _alpha_filter_np(P, np.full(P.shape[0], alpha), dt.simplices)
    assert _canon(r["elem"]) == _canon(expect)
    # orientation: every output triangle counter-clockwise
    e = r["elem"]
    a, b, c = P[e[:, 0]], P[e[:, 1]], P[e[:, 2]]
    cr = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    assert (cr > 0).all()
    # owner is the lowest id
    assert (e[:, 0] < e[:, 1]).all() and (e[:, 0] < e[:, 2]).all()


def test_alpha_shape_invariants_jittered_square():
    pts, al = clouds.jittered_square(128, seed=0)
    r = O.ashape2d(pts, al, bits=15)
    n = pts.shape[0]
    e = r["elem"].astype(np.int64)
    # edge incidence: each undirected edge in 1 (boundary) or 2 triangles
    edges = {}
    for t in e:
        for u, v in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            k = (min(u, v), max(u, v))
            edges[k] = edges.get(k, 0) + 1
    assert set(edges.values()) <= {1, 2}
    # CSR edges are the triangle edges, flagged boundary iff incidence 1
    row, neig, eb = r["row"], r["neig"], r["edge_is_bnd"]
    for i in range(n):
        for k in range(row[i], row[i + 1]):
            j = int(neig[k])
            cnt = edges[(min(i, j), max(i, j))]
            assert eb[k] == (cnt == 1)
    assert row[n] == 2 * len(edges)
    # boundary nodes: endpoint of a boundary edge or isolated
    on_bnd_edge = np.zeros(n, bool)
    for (u, v), c in edges.items():
        if c == 1:
            on_bnd_edge[u] = on_bnd_edge[v] = True
    deg = np.diff(row.astype(np.int64))
    assert np.array_equal(r["node_is_bnd"].astype(bool), on_bnd_edge | (deg == 0))
    # jittered grid covers the square: Euler characteristic of a disk
    V = len(np.unique(e))
    assert V - len(edges) + e.shape[0] == 1


def test_alpha_shape_exact_path_on_unjittered_grid():
    # an unjittered grid is all cocircular squares: every square goes
    # through the exact stage and the simulation of simplicity; the result
    # must still be a valid triangulation of the square
    h = 1.0 / 24
    g = (np.stack(np.meshgrid(np.arange(24), np.arange(24), indexing="ij"), -1).reshape(-1, 2) + 0.5) * h
    r = O.ashape2d(g, np.full(g.shape[0], 1.3 * h), bits=32)
    assert r["n_exact_calls"] > 0
    assert r["elem"].shape[0] == 2 * 23 * 23


# ---------------------------------------------------------------------------
# P1 assembly: dense gradient assembly, symmetry, patch test.
# ---------------------------------------------------------------------------
def _dense_p1(xy, elem):
    n = xy.shape[0]
    K = np.zeros((n, n))
    M = np.zeros(n)
    for t in elem:
        p = xy[list(t)]
        B = np.array([[p[1, 0] - p[0, 0], p[2, 0] - p[0, 0]], [p[1, 1] - p[0, 1], p[2, 1] - p[0, 1]]])
        area = 0.5 * abs(np.linalg.det(B))
        G = np.linalg.inv(B).T @ np.array([[-1.0, 1.0, 0.0], [-1.0, 0.0, 1.0]])
        K[np.ix_(t, t)] += area * (G.T @ G)
        M[list(t)] += area / 3
    return K, M


def test_p1_square_closed_form():
    xy = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    elem = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    row, col, val, b = O.assemble_p1(xy, elem, np.zeros(4, np.uint8), np.ones(4))
    K = np.zeros((4, 4))
    for i in range(4):
        for k in range(row[i], row[i + 1]):
            K[i, col[k]] = val[k]
    # unit square split along (0,0)-(1,1)
    expect = 0.5 * np.array([[2, -1, 0, -1], [-1, 2, -1, 0], [0, -1, 2, -1], [-1, 0, -1, 2]], float)
    assert np.allclose(K, expect, atol=1e-15)
    assert np.allclose(b, [1 / 3, 1 / 6, 1 / 3, 1 / 6])


def test_p1_against_dense_assembly_and_patch_test():
    pts, al = clouds.jittered_square(24, seed=4)
    r = O.ashape2d(pts, al, bits=32)
    xy = r["ordered"]
    n = xy.shape[0]
    bnd = np.zeros(n, np.uint8)
    row, col, val, b = O.assemble_p1(xy, r["elem"], bnd, np.ones(n))
    K = np.zeros((n, n))
    for i in range(n):
        for k in range(row[i], row[i + 1]):
            K[i, col[k]] = val[k]
        assert np.all(np.diff(col[row[i]:row[i + 1]].astype(np.int64)) > 0)
    Kd, M = _dense_p1(xy, r["elem"].astype(np.int64))
    assert np.allclose(K, Kd, rtol=1e-12, atol=1e-12)
    assert np.array_equal(K, K.T)  # bitwise symmetric
    assert np.allclose(b, M, rtol=1e-14)
    assert np.abs(K.sum(axis=1)).max() < 1e-12
    # linear field is reproduced on interior nodes
    u = 0.3 * xy[:, 0] - 1.7 * xy[:, 1] + 0.25
    interior = r["node_is_bnd"] == 0
    assert np.abs((K @ u)[interior]).max() < 1e-11


# ---------------------------------------------------------------------------
# CG: the diagonal KAT and a direct solve.
# ---------------------------------------------------------------------------
def test_cg_diagonal_kat():
    # tests/numerics/cg_jacobi.cpp:36-83: A = diag(1..100), b = 1
    n = 100
    row = np.arange(n + 1, dtype=np.uint32)
    col = np.arange(n, dtype=np.uint32)
    val = np.arange(1, n + 1, dtype=np.float64)
    x, it, rel = O.jacobi_pcg(row, col, val, np.ones(n))
    assert np.allclose(x, 1.0 / np.arange(1, n + 1), rtol=0, atol=1e-15)
    assert it == 1 and rel <= 1e-8


def test_cg_against_direct_solve_square():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    pts, al = clouds.jittered_square(128, seed=0)
    r = O.ashape2d(pts, al, bits=15)
    xy = r["ordered"]
    n = xy.shape[0]
    row, col, val, b = O.assemble_p1(xy, r["elem"], r["node_is_bnd"], np.ones(n))
    A = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    xd = spl.spsolve(A.tocsc(), b)
    x, it, rel = O.jacobi_pcg(row, col, val, b, rtol=1e-8)
    assert rel <= 1e-8
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-8
    err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    assert err < 1e-6, err
    assert 100 < it < 2000


# ---------------------------------------------------------------------------
# Agglomerative LBVH (include/geometry/lbvh_agglo.hpp)
# ---------------------------------------------------------------------------
def test_lbvha_line_of_spheres_kat():
    """tests/geometry/lbvh_agglo_simple_collision.cpp:48-160: n spheres of
    radius 0.55 on a line; root box [-0.55, n-1+0.55] x [-0.55, 0.55] and
    2(n-2)+2 collisions for every leaf bucket size 1..16."""
    n = 100
    c = np.stack([np.arange(n, dtype=np.float64), np.zeros(n)], 1)
    r = ol.lbvha2d(c, 0.55, leaf_mode=1)
    root = int(r["root"][0])
    assert np.allclose(r["boxes"][root], [-0.55, -0.55, n - 1 + 0.55, 0.55])
    for leaf in (1, 2, 4, 8, 16):
        assert ol.lbvha_collisions(c.astype(np.float32), np.float32(0.55), leaf, dtype=np.float32) == 2 * (n - 2) + 2


@pytest.mark.parametrize("seed", [0, 1])
def test_lbvha_equals_karras_radix_tree(seed):
    """With unique codes the binary radix tree is unique: every Karras
    internal node (range [a, b], split g) appears as agglomerative node g + n
    with the same range, and the boxes agree."""
    rng = np.random.default_rng(seed)
    n = 3000
    c = rng.random((n, 2))
    a = ol.lbvha2d(c, 0.003, leaf_mode=1)
    k = ol.lbvh2d(c, np.full(n, 0.003), bits=32, leaf_mode=1)
    assert np.array_equal(a["perm"], k["perm"])
    codes = k["codes_sorted"]
    assert len(np.unique(codes)) == n
    # Karras ranges by recursion from the root
    sep = k["sep"].astype(np.int64)
    cl = k["child_is_leaf"]
    ranges = {}
    stack = [(0, 0, n - 1)]
    while stack:
        node, lo, hi = stack.pop()
        ranges[node] = (lo, hi)
        g = int(sep[node])
        if not (cl[node] & 1):
            stack.append((g, lo, g))
        if not (cl[node] & 2):
            stack.append((g + 1, g + 1, hi))
    got = {}
    for kk in range(n - 1):
        got[int(a["split"][kk])] = tuple(int(v) for v in a["range"][kk])
    for node, (lo, hi) in ranges.items():
        g = int(sep[node])
        assert got[g] == (lo, hi)
        assert np.allclose(a["boxes"][g + n], k["internal_box"][node])


def test_lbvha_collisions_against_brute_force():
    rng = np.random.default_rng(3)
    n = 2000
    c = rng.random((n, 2))
    rad = 0.01 + 0.01 * rng.random(n)
    d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    rs = rad[:, None] + rad[None, :]
    brute = int(np.sum(d2 <= rs * rs) - n)
    for leaf in (1, 4, 16):
        assert ol.lbvha_collisions(c, rad, leaf) == brute
    c3 = rng.random((n, 3))
    d2 = ((c3[:, None, :] - c3[None, :, :]) ** 2).sum(-1)
    rs3 = 0.04
    brute3 = int(np.sum(d2 <= (2 * rs3) ** 2) - n)
    assert ol.lbvha_collisions(c3, rs3, 4, bits=21) == brute3


def test_vec2_sqnorm_and_cross_against_reference():
    """The alpha filter's squared edge lengths and signed area
    (oracle.hpp: fmac::sumsq and fmac::diff in plain mode) in the operation
    order of the reference's Vec2f::sqnorm and cross (vec.hpp:216-238,
    compiled in oracle/_ref): bit for bit on float inputs across a wide
    range of magnitudes, where any other association or a fused product
    would round differently."""
    R = ol.ref()
    if R is None:
        pytest.skip("reference headers not built (oracle/_ref)")
    L = ol.lib()
    f = ctypes.c_float
    R.ref_vec2_sqnorm.restype = f
    R.ref_vec2_sqnorm.argtypes = [f, f]
    R.ref_vec2_cross.restype = f
    R.ref_vec2_cross.argtypes = [f, f, f, f]
    L.orc_sumsq2_f32.restype = f
    L.orc_sumsq2_f32.argtypes = [f, f]
    L.orc_cross2_f32.restype = f
    L.orc_cross2_f32.argtypes = [f, f, f, f]
    rng = np.random.default_rng(17)
    v = (rng.standard_normal((3000, 4)) * np.exp2(rng.integers(-20, 20, (3000, 4)))).astype(np.float32)
    differs_from_fma = 0
    for ax, ay, bx, by in v:
        s_ref = np.float32(R.ref_vec2_sqnorm(ax, ay))
        assert np.float32(L.orc_sumsq2_f32(ax, ay)).tobytes() == s_ref.tobytes()
        c_ref = np.float32(R.ref_vec2_cross(ax, ay, bx, by))
        assert np.float32(L.orc_cross2_f32(ax, ay, bx, by)).tobytes() == c_ref.tobytes()
        # the same order in numpy float32 (each product rounded, then the sum)
        assert (np.float32(ax * ax) + np.float32(ay * ay)).tobytes() == s_ref.tobytes()
        fused = np.float32(float(ax) * float(by) - np.float64(np.float32(ay * bx)))
        differs_from_fma += fused.tobytes() != c_ref.tobytes()
    assert differs_from_fma > 0  # the check can tell a fused evaluation apart

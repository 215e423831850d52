"""Pins of the Mesh2D oracle (src/mesh/mesh.cpp:45-796, full 2D Delaunay).

* The result equals Qhull's Delaunay triangulation of the nodes plus the
  three super-triangle vertices, minus the triangles on a super vertex (the
  reference's construction: its super-triangle sits at max extent + 10, so
  thin hull triangles whose circumcircle reaches a super vertex are absent,
  exactly as here) -- for the Morton and the BVH-neighbour seeding.
* The star after Morton seeding equals the Delaunay star of the node among
  its inserted neighbours and the super vertices (Qhull on that subset).
* The BVH kd cells tile the global box: each node lies in its own cell and
  the leaf cells' areas sum to the box area.
* On an exact lattice (every 4 neighbours cocircular) no two triangles
  overlap and no node lies strictly inside a circumcircle (exact integer
  arithmetic).  The reference's tie-break (mesh.hpp:311-322) stops after two
  orientation terms and then decides by swap parity; it is not a complete
  symbolic perturbation, and next to the top super vertex a node's whole
  star can fall into one cavity, leaving that node with no triangles (two
  unit squares of a 24 x 24 lattice).  The restatement keeps that.
"""
import numpy as np
import pytest
from scipy.spatial import Delaunay

import oracle_lib as ol


def _tri_set(e):
    return set(map(tuple, np.sort(np.asarray(e, np.int64), axis=1)))


def _qhull_reference(xym):
    n = xym.shape[0]
    allp = np.concatenate([xym, ol.super_nodes(xym)]).astype(np.float64)
    d = Delaunay(allp)
    return {tuple(t) for t in np.sort(d.simplices, axis=1) if t.max() < n}


def _run(xy, seeding):
    m = ol.Mesh2D(xy)
    if seeding == "morton":
        m.insert_morton_neighbors()
    else:
        m.insert_BVH_neighbors()
    m.insert_iterative()
    m.remove_super_nodes()
    return m, m.get_elem()


@pytest.mark.parametrize("n", [2, 3, 10, 100, 1000, 20000])
@pytest.mark.parametrize("seeding", ["morton", "bvh"])
def test_equals_qhull_with_super_vertices(n, seeding):
    xy = np.random.default_rng(n).random((n, 2)).astype(np.float32)
    m, e = _run(xy, seeding)
    assert m.overflow() == 0
    xym = m.stars()["xy"]
    assert _tri_set(e) == _qhull_reference(xym)
    assert len(e) == len(_tri_set(e))


def test_reference_demo_cloud():
    # tests/mesh/mesh_python_api.py:25-35: 100 x 100 grid + N(0, 1e-4) noise
    rs = np.random.RandomState(42)
    t = np.linspace(0, 1, 100, endpoint=True)
    nodes = np.array([[x, y] for x in t for y in t]).astype(np.float32)
    nodes = (nodes + rs.normal(0, 0.0001, size=(100 * 100, 2))).astype(np.float32)
    m, e = _run(nodes, "bvh")
    assert m.overflow() == 0
    assert _tri_set(e) == _qhull_reference(m.stars()["xy"])


def test_morton_star_is_delaunay_of_its_subset():
    n = 3000
    xy = np.random.default_rng(9).random((n, 2)).astype(np.float32)
    m = ol.Mesh2D(xy)
    m.insert_morton_neighbors()
    st = m.stars()
    xym = st["xy"]
    sup = ol.super_nodes(xym)
    for tid in (0, 3, 7, 1500, n - 6, n - 1):
        k = int(st["nlocal"][tid]) - 3
        nb = st["neig"][tid, :k].astype(np.int64)
        ids = np.concatenate([[tid], nb, [n, n + 1, n + 2]])
        pts = np.concatenate([xym[[tid]], xym[nb], sup]).astype(np.float64)
        d = Delaunay(pts)
        want = {tuple(sorted(ids[s])) for s in d.simplices if 0 in s}
        glob = lambda j: int(n + j) if j < 3 else int(st["neig"][tid, j - 3])
        got = set()
        for v in st["elem"][tid, : st["nelem"][tid]]:
            a, b = int(v) & 255, int(v) >> 8
            got.add(tuple(sorted((tid, glob(a), glob(b)))))
        assert got == want, tid


def test_kd_cells_tile_the_box():
    n = 5000
    xy = np.random.default_rng(2).random((n, 2)).astype(np.float32)
    m = ol.Mesh2D(xy)
    c = m.cells().astype(np.float64)
    xym = m.stars()["xy"].astype(np.float64)
    lo, hi = xym.min(axis=0), xym.max(axis=0)
    leaf = c[:n]
    assert np.all(leaf[:, 0] <= xym[:, 0]) and np.all(xym[:, 0] <= leaf[:, 2])
    assert np.all(leaf[:, 1] <= xym[:, 1]) and np.all(xym[:, 1] <= leaf[:, 3])
    area = ((leaf[:, 2] - leaf[:, 0]) * (leaf[:, 3] - leaf[:, 1])).sum()
    assert abs(area - (hi - lo).prod()) < 1e-6
    assert np.all(leaf[:, 2] >= leaf[:, 0]) and np.all(leaf[:, 3] >= leaf[:, 1])


def test_lattice_sos():
    k = 24
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
    for seeding in ("morton", "bvh"):
        m, e = _run(g, seeding)
        assert m.overflow() == 0
        P = m.stars()["xy"].astype(np.int64)
        tri = P[e.astype(np.int64)]
        area2 = np.abs((tri[:, 1, 0] - tri[:, 0, 0]) * (tri[:, 2, 1] - tri[:, 0, 1])
                       - (tri[:, 1, 1] - tri[:, 0, 1]) * (tri[:, 2, 0] - tri[:, 0, 0]))
        assert np.all(area2 > 0)
        # no overlap: every unit square carries at most its two halves
        lo = tri.min(axis=1)
        sq = np.bincount(lo[:, 0] * k + lo[:, 1], minlength=k * k)
        assert sq.max() <= 2 and area2.sum() <= 2 * (k - 1) ** 2
        assert area2.sum() >= 2 * (k - 1) ** 2 - 8
        # empty circumcircles (exact integers, orientation-normalised)
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        orient = np.sign((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
        for p in P[:: 7]:
            rows = []
            for q in (a, b, c):
                dx, dy = q[:, 0] - p[0], q[:, 1] - p[1]
                rows.append((dx, dy, dx * dx + dy * dy))
            (ax, ay, al), (bx, by, bl), (cx, cy, cl) = rows
            det = ax * (by * cl - bl * cy) - ay * (bx * cl - bl * cx) + al * (bx * cy - by * cx)
            assert np.all(det * orient <= 0)

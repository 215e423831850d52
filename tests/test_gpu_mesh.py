"""GPU parity: scan, radix sort, LBVH and the 2D alpha-shape against the
CPU oracle, bit for bit (integer/index outputs and Morton-ordered
coordinates)."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

import oracle_lib as O
from stream_amd import clouds

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _lib():
    import stream_amd

    return stream_amd


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [0, 1, 2, 2047, 2048, 2049, 100_003, 1_000_000])
def test_exclusive_scan(n):
    S = _lib()
    _say(f"scan n={n}")
    rng = np.random.default_rng(n)
    a = rng.integers(0, 300, n).astype(np.uint32)
    out = np.empty(n + 1, np.uint32)
    fn = S.lib.Stream_exclusive_scan_u32
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    S.check(fn(n, a.ctypes.data, out.ctypes.data))
    exp = np.zeros(n + 1, np.uint64)
    exp[1:] = np.cumsum(a, dtype=np.uint64)
    assert np.array_equal(out, exp.astype(np.uint32))
    b = rng.integers(0, 256, n).astype(np.uint8)
    fn8 = S.lib.Stream_exclusive_scan_u8
    fn8.restype = ctypes.c_int32
    fn8.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    S.check(fn8(n, b.ctypes.data, out.ctypes.data))
    exp[1:] = np.cumsum(b, dtype=np.uint64)
    assert np.array_equal(out, exp.astype(np.uint32))


@pytest.mark.parametrize("n,bits,dups", [(1, 64, False), (2, 64, False), (2047, 64, False), (2048, 32, True),
                                         (2049, 64, True), (300_000, 64, True), (1_000_000, 64, False),
                                         (65_537, 16, True)])
def test_radix_sort_pairs_stable(n, bits, dups):
    S = _lib()
    _say(f"sort n={n} bits={bits} dups={dups}")
    rng = np.random.default_rng(n + bits)
    if dups:
        keys = rng.integers(0, 1000, n).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    else:
        keys = rng.integers(0, 2 ** 63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + np.uint64(
            rng.integers(0, 2))
    if bits < 64:
        keys &= np.uint64((1 << bits) - 1)
    vals = np.arange(n, dtype=np.uint32)
    ko = np.empty(n, np.uint64)
    vo = np.empty(n, np.uint32)
    fn = S.lib.Stream_sort_pairs_u64
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
    S.check(fn(n, keys.ctypes.data, vals.ctypes.data, ko.ctypes.data, vo.ctypes.data, bits))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ko, keys[order])
    assert np.array_equal(vo, order.astype(np.uint32))


# ---------------------------------------------------------------------------
# LBVH
# ---------------------------------------------------------------------------
def _lbvh_cases():
    yield "square128_b15", clouds.jittered_square(128, seed=0), 15, np.float64
    yield "square128_b32_f32", clouds.jittered_square(128, seed=0), 32, np.float32
    rng = np.random.default_rng(3)
    p = rng.uniform(-2, 3, (50_001, 2))
    yield "uniform50k", (p, np.full(p.shape[0], 0.01)), 32, np.float64
    yield "dam100k", clouds.dam_break_scaled(100_000), 32, np.float64


@pytest.mark.parametrize("case", list(_lbvh_cases()), ids=lambda c: c[0])
def test_lbvh_matches_oracle(case):
    from stream_amd import geometry

    name, (pts, al), bits, dt = case
    _say(f"lbvh {name}")
    g = geometry.build(pts, al, bits=bits, leaf_mode=0, dtype=dt)
    o = O.lbvh2d(pts, al, bits=bits, leaf_mode=0, dtype=dt)
    codes = geometry.morton_codes(pts, al, bits=bits, dtype=dt)
    assert np.array_equal(codes, o["codes"])
    for k in ("codes_sorted", "perm", "sep", "child_is_leaf", "leaf_parent", "internal_parent"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(g["internal_box"], o["internal_box"])
    assert np.array_equal(g["global_box"], o["global_box"])


def test_lbvh_line_of_spheres_kat():
    from stream_amd import geometry

    _say("lbvh line KAT")
    for dt in (np.float32, np.float64):
        n = 100
        pts = np.stack([np.arange(n), np.zeros(n)], axis=1).astype(dt)
        rad = np.full(n, 0.55, dt)
        g = geometry.build(pts, rad, leaf_mode=geometry.LEAF_SPHERE, dtype=dt)
        root = g["internal_box"][0]
        assert root[0] == dt(-0.55) and root[2] == dt(n - 1) + dt(0.55)
        assert root[1] == dt(-0.55) and root[3] == dt(0.55)
        assert geometry.collisions(pts, rad, dtype=dt) == 2 * (n - 2) + 2


# ---------------------------------------------------------------------------
# 2D alpha-shape
# ---------------------------------------------------------------------------
def _run_gpu_ashape(pts, al, bits, dt):
    from stream_amd.mesh import AlphaShape2D

    a = AlphaShape2D(dtype=dt, morton_bits=bits)
    a.set_nodes(pts, al)
    a.compute()
    row, neig, eb = a.get_edges()
    return dict(elem=a.get_elem(), row=row, neig=neig, edge_is_bnd=eb, node_is_bnd=a.get_boundary_nodes(),
                perm=a.get_permutation(), ordered=a.get_ordered_nodes(), nexact=a.get_nexact())


def _compare(g, o):
    assert g["elem"].shape == o["elem"].shape
    for k in ("perm", "ordered", "node_is_bnd", "row", "neig", "edge_is_bnd", "elem"):
        assert np.array_equal(g[k], o[k]), k


def _ashape_cases():
    yield "square128_b15_f64", clouds.jittered_square(128, seed=0), 15, np.float64
    yield "dam100k_f64", clouds.dam_break_scaled(100_000), 32, np.float64
    rng = np.random.default_rng(9)
    p = rng.uniform(0, 1, (20_000, 2))
    yield "uniform20k_varalpha_f64", (p, rng.uniform(0.005, 0.02, p.shape[0])), 32, np.float64
    h = 1.0 / 40
    g = (np.stack(np.meshgrid(np.arange(40), np.arange(40), indexing="ij"), -1).reshape(-1, 2) + 0.5) * h
    yield "grid40_cocircular_f64", (g, np.full(g.shape[0], 1.3 * h)), 32, np.float64
    yield "grid40_cocircular_f32", (g.astype(np.float32), np.full(g.shape[0], 1.3 * h, np.float32)), 32, np.float32


@pytest.mark.parametrize("case", list(_ashape_cases()), ids=lambda c: c[0])
def test_ashape2d_matches_oracle(case):
    name, (pts, al), bits, dt = case
    _say(f"ashape2d {name}")
    g = _run_gpu_ashape(pts, al, bits, dt)
    o = O.ashape2d(pts, al, bits=bits, dtype=dt)
    _compare(g, o)
    if name.startswith("grid40"):
        assert g["nexact"] > 0


def test_ashape2d_readme_kat_fp32_1m():
    # testcases/IMR26/README.md:49-54: 1,877,465 triangles; full outputs
    # compared with the oracle
    _say("ashape2d README KAT 1M fp32")
    pts = O.glibc_uniform(1_000_000, 2, 42)
    al = np.full(pts.shape[0], 0.0012, np.float32)
    g = _run_gpu_ashape(pts, al, 32, np.float32)
    assert g["elem"].shape[0] == 1_877_465
    o = O.ashape2d(pts, al, bits=32, dtype=np.float32)
    _compare(g, o)


def test_ashape2d_dam_break_1m_fp64():
    _say("ashape2d dam-break 1M fp64")
    pts, al = clouds.dam_break()
    g = _run_gpu_ashape(pts, al, 32, np.float64)
    o = O.ashape2d(pts, al, bits=32, dtype=np.float64)
    _compare(g, o)


def test_ashape2d_repeat_is_deterministic():
    from stream_amd.mesh import AlphaShape2D

    _say("ashape2d repeat")
    pts, al = clouds.dam_break_scaled(200_000)
    a = AlphaShape2D(dtype=np.float64)
    a.set_nodes(pts, al)
    a.compute()
    e1 = a.get_elem()
    a.init()
    a.compute()
    e2 = a.get_elem()
    assert np.array_equal(e1, e2)


# ---------------------------------------------------------------------------
# Agglomerative LBVH
# ---------------------------------------------------------------------------
def test_lbvha_gpu_equals_oracle():
    from stream_amd import geometry

    rng = np.random.default_rng(21)
    n = 200_000
    c = rng.random((n, 2))
    g = geometry.build_agglomerative(c, 0.002, bits=32, leaf_mode=geometry.LEAF_SPHERE)
    o = O.lbvha2d(c, 0.002, bits=32, leaf_mode=1)
    for k in ("perm", "children", "range", "parent", "root", "split"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(g["boxes"], o["boxes"])


def test_lbvha_gpu_collision_kats():
    from stream_amd import geometry

    n = 100
    c = np.stack([np.arange(n, dtype=np.float32), np.zeros(n, np.float32)], 1)
    for leaf in (1, 2, 4, 8, 16):
        assert geometry.collisions_agglomerative(c, np.float32(0.55), leaf, bits=32, dtype=np.float32) == 2 * n - 2
    rng = np.random.default_rng(4)
    c3 = rng.random((50_000, 3))
    got = geometry.collisions_agglomerative(c3, 0.01, 4, bits=21)
    assert got == O.lbvha_collisions(c3, 0.01, 4, bits=21)


# ---------------------------------------------------------------------------
# TRBVH
# ---------------------------------------------------------------------------
TRBVH_KEYS = ("perm", "children", "range", "parent", "root", "boxes", "sah", "subtree")


@pytest.mark.parametrize("case", ["ref100", "ref5000", "rand2d_200k", "rand3d_50k"])
@pytest.mark.parametrize("rounds", [0, 1, 3])
def test_trbvh_gpu_equals_oracle(case, rounds):
    """Topology, boxes, costs and subtree counts bit-equal to the oracle."""
    from stream_amd import geometry

    _say(f"trbvh {case} rounds={rounds}")
    if case.startswith("ref"):
        c, r = O.trbvh_test_spheres(int(case[3:]))
    else:
        n = 200_000 if case == "rand2d_200k" else 50_000
        d = 2 if case.startswith("rand2d") else 3
        rng = np.random.default_rng(n + d)
        c = rng.random((n, d)).astype(np.float32)
        r = (np.float32(0.25) / np.sqrt(np.float32(n)) * (1 + rng.random(n))).astype(np.float32)
    g = geometry.build_trbvh(c, r, rounds)
    o = O.trbvh(c, r, rounds)
    for k in TRBVH_KEYS:
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("n", [100, 1000, 100_000])
def test_trbvh_gpu_collisions(n):
    from stream_amd import geometry

    _say(f"trbvh collisions {n}")
    c, r = O.trbvh_test_spheres(n)
    assert geometry.collisions_trbvh(c, r, 3) == O.trbvh_collisions(c, r, 3)
    c3 = np.random.default_rng(n).random((n, 3)).astype(np.float32)
    r3 = np.full(n, 0.6 / n ** (1 / 3), np.float32)
    assert geometry.collisions_trbvh(c3, r3, 3) == O.trbvh_collisions(c3, r3, 3)


def test_trbvh_gpu_1m_timing_and_structure():
    """Reference test cloud at 1M spheres: the tree is complete and
    consistent; prints the GPU build / restructuring times."""
    from stream_amd import geometry

    n = 1_000_000
    _say("trbvh 1M")
    rng = np.random.default_rng(11)
    c = rng.random((n, 2)).astype(np.float32)
    r = (np.float32(0.25) / np.sqrt(np.float32(n)) * (1 + rng.random(n))).astype(np.float32)
    g = geometry.build_trbvh(c, r, 3)
    geometry.build_trbvh(c, r, 3)
    g = geometry.build_trbvh(c, r, 3)
    print(f"trbvh 1M: build {g['ms'][0]:.3f} ms, 3 rounds {g['ms'][1]:.3f} ms", flush=True)
    root = int(g["root"][0])
    assert g["parent"][root] == 2 * n and g["subtree"][root] == n
    ch = g["children"]
    inner = np.arange(n, 2 * n - 1)
    assert np.array_equal(g["parent"][ch[:, 0]], inner) and np.array_equal(g["parent"][ch[:, 1]], inner)
    assert np.array_equal(g["subtree"][inner], g["subtree"][ch[:, 0]] + g["subtree"][ch[:, 1]])
    assert np.array_equal(g["boxes"][inner, :2], np.minimum(g["boxes"][ch[:, 0], :2], g["boxes"][ch[:, 1], :2]))
    assert np.array_equal(g["boxes"][inner, 2:], np.maximum(g["boxes"][ch[:, 0], 2:], g["boxes"][ch[:, 1], 2:]))
    seen = np.bincount(ch.ravel(), minlength=2 * n - 1)
    seen[root] += 1
    assert np.all(seen == 1)
    o = O.trbvh(c, r, 3)
    for k in TRBVH_KEYS:
        assert np.array_equal(g[k], o[k]), k


def _heap_ref(items):
    """The reference's sift order (traversal_stack.hpp:50-90) in Python."""
    h = []
    for it in items:
        if len(h) >= 32:
            continue
        h.append(it)
        i = len(h) - 1
        p = (i - 1) >> 1
        while i > 0 and h[p][1] > h[i][1]:
            h[i], h[p] = h[p], h[i]
            i = p
            p = (i - 1) >> 1
    out = []
    while h:
        out.append(h[0])
        h[0] = h[-1]
        h.pop()
        i = 0
        while True:
            l, r, s = 2 * i + 1, 2 * i + 2, i
            if l < len(h) and h[l][1] < h[s][1]:
                s = l
            if r < len(h) and h[r][1] < h[s][1]:
                s = r
            if s == i:
                break
            h[i], h[s] = h[s], h[i]
            i = s
    return out


def test_traversal_min_heap_device():
    import ctypes

    import stream_amd

    fn = stream_amd.lib.Stream_heap_order
    fn.restype = ctypes.c_int32
    # tests/geometry/traversal_stack.cpp: costs 10..0 pushed, popped ascending
    cases = [[(k, float(k)) for k in range(10, -1, -1)]]
    rng = np.random.default_rng(2)
    cases.append([(k, float(v)) for k, v in enumerate(rng.integers(0, 5, 40))])  # ties + overflow
    for items in cases:
        n = len(items)
        idx = np.array([a for a, _ in items], np.uint32)
        cost = np.array([b for _, b in items], np.float32)
        oi = np.zeros(32, np.uint32)
        oc = np.zeros(32, np.float32)
        pushed = ctypes.c_int32(0)
        stream_amd.check(fn(n, idx.ctypes.data_as(ctypes.c_void_p), cost.ctypes.data_as(ctypes.c_void_p),
                            oi.ctypes.data_as(ctypes.c_void_p), oc.ctypes.data_as(ctypes.c_void_p),
                            ctypes.byref(pushed)))
        ref = _heap_ref(items)
        assert pushed.value == len(ref)
        assert [(int(a), float(b)) for a, b in zip(oi[:len(ref)], oc[:len(ref)])] == ref
    assert [a for a, _ in _heap_ref(cases[0])] == list(range(11))

"""Pins of the TRBVH oracle (include/geometry/trbvh.hpp:64-681).

* tree invariants after 0, 1 and 3 restructuring rounds (every node reached
  once from the root, parent links, boxes = exact union of the children,
  subtree counts = sums, leaf boxes = sphere boxes);
* overlap counts = brute force on the reference test's own sphere cloud
  (tests/geometry/trbvh_random_collision.cpp:26-34) and on 3D clouds;
* cost known answers on 2 and 3 spheres, worked by hand, that fix which
  child's running cost is carried (a wrong carrier changes the value);
* the subset programme = the minimum over all 105 binary trees on 5 leaves
  (explicit enumeration), and its tie rule (first partition in the
  enumeration order) on leaves that all tie.
The restructuring's orientation (the chosen partition becomes the right
child, trbvh.hpp:567-571) is fixed by reading the reference; no output of
the reference's test observes it.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as ol

F = np.float32
CI, CT = F(1.2), F(1.0)


def check_tree(t, coords, radii):
    n = coords.shape[0]
    d = coords.shape[1]
    ch, par, box, sub = t["children"], t["parent"], t["boxes"], t["subtree"]
    root = int(t["root"][0])
    assert par[root] == 2 * n
    seen = np.zeros(2 * n - 1, int)
    obj = coords[t["perm"]]
    rad = radii[t["perm"]]
    st = [root]
    while st:
        v = st.pop()
        seen[v] += 1
        if v < n:
            assert np.array_equal(box[v, :d], (obj[v] - rad[v]).astype(F))
            assert np.array_equal(box[v, d:], (obj[v] + rad[v]).astype(F))
            assert sub[v] == 1
            continue
        l, r = ch[v - n]
        assert par[l] == v and par[r] == v
        assert np.array_equal(box[v, :d], np.minimum(box[l, :d], box[r, :d]))
        assert np.array_equal(box[v, d:], np.maximum(box[l, d:], box[r, d:]))
        assert sub[v] == sub[l] + sub[r]
        st += [int(l), int(r)]
    assert np.all(seen == 1)


def brute_overlaps(coords, radii):
    c = coords.astype(F)
    r = radii.astype(F)
    total = 0
    for i in range(c.shape[0]):
        dd = c - c[i]
        d2 = (dd * dd).sum(axis=1, dtype=F)
        rs = r + r[i]
        total += int(np.count_nonzero(d2 <= rs * rs)) - 1
    return total


def random_spheres(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, d)).astype(F)
    r = (F(0.25) / np.sqrt(F(n)) * (1 + rng.random(n))).astype(F)
    return c, r


@pytest.mark.parametrize("n", [2, 3, 5, 8, 17, 100, 1000, 4000])
@pytest.mark.parametrize("rounds", [0, 1, 3])
def test_tree_invariants_2d(n, rounds):
    c, r = random_spheres(n, 2, n)
    check_tree(ol.trbvh(c, r, rounds), c, r)


@pytest.mark.parametrize("n", [2, 9, 300, 3000])
def test_tree_invariants_3d(n):
    c, r = random_spheres(n, 3, 7 + n)
    for rounds in (0, 3):
        check_tree(ol.trbvh(c, r, rounds), c, r)


@pytest.mark.parametrize("n", [100, 1000])
def test_reference_cloud_overlaps(n):
    c, r = ol.trbvh_test_spheres(n)
    assert ol.trbvh_collisions(c, r, 3) == brute_overlaps(c, r)
    assert ol.trbvh_collisions(c, r, 0) == brute_overlaps(c, r)


def test_overlaps_3d():
    c, r = random_spheres(2000, 3, 5)
    r *= 3
    assert ol.trbvh_collisions(c, r, 3) == brute_overlaps(c, r)


def test_cost_two_spheres():
    # Boxes [-0.5,0.5]^2 (area 1) and [1.75,2.25]x[0.75,1.25] (area 0.25);
    # root box 2.75 x 1.75.  The right leaf carries Ct * 0.25.
    c = np.array([[0, 0], [2, 1]], F)
    r = np.array([0.5, 0.25], F)
    t = ol.trbvh(c, r, 0)
    root = int(t["root"][0])
    assert list(t["perm"]) == [0, 1]
    want = F(F(0.25) * CT) + F(F(CI * F(2.75)) * F(1.75))
    assert t["sah"][root] == want
    assert t["sah"][root] != F(1.0) + F(F(CI * F(2.75)) * F(1.75))  # left carrier
    assert t["sah"][0] == 0 and t["sah"][1] == 0


def test_cost_three_spheres():
    # Leaves 0 | (1, 2): node (1,2) has box [2.75,3.75]^2 and cost
    # Ct*0.25 (leaf 2 carries) + Ci*1*1; the root (box 4.25^2) adds Ci*4.25^2
    # to that carried cost, plus 0 for leaf 0.
    c = np.array([[0, 0], [3, 3], [3.5, 3.5]], F)
    r = np.array([0.5, 0.25, 0.25], F)
    t = ol.trbvh(c, r, 0)
    n = 3
    root = int(t["root"][0])
    l, rr = t["children"][root - n]
    assert l == 0 and rr >= n and list(t["children"][rr - n]) == [1, 2]
    inner = F(F(0.5) * F(0.5)) + F(F(CI * F(1.0)) * F(1.0))
    assert t["sah"][rr] == inner
    assert t["sah"][root] == inner + F(F(CI * F(4.25)) * F(4.25))
    assert list(t["subtree"][[root, rr]]) == [3, 2]


def _trees(leaves):
    if len(leaves) == 1:
        yield leaves[0]
        return
    first, rest = leaves[0], leaves[1:]
    for k in range(0, len(rest)):
        for other in itertools.combinations(rest, k):
            a = (first,) + other
            b = tuple(x for x in rest if x not in other)
            for ta in _trees(a):
                for tb in _trees(b):
                    yield (ta, tb)


def _tree_cost(t, boxes, cost, size):
    """(cost, leaf set) of an explicit tree under c = min(Ci A + cl + cr,
    Ct A N), float arithmetic in the optimiser's order."""
    if isinstance(t, int):
        return cost[t], (t,)
    cl, sl = _tree_cost(t[0], boxes, cost, size)
    cr, sr = _tree_cost(t[1], boxes, cost, size)
    s = sl + sr
    lo = boxes[list(s), :2].min(axis=0)
    hi = boxes[list(s), 2:].max(axis=0)
    area = F(F(hi[0] - lo[0]) * F(hi[1] - lo[1]))
    nobj = F(sum(int(size[i]) for i in s))
    split = F(F(CI * area) + F(cl + cr))
    flat = F(F(CT * area) * nobj)
    return min(split, flat), s


def test_treelet_dp_is_minimum_over_all_trees():
    trees = list(_trees((0, 1, 2, 3, 4)))
    assert len(trees) == 105
    rng = np.random.default_rng(3)
    for trial in range(40):
        lo = rng.random((5, 2)).astype(F)
        ext = (rng.random((5, 2)) * (0.05 if trial % 2 else 0.6)).astype(F)
        boxes = np.concatenate([lo, lo + ext], axis=1).astype(F)
        cost = (rng.random(5) * (0 if trial % 3 == 0 else 2)).astype(F)
        size = rng.integers(1, 40, 5).astype(np.uint32)
        copt, popt, sz = ol.treelet_dp(boxes, cost, size)
        best = min(_tree_cost(t, boxes, cost, size)[0] for t in trees)
        assert copt[31] == best
        assert sz[31] == size.sum()

        # the tree the partitions describe realises the optimum
        def build(s):
            if bin(s).count("1") == 1:
                return s.bit_length() - 1
            p = int(popt[s])
            assert p and (p & s) == p and p != s
            return (build(s & ~p), build(p))

        assert _tree_cost(build(31), boxes, cost, size)[0] == copt[31]


def test_treelet_dp_tie_rule():
    # Identical leaves: every partition of a subset ties; the first one in
    # the enumeration is the lowest member of s & (s - 1).
    boxes = np.tile(np.array([[0, 0, 1, 1]], F), (5, 1))
    copt, popt, _ = ol.treelet_dp(boxes, np.zeros(5, F), np.ones(5, np.uint32))
    for s in range(32):
        if bin(s).count("1") >= 2:
            d = s & (s - 1)
            assert popt[s] == d & -d

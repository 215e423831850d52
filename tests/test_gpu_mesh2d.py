"""GPU Mesh2D (full 2D Delaunay) against the oracle: stars bit-equal after
every phase, triangles equal as arrays, for random clouds, the reference's
demo cloud (tests/mesh/mesh_python_api.py:25-35), an exact lattice (tie
breaks) and 1M nodes."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _stars_equal(g, o, what):
    n = len(o["nelem"])
    assert np.array_equal(g["nelem"], o["nelem"]), f"{what}: nelem"
    assert np.array_equal(g["nlocal"], o["nlocal"]), f"{what}: nlocal"
    mask = np.arange(64)[None, :] < o["nelem"][:, None]
    assert np.array_equal(np.where(mask, g["elem"], 0), np.where(mask, o["elem"], 0)), f"{what}: elem"
    mask = np.arange(64)[None, :] < (o["nlocal"][:, None].astype(np.int64) - 3)
    assert np.array_equal(np.where(mask, g["neig"], 0), np.where(mask, o["neig"], 0)), f"{what}: neig"
    assert n == len(g["nelem"])


def _clouds():
    rng = np.random.default_rng(17)
    rs = np.random.RandomState(42)
    t = np.linspace(0, 1, 100, endpoint=True)
    demo = np.array([[x, y] for x in t for y in t]).astype(np.float32)
    demo = (demo + rs.normal(0, 0.0001, size=(100 * 100, 2))).astype(np.float32)
    k = 40
    lattice = np.stack(np.meshgrid(np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
    return {
        "rand2": rng.random((2, 2)).astype(np.float32),
        "rand100": rng.random((100, 2)).astype(np.float32),
        "rand50k": rng.random((50_000, 2)).astype(np.float32),
        "demo10k": demo,
        "lattice40": lattice,
    }


CLOUDS = _clouds()


@pytest.mark.parametrize("cloud", list(CLOUDS))
@pytest.mark.parametrize("seeding", ["morton", "bvh"])
def test_mesh2d_gpu_equals_oracle(cloud, seeding):
    from stream_amd.mesh import Mesh2D

    _say(f"mesh2d {cloud} {seeding}")
    xy = CLOUDS[cloud]
    g = Mesh2D()
    g.set_nodes(xy)
    o = O.Mesh2D(xy)
    assert np.array_equal(g.get_ordered_nodes(), o.stars()["xy"])
    if seeding == "morton":
        g.insert_morton_neighbors()
        o.insert_morton_neighbors()
    else:
        g.insert_BVH_neighbors()
        o.insert_BVH_neighbors()
        assert np.array_equal(g.get_cells(), o.cells())
    _stars_equal(g.get_stars(), o.stars(), "seeding")
    g.insert_iterative()
    rounds = o.insert_iterative()
    st = g.get_stats()
    assert st["overflow"] == 0 and o.overflow() == 0
    assert st["rounds"] == rounds
    _stars_equal(g.get_stars(), o.stars(), "iterative")
    g.remove_super_nodes()
    o.remove_super_nodes()
    _stars_equal(g.get_stars(), o.stars(), "remove")
    eg, eo = g.get_elem(), o.get_elem()
    assert np.array_equal(eg, eo)
    # get_nelem compresses again: same result
    assert np.array_equal(g.get_elem(), eo)


def test_mesh2d_gpu_1m():
    from stream_amd.mesh import Mesh2D

    _say("mesh2d 1M")
    xy = np.random.default_rng(5).random((1_000_000, 2)).astype(np.float32)
    g = Mesh2D()
    g.set_nodes(xy)
    g.insert_BVH_neighbors()
    g.insert_iterative()
    g.remove_super_nodes()
    eg = g.get_elem()
    st = g.get_stats()
    print(f"mesh2d 1M: {len(eg)} triangles, stats {st}", flush=True)
    o = O.Mesh2D(xy)
    o.insert_BVH_neighbors()
    o.insert_iterative()
    o.remove_super_nodes()
    assert st["overflow"] == 0 and o.overflow() == 0
    assert np.array_equal(eg, o.get_elem())


def test_mesh2d_errors():
    from stream_amd import StreamError
    from stream_amd.mesh import Mesh2D

    g = Mesh2D()
    with pytest.raises(StreamError):
        g.insert_iterative()  # before set_nodes / init
    with pytest.raises(StreamError):
        g.set_nodes(np.zeros((1, 2), np.float32))

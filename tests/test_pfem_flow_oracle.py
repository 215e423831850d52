"""The oracle's PFEM projection step (oracle/pfem_flow.py) pinned by closed
forms on an alpha-shape mesh of the oracle: the lumped mass sums to the
mesh area; D of a linear velocity is m_i div u at every node; G of a
linear pressure is m_i grad p at every node; K of a linear field vanishes
on interior nodes (patch test) and K equals the assembled P1 matrix before
its Dirichlet rows; the correction with the exact pressure of a potential
flow removes it."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
from oracle import pfem_flow as F
from stream_amd import clouds


@pytest.fixture(scope="module")
def mesh():
    pts, al = clouds.jittered_square(40, seed=5)
    r = O.ashape2d(pts, al, bits=32, dtype=np.float64)
    return r["ordered"], r["elem"].astype(np.int64), r["node_is_bnd"]


def test_mass_sums_to_area(mesh):
    xy, elem, _ = mesh
    area, _, _ = F.geometry(xy, elem)
    m = F.lumped_mass(xy, elem)
    assert abs(m.sum() - area.sum()) < 1e-14
    # a unit right triangle: each vertex gets 1/6
    assert np.allclose(F.lumped_mass(np.array([[0, 0], [1, 0], [0, 1.0]]), np.array([[0, 1, 2]])), 1 / 6)


def test_divergence_and_gradient_of_linear_fields(mesh):
    xy, elem, _ = mesh
    m = F.lumped_mass(xy, elem)
    A = np.array([[0.3, -1.2], [0.7, 0.5]])  # u = A x + c, div u = trace A
    u = xy @ A.T + np.array([0.1, -0.2])
    assert np.allclose(F.divergence(xy, elem, u), m * np.trace(A), rtol=0, atol=1e-13)
    gp = np.array([2.0, -3.5])
    p = xy @ gp + 0.25
    assert np.allclose(F.gradient(xy, elem, p), m[:, None] * gp[None, :], rtol=0, atol=1e-13)


def test_stiffness_patch_test_and_assembled_matrix(mesh):
    xy, elem, bnd = mesh
    n = len(xy)
    lin = xy @ np.array([0.8, -0.3]) + 1.0
    Ku = F.stiffness_apply(xy, elem, lin)
    # interior nodes of a complete star: the linear field is in the kernel
    row, col, val, _ = O.assemble_p1(xy, elem.astype(np.uint32), np.zeros(n, np.uint8), np.ones(n))
    K = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    assert np.allclose(K @ lin, Ku, rtol=0, atol=1e-12)
    inner = np.flatnonzero(bnd == 0)
    assert np.max(np.abs(Ku[inner])) < 1e-12


def test_projection_removes_a_potential_flow(mesh):
    """u = grad phi with phi = x^2 + y^2 (div u = 4): u* = u (g = nu = 0);
    the pressure solves Laplace(p) = 4 rho / dt with p = 0 on the free
    surface, so p < 0 inside, and the (approximate, lumped) projection
    removes most of the discrete divergence: to 10% of it on nodes not next
    to the free surface, 20% on all free nodes."""
    import scipy.sparse.linalg as spl

    xy, elem, bnd = mesh
    n = len(xy)
    u = 2.0 * xy
    dt, rho = 0.1, 1.0
    us, m = F.predictor(xy, elem, u, dt, 0.0, (0.0, 0.0))
    assert np.array_equal(us, u)
    b = F.pressure_rhs(xy, elem, bnd, us, dt, rho)
    row, col, val, _ = O.assemble_p1(xy, elem.astype(np.uint32), bnd, np.ones(n))
    K = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    p = spl.spsolve(K.tocsc(), b)
    assert np.all(p[bnd != 0] == 0.0)
    assert np.all(p[bnd == 0] < 0.0)
    un = F.correct(xy, elem, us, m, p, dt, rho)
    near = np.zeros(n, bool)
    for k in range(3):
        for j in range(3):
            near[elem[:, k]] |= bnd[elem[:, j]] != 0
    for idx, bound in ((np.flatnonzero(bnd == 0), 0.25), (np.flatnonzero(~near), 0.15)):
        d0 = np.linalg.norm(F.divergence(xy, elem, us)[idx])
        d1 = np.linalg.norm(F.divergence(xy, elem, un)[idx])
        assert d1 < bound * d0, (len(idx), d1 / d0)

"""The PFEM projection (fractional) velocity step on the GPU against the
oracle's restatement (oracle/pfem_flow.py) on the step's own mesh: u*, the
pressure right-hand side, the corrected velocity (with the GPU's pressure)
and its discrete divergence within 1e-12 relative (each node sums its own
triangles on the GPU, the oracle loops over elements); the pressure solve
to a true FP64 relative residual of 1e-8; the Lagrangian move; and the 1M
dam-break with the 8-mode Fourier velocity, solved by the learned and the
FP64 Jacobi PCG."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pfem_flow as F
from stream_amd import clouds

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _velocity(xy):
    x, y = xy[:, 0], xy[:, 1]
    return np.stack([np.sin(np.pi * x) * np.cos(np.pi * y) + 0.3 * y,
                     -0.5 * np.cos(np.pi * x) * np.sin(np.pi * y) + 0.2 * x * x], 1)


@pytest.mark.parametrize("case", ["square64", "dam20k"])
def test_projection_step_matches_oracle(case):
    from stream_amd.pfem import PfemStep2D

    _say(f"projection step {case}")
    if case == "square64":
        pts, al = clouds.jittered_square(64, seed=3)
    else:
        pts, al = clouds.dam_break_scaled(20_000)
    n = len(pts)
    st = PfemStep2D(pts, al, np.zeros(n))
    u = _velocity(pts)
    st.set_velocity(u)
    dt, rho, nu, g = 1e-2, 1000.0, 1e-3, (0.0, -9.81)
    r = st.projection_step("fp64", dt=dt, rho=rho, nu=nu, g=g, move=False)
    elem, bnd, xy, perm = st.mesh(r["n_elems"])
    elem = elem.astype(np.int64)
    us, un, m, div = st.flow_fields()
    us_o, m_o = F.predictor(xy, elem, u[perm], dt, nu, g)
    assert _rel(m, m_o) < 1e-13
    assert _rel(us, us_o) < 1e-12
    row, col, val, b = st.system()
    b_o = F.pressure_rhs(xy, elem, bnd, us_o, dt, rho)
    assert _rel(b, b_o) < 1e-12
    p = st.solution(morton=True)
    A = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    assert np.linalg.norm(b - A @ p) / np.linalg.norm(b) <= 1e-8
    un_o = F.correct(xy, elem, us, m, p, dt, rho)
    assert _rel(un, un_o) < 1e-12
    assert _rel(div, F.divergence(xy, elem, un)) < 1e-12
    assert np.array_equal(st.velocity(), st.permute_backward(un))
    inner = np.flatnonzero(bnd == 0)
    d0 = np.linalg.norm(F.divergence(xy, elem, us)[inner])
    print(f"{case}: {r['iters']} it, rel {r['true_rel']:.1e}, D(u*) {d0:.3e} -> D(u_new) "
          f"{np.linalg.norm(div[inner]):.3e} on free nodes, flow kernels {r['flow_ms']:.3f} ms", flush=True)
    assert np.linalg.norm(div[inner]) < 0.5 * d0


def test_projection_step_moves_points():
    from stream_amd.pfem import PfemStep2D

    _say("projection step move")
    pts, al = clouds.jittered_square(48, seed=6)
    st = PfemStep2D(pts, al, np.zeros(len(pts)))
    st.set_velocity(_velocity(pts))
    dt = 5e-3
    x0 = st.points()
    st.projection_step("mixed", dt=dt, g=(0.0, -1.0), move=True)
    v = st.velocity()
    assert np.array_equal(st.points(), x0 + dt * v)
    # the next step meshes the moved points
    r = st.projection_step("mixed", dt=dt, g=(0.0, -1.0), move=True)
    assert r["true_rel"] <= 1e-8


def test_dam_break_1m_projection_learned_and_fp64():
    from stream_amd.learned import PrecLearned
    from stream_amd.pfem import PfemStep2D

    _say("1M dam-break projection step")
    pts, al = clouds.dam_break()
    u = clouds.fourier_velocity(pts, clouds.fourier_modes(11))
    st = PfemStep2D(pts, al, np.zeros(len(pts)))
    wts = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "stream_amd", "weights",
                       "mpnn_dam2d")
    st.set_learned(PrecLearned(PrecLearned.load_params(wts)))
    for solver in ("learned", "fp64"):
        st.set_velocity(u)
        r = st.projection_step(solver, dt=1e-3, rho=1.0, nu=0.0, g=(0.0, 0.0), move=False)
        _, _, _, div = st.flow_fields()
        print(f"1M projection {solver}: {r['iters']} it, rel {r['true_rel']:.2e}, solve {r['solve_ms']:.1f} ms, "
              f"velocity kernels {r['flow_ms']:.3f} ms", flush=True)
        assert r["true_rel"] <= 1e-8 and not r["fallback"]

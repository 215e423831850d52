"""Pins of the learned-preconditioner oracle (oracle/mpnn.py) on the CPU.

The network has no counterpart in the reference (its only preconditioner is
Jacobi, src/numerics/preconditioner_jacobi.cpp:29-60), so the oracle is pinned
by mathematics: finite differences of every parameter gradient, the
Jacobi closed form at zero output weights, an independent dense loss, and
Kaporin's theorem (the loss minimiser over the pattern is the FSAI factor).
"""
import numpy as np
import pytest
import scipy.optimize as so
import scipy.sparse as sp

import oracle_lib as ol
from oracle import mpnn
from stream_amd import clouds


def _system(n_side=20, seed=0, source="ones"):
    pts, al = clouds.jittered_square(n_side, seed=seed)
    r = ol.ashape2d(pts, al, bits=32)
    xy = r["ordered"]
    f = np.ones(len(pts)) if source == "ones" else clouds.divergence_source(xy, clouds.fourier_modes(11))
    row, col, val, b = ol.assemble_p1(xy, r["elem"], r["node_is_bnd"], f)
    return row, col, val, b, xy, r["node_is_bnd"]


@pytest.fixture(scope="module")
def graph():
    row, col, val, b, xy, bnd = _system()
    return mpnn.Graph(row, col, val, xy, bnd)


def _random_params(seed=3):
    p = mpnn.init_params(7)
    rng = np.random.default_rng(seed)
    P = mpnn.unflatten(p)
    for k in ("o_w", "d_w", "de_b", "dd_b", "m_b0", "u_b1"):
        P[k][:] = 0.03 * rng.standard_normal(P[k].shape)
    return mpnn.flatten(P)


def test_param_count():
    # 5x64+64 + 3 x (64x128 + 3x64 + 64 + 128x64 + 64) + 64x192 + 3x64 + 4x64
    assert mpnn.n_params() == 384 + 3 * 16704 + 12288 + 192 + 256


def test_bf16_rounding_kat():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 3.0e38, 1.0 + 2.0 ** -9])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 2.9908441e38, 1.0])
    got = mpnn.bf16(x)
    assert got[:5].tolist() == want[:5].tolist()
    assert got[6] == 1.0
    # every result is a bf16 value: low 16 bits of the float32 pattern clear
    assert np.all(got.astype(np.float32).view(np.uint32) & 0xFFFF == 0)


def test_zero_output_weights_is_jacobi(graph):
    p = mpnn.init_params(7)
    gup, gdiag, _ = mpnn.forward(mpnn.unflatten(p), graph, emu=True)
    assert np.all(gup == 0.0)
    assert np.allclose(gdiag, 1.0, rtol=0, atol=1e-15)
    loss, _, _, T = mpnn.loss_and_grad_ghat(graph, gup, gdiag)
    assert abs(loss) < 1e-12 and abs(T - graph.n) < 1e-9


def test_loss_dense(graph):
    p = _random_params()
    gup, gdiag, _ = mpnn.forward(mpnn.unflatten(p), graph, emu=False)
    loss, _, _, _ = mpnn.loss_and_grad_ghat(graph, gup, gdiag)
    G = mpnn.ghat_matrix(graph, gup, gdiag).toarray()
    assert np.allclose(G, np.triu(G))
    L = mpnn.factor_csr(graph, gup, gdiag).toarray()
    A = graph.ahat.toarray() / np.outer(graph.sd, graph.sd)
    X = L.T @ A @ L
    n = graph.n
    # log det X = log det A + 2 sum log L_ii with L_ii = sd_i Ghat_ii
    dense = (np.log(np.trace(X) / n) - np.linalg.slogdet(X)[1] / n + np.linalg.slogdet(A)[1] / n
             + 2.0 / n * np.log(graph.sd).sum())
    assert abs(dense - loss) < 1e-9 * max(1.0, abs(loss))


@pytest.mark.parametrize("emu", [False])
def test_gradient_finite_differences(graph, emu):
    p = _random_params()
    _, grad = mpnn.backward(mpnn.unflatten(p), graph, emu=emu)
    rng = np.random.default_rng(5)
    offs = 0
    worst = 0.0
    for name, shape in mpnn.param_layout():
        k = int(np.prod(shape))
        d = np.zeros_like(p)
        d[offs:offs + k] = rng.standard_normal(k)
        eps = 1e-7
        lp = mpnn.loss(mpnn.unflatten(p + eps * d), graph, emu)
        lm = mpnn.loss(mpnn.unflatten(p - eps * d), graph, emu)
        fd = (lp - lm) / (2 * eps)
        an = float(grad[name].ravel() @ d[offs:offs + k])
        err = abs(fd - an) / max(1e-8, abs(fd), abs(an))
        worst = max(worst, err)
        assert err < 5e-4, (name, fd, an)
        offs += k
    print("worst relative FD error", worst)


def _system3d():
    pts, al = clouds.sloshing_tank(7, 4, 4)
    r = ol.ashape3d(pts, al, bits=21)
    xyz = r["ordered"]
    row, col, val, b = ol.assemble_p1_tet(xyz, r["elem"], r["node_is_bnd"], xyz[:, 2] + 1.0)
    return mpnn.Graph(row, col, val, xyz, r["node_is_bnd"])


def test_gradient_finite_differences_3d():
    """3D systems: the z feature is live (non-zero gradient) and every
    gradient block matches finite differences."""
    g3 = _system3d()
    assert g3.nfeat[:, 2].max() > 0.5
    p = _random_params()
    _, grad = mpnn.backward(mpnn.unflatten(p), g3, emu=False)  # exact arithmetic: FD sees every weight
    assert np.abs(grad["enc_w"][2]).max() > 0
    rng = np.random.default_rng(8)
    offs = 0
    for name, shape in mpnn.param_layout():
        k = int(np.prod(shape))
        d = np.zeros_like(p)
        d[offs:offs + k] = rng.standard_normal(k)
        eps = 1e-7
        fd = (mpnn.loss(mpnn.unflatten(p + eps * d), g3, False) - mpnn.loss(mpnn.unflatten(p - eps * d), g3, False)) / (
            2 * eps)
        an = float(grad[name].ravel() @ d[offs:offs + k])
        assert abs(fd - an) / max(1e-8, abs(fd), abs(an)) < 5e-4, (name, fd, an)
        offs += k


def test_2d_systems_ignore_the_z_weights(graph):
    p = _random_params()
    P = mpnn.unflatten(p)
    a = mpnn.forward(P, graph, emu=True)
    P2 = {k: v.copy() for k, v in P.items()}
    P2["enc_w"][2] += 1.0
    b = mpnn.forward(P2, graph, emu=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _fsai_dense(A):
    n = A.shape[0]
    G = np.zeros((n, n))
    for i in range(n):
        P = [i] + [j for j in range(i + 1, n) if A[i, j] != 0.0]
        g = np.linalg.solve(A[np.ix_(P, P)], np.eye(len(P))[0])
        G[i, P] = g / np.sqrt(g[0])
    return G


def test_kaporin_minimiser_is_fsai():
    row, col, val, b, xy, bnd = _system(n_side=8)
    g = mpnn.Graph(row, col, val, xy, bnd)
    sel = g.mask
    idx = np.flatnonzero(sel)
    n = g.n

    def f(v):
        gup = np.zeros(len(g.ei))
        gup[idx] = v[:len(idx)]
        gd = np.where(g.bnd != 0, 1.0, v[len(idx):])
        if np.any(gd <= 0):
            return 1e30, np.zeros_like(v)
        L, dgup, dgd, _ = mpnn.loss_and_grad_ghat(g, gup, gd)
        return L, np.concatenate([dgup[idx], np.where(g.bnd != 0, 0.0, dgd)])

    v0 = np.concatenate([np.zeros(len(idx)), np.ones(n)])
    res = so.minimize(f, v0, jac=True, method="L-BFGS-B", options=dict(maxiter=2000, gtol=1e-12, ftol=1e-15))
    gup = np.zeros(len(g.ei))
    gup[idx] = res.x[:len(idx)]
    gd = np.where(g.bnd != 0, 1.0, res.x[len(idx):])
    G = mpnn.ghat_matrix(g, gup, gd).toarray()
    Ah = g.ahat.toarray()
    act = g.bnd == 0
    # rows of the minimiser all have ghat_k Ahat ghat_k = T/n; FSAI has 1
    Gs = G[np.ix_(act, act)]
    Ga = Gs / np.sqrt(np.diag(Gs @ Ah[np.ix_(act, act)] @ Gs.T))[:, None]
    F = _fsai_dense(Ah[np.ix_(act, act)])
    assert np.max(np.abs(Ga - F)) < 1e-5


def test_training_lowers_loss(graph):
    p = mpnn.init_params(7)
    opt = mpnn.Adam(len(p), lr=2e-3)
    losses = []
    for _ in range(25):
        L, grad = mpnn.backward(mpnn.unflatten(p), graph, emu=True)
        losses.append(L)
        p = opt.step(p, mpnn.flatten(grad))
    assert abs(losses[0]) < 1e-12
    assert losses[-1] < -0.02


def test_adam_hand_worked_steps():
    """The oracle's Adam against values worked by hand (Kingma & Ba,
    Algorithm 1, with bias correction).
    * Constant gradient g: m_t = (1 - b1^t) g and v_t = (1 - b2^t) g^2, so
      the corrected moments are g and g^2 at every t and each step moves w
      by lr g / (|g| + eps): after k steps w_k = w_0 - k lr g / (|g| + eps).
    * Gradients +1 then -1 (b1 = 0.9, b2 = 0.999): step 1 moves by
      -lr / (1 + eps); at t = 2, m = 0.9 * 0.1 - 0.1 = -0.01 and
      1 - b1^2 = 0.19, so m_hat = -1/19; v = 0.999 * 0.001 + 0.001 =
      0.001999 = 1 - b2^2, so v_hat = 1: step 2 moves by +lr (1/19)/(1 + eps).
      Without the bias correction step 2 would move by
      lr * 0.01 / (sqrt(0.001999) + eps) = 0.2237 lr instead of 0.0526 lr."""
    lr, eps = 1e-2, 1e-8
    g = np.array([0.5, -2.0, 1e-3, 3.0])
    opt = mpnn.Adam(4, lr=lr, eps=eps)
    w = np.zeros(4)
    for _ in range(5):
        w = opt.step(w, g)
    assert np.allclose(w, -5 * lr * g / (np.abs(g) + eps), rtol=1e-12, atol=0)

    opt = mpnn.Adam(1, lr=lr, b1=0.9, b2=0.999, eps=eps)
    w1 = opt.step(np.zeros(1), np.ones(1))
    assert np.allclose(w1, -lr / (1 + eps), rtol=1e-14, atol=0)
    w2 = opt.step(w1, -np.ones(1))
    assert np.allclose(w2 - w1, lr * (1.0 / 19.0) / (1 + eps), rtol=1e-12, atol=0)

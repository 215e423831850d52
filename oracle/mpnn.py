"""CPU oracle of the learned L L^T preconditioner network.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
bench's CPU baseline leg, never by the stream_amd package.

The reference has no learned preconditioner (its only preconditioner is
Jacobi, include/numerics/preconditioner.hpp:60-72 and
src/numerics/preconditioner_jacobi.cpp:29-60), so this module is the
specification of the network, its loss and its gradients.  It is pinned by
  * central finite differences of the loss w.r.t. every parameter tensor
    (FP64, bf16 emulation off),
  * closed forms: zero output weights give G = I (Jacobi) and loss 0; the
    loss minimiser over the pattern is the FSAI factor (Kaporin), checked
    by optimising the factor values directly with the oracle gradient,
  * an independent dense evaluation of the loss (numpy on Ahat).

Notation.  A (n x n, symmetric, CSR with structural zeros), sd_i =
1/sqrt(A_ii) (1 when A_ii <= 0), Ahat = D^-1/2 A D^-1/2.  The network emits
an upper-triangular Ghat on the pattern of A (row k: its diagonal and the
entries A_ki with i > k); the preconditioner is M^-1 = L L^T with
L = D^-1/2 Ghat^T, so L^T A L = Ghat Ahat Ghat^T.

Network (hidden 64, three message-passing layers):
  F_i   = [(x_i - xmin)/ext, (y_i - ymin)/ext, (z_i - zmin)/ext, bnd_i, A_ii / 8]
          (ext = largest box extent; z term 0 for 2D systems)
  e_ij  = [Ahat_ij, |x_i - x_j| / lbar, +1 if j > i else -1]  (lbar: mean
          length over the off-diagonal pattern)
  H0    = relu(F W_enc + b_enc)
  layer l:  [P | Q] = H_l W_pq
            z_ij = P_i + Q_j + e_ij W_e + b_m,  M_i = mean_j relu(z_ij)
            U = [H_l | M] W_u + b_u,  H_l+1 = H_l + relu(U)
  decoder:  [R | S | Dg] = H_3 W_dec
            upper edge (k, i > k): t = relu(R_k + S_i + e_ki W_de + b_de),
                                   Ghat_ki = (t . w_o) * [A_ki != 0]
            diagonal: Ghat_kk = 1 on Dirichlet rows, else
                      softplus(relu(Dg_k + b_dd) . w_d + log(e - 1))
Every product of a weight with an activation rounds both operands to bf16
(round to nearest even) and accumulates in FP32 on the GPU (FP64 here) when
`emu` is on; the backward GEMMs (dX = dY W^T, dW = X^T dY) round dY too,
the VALU parts (edge features, output dots, encoder) keep dY unrounded.

Loss (Kaporin's condition number per node; its minimiser over the pattern
is the FSAI factor):
  T = tr(Ghat Ahat Ghat^T),  loss = log(T / n) - (2/n) sum_k log Ghat_kk.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

HIDDEN = 64
LAYERS = 3
NODE_FEATS = 5
EDGE_FEATS = 3
DIAG_SHIFT = float(np.log(np.e - 1.0))  # softplus(DIAG_SHIFT) = 1


def param_layout():
    """(name, shape) of every parameter tensor in flat-array order."""
    H = HIDDEN
    out = [("enc_w", (NODE_FEATS, H)), ("enc_b", (H,))]
    for l in range(LAYERS):
        out += [(f"pq_w{l}", (H, 2 * H)), (f"e_w{l}", (EDGE_FEATS, H)), (f"m_b{l}", (H,)),
                (f"u_w{l}", (2 * H, H)), (f"u_b{l}", (H,))]
    out += [("dec_w", (H, 3 * H)), ("de_w", (EDGE_FEATS, H)), ("de_b", (H,)), ("o_w", (H,)),
            ("dd_b", (H,)), ("d_w", (H,))]
    return out


def n_params():
    return int(sum(np.prod(s) for _, s in param_layout()))


def unflatten(flat):
    out, o = {}, 0
    for name, shape in param_layout():
        k = int(np.prod(shape))
        out[name] = flat[o:o + k].reshape(shape)
        o += k
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in param_layout()])


def bf16(x):
    """Round to bf16 (nearest even) through float32; returns float64."""
    u = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """Scaled system and features of one pressure matrix."""

    def __init__(self, row, col, val, xy, bnd):
        row = np.asarray(row, np.int64)
        col = np.asarray(col, np.int64)
        val = np.asarray(val, np.float64)
        n = row.shape[0] - 1
        self.n = n
        rid = np.repeat(np.arange(n), np.diff(row))
        dmask = rid == col
        diag = np.zeros(n)
        diag[rid[dmask]] = val[dmask]
        self.diag = diag
        sd = np.where(diag > 0, 1.0 / np.sqrt(np.where(diag > 0, diag, 1.0)), 1.0)
        self.sd = sd
        off = ~dmask
        self.ei = rid[off]          # message receiver i
        self.ej = col[off]          # sender j
        self.aval = val[off]
        self.deg = np.bincount(self.ei, minlength=n).astype(np.float64)
        lo = np.minimum(self.ei, self.ej)
        hi = np.maximum(self.ei, self.ej)
        ahat = self.aval * (sd[lo] * sd[hi])
        xy = np.asarray(xy, np.float64)
        dim = xy.shape[1]
        d2 = np.zeros(lo.shape[0])
        for a in range(dim):
            d = xy[lo, a] - xy[hi, a]
            d2 = d2 + d * d
        length = np.sqrt(d2)
        lbar = length.mean() if length.size else 1.0
        sgn = np.where(self.ej > self.ei, 1.0, -1.0)
        self.efeat = np.stack([ahat.astype(np.float32), (length / lbar).astype(np.float32), sgn], 1).astype(np.float64)
        lo_xy = xy.min(0)
        ext = float((xy.max(0) - lo_xy).max()) if n else 1.0
        ext = ext if ext > 0 else 1.0
        bnd = np.asarray(bnd, np.float64)
        self.bnd = bnd
        pos = [((xy[:, a] - lo_xy[a]) / ext).astype(np.float32) if a < dim else np.zeros(n, np.float32)
               for a in range(3)]
        self.nfeat = np.stack(pos + [bnd, (diag / 8.0).astype(np.float32)], 1).astype(np.float64)
        self.up = self.ej > self.ei          # upper entries among off-diagonals
        self.mask = (self.aval != 0.0) & self.up
        # Ahat as a sparse matrix (full pattern, diagonal included)
        rows_all = rid
        self.ahat = sp.csr_matrix((val * sd[rid] * sd[col], (rows_all, col)), shape=(n, n))


def _lin(X, W, emu):
    return (bf16(X) @ bf16(W)) if emu else (X @ W)


def _mean_scatter(vals, idx, n, deg):
    out = np.zeros((n, vals.shape[1]))
    np.add.at(out, idx, vals)
    return out / np.maximum(deg, 1.0)[:, None]


def forward(p, g: Graph, emu=True):
    """Returns (Ghat_upper_values [E_off], Ghat_diag [n], cache)."""
    c = {}
    W = p
    bw = (lambda x: bf16(x)) if emu else (lambda x: x)
    z0 = bw(g.nfeat) @ bw(W["enc_w"]) + W["enc_b"]
    c["zenc"] = z0
    H = np.maximum(z0, 0.0)
    Hs = [H]
    ef = bw(g.efeat)
    for l in range(LAYERS):
        PQ = _lin(H, W[f"pq_w{l}"], emu)
        P, Q = PQ[:, :HIDDEN], PQ[:, HIDDEN:]
        z = P[g.ei] + Q[g.ej] + ef @ bw(W[f"e_w{l}"]) + W[f"m_b{l}"]
        M = _mean_scatter(np.maximum(z, 0.0), g.ei, g.n, g.deg)
        X = np.concatenate([H, M], 1)
        U = _lin(X, W[f"u_w{l}"], emu) + W[f"u_b{l}"]
        c[f"PQ{l}"], c[f"z{l}"], c[f"M{l}"], c[f"U{l}"] = PQ, z, M, U
        H = H + np.maximum(U, 0.0)
        Hs.append(H)
    c["H"] = Hs
    RSD = _lin(H, W["dec_w"], emu)
    c["RSD"] = RSD
    R, S, Dg = RSD[:, :HIDDEN], RSD[:, HIDDEN:2 * HIDDEN], RSD[:, 2 * HIDDEN:]
    ze = R[g.ei] + S[g.ej] + ef @ bw(W["de_w"]) + W["de_b"]
    te = np.maximum(ze, 0.0)
    oe = bw(te) @ bw(W["o_w"])
    gup = np.where(g.mask, oe, 0.0)
    zd = Dg + W["dd_b"]
    td = np.maximum(zd, 0.0)
    od = bw(td) @ bw(W["d_w"]) + DIAG_SHIFT
    gdiag = np.where(g.bnd != 0, 1.0, np.logaddexp(0.0, od))
    c.update(ze=ze, te=te, zd=zd, td=td, od=od)
    return gup, gdiag, c


def ghat_matrix(g: Graph, gup, gdiag):
    sel = g.up
    n = g.n
    return sp.csr_matrix((np.concatenate([gup[sel], gdiag]),
                          (np.concatenate([g.ei[sel], np.arange(n)]), np.concatenate([g.ej[sel], np.arange(n)]))),
                         shape=(n, n))


def loss_and_grad_ghat(g: Graph, gup, gdiag):
    """Kaporin loss and its gradient w.r.t. the upper values and diagonal."""
    G = ghat_matrix(g, gup, gdiag)
    Y = (G @ g.ahat).tocsr()        # y_ka = (Ghat Ahat)_ka
    T = float(G.multiply(Y).sum())
    n = g.n
    loss = np.log(T / n) - 2.0 / n * np.log(gdiag).sum()
    yup = np.asarray(Y[g.ei, g.ej]).ravel()
    yd = Y.diagonal()
    dgup = np.where(g.up, 2.0 * yup / T, 0.0)
    dgdiag = 2.0 * yd / T - 2.0 / n / gdiag
    return loss, dgup, dgdiag, T


def loss(p, g, emu=True):
    gup, gdiag, _ = forward(p, g, emu)
    return loss_and_grad_ghat(g, gup, gdiag)[0]


def backward(p, g: Graph, emu=True):
    """Loss and gradient (dict of the parameter tensors)."""
    W = p
    bw = (lambda x: bf16(x)) if emu else (lambda x: x)
    gup, gdiag, c = forward(p, g, emu)
    L, dgup, dgdiag, T = loss_and_grad_ghat(g, gup, gdiag)
    grad = {name: np.zeros(shape) for name, shape in param_layout()}
    ef = bw(g.efeat)
    # diagonal head
    sig = 1.0 / (1.0 + np.exp(-c["od"]))
    dod = np.where(g.bnd != 0, 0.0, dgdiag * sig)
    grad["d_w"] = bw(c["td"]).T @ dod
    dzd = (dod[:, None] * bw(W["d_w"])[None, :]) * (c["zd"] > 0)
    grad["dd_b"] = dzd.sum(0)
    # edge head
    doe = np.where(g.mask, dgup, 0.0)
    grad["o_w"] = bw(c["te"]).T @ doe
    dze = (doe[:, None] * bw(W["o_w"])[None, :]) * (c["ze"] > 0)
    grad["de_w"] = ef.T @ dze
    grad["de_b"] = dze.sum(0)
    dRSD = np.zeros((g.n, 3 * HIDDEN))
    np.add.at(dRSD[:, :HIDDEN], g.ei, dze)
    np.add.at(dRSD[:, HIDDEN:2 * HIDDEN], g.ej, dze)
    dRSD[:, 2 * HIDDEN:] = dzd
    H3 = c["H"][LAYERS]
    grad["dec_w"] = bw(H3).T @ bw(dRSD)
    dH = bw(dRSD) @ bw(W["dec_w"]).T
    for l in reversed(range(LAYERS)):
        H = c["H"][l]
        U = c[f"U{l}"]
        dU = dH * (U > 0)
        X = np.concatenate([H, c[f"M{l}"]], 1)
        grad[f"u_w{l}"] = bw(X).T @ bw(dU)
        grad[f"u_b{l}"] = dU.sum(0)
        dX = bw(dU) @ bw(W[f"u_w{l}"]).T
        dM = dX[:, HIDDEN:]
        dHl = dH + dX[:, :HIDDEN]
        z = c[f"z{l}"]
        dz = (dM[g.ei] / np.maximum(g.deg[g.ei], 1.0)[:, None]) * (z > 0)
        grad[f"e_w{l}"] = ef.T @ dz
        grad[f"m_b{l}"] = dz.sum(0)
        dPQ = np.zeros((g.n, 2 * HIDDEN))
        np.add.at(dPQ[:, :HIDDEN], g.ei, dz)
        np.add.at(dPQ[:, HIDDEN:], g.ej, dz)
        grad[f"pq_w{l}"] = bw(H).T @ bw(dPQ)
        dH = dHl + bw(dPQ) @ bw(W[f"pq_w{l}"]).T
    dz0 = dH * (c["zenc"] > 0)
    grad["enc_w"] = bw(g.nfeat).T @ dz0
    grad["enc_b"] = dz0.sum(0)
    return L, grad


def init_params(seed=7):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from SplitMix64; output
    weights (o_w, d_w) and biases zero, so the initial factor is Jacobi."""
    from_seed = _uniform(seed, n_params())
    flat = np.zeros(n_params())
    o = 0
    for name, shape in param_layout():
        k = int(np.prod(shape))
        if name.endswith("_w") or name[:-1].endswith("_w"):
            if name not in ("o_w", "d_w"):
                fan_in = shape[0]
                flat[o:o + k] = (2.0 * from_seed[o:o + k] - 1.0) / np.sqrt(fan_in)
        o += k
    return flat


def _uniform(seed, n):
    golden = np.uint64(0x9E3779B97F4A7C15)
    m1 = np.uint64(0xBF58476D1CE4E5B9)
    m2 = np.uint64(0x94D049BB133111EB)
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * golden
        z = (z ^ (z >> np.uint64(30))) * m1
        z = (z ^ (z >> np.uint64(27))) * m2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Adam:
    """Adam (Kingma & Ba) on the flat FP32 master parameters."""

    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, w, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return w - self.lr * mh / (np.sqrt(vh) + self.eps)


def factor_csr(g: Graph, gup, gdiag):
    """L = D^-1/2 Ghat^T as a scipy CSR matrix (lower triangular)."""
    G = ghat_matrix(g, gup, gdiag)
    return (sp.diags(g.sd) @ G.T).tocsr()

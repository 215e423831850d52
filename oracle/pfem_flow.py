"""CPU oracle of the PFEM fractional (projection) velocity step on a P1
triangle mesh with lumped mass (test infrastructure only).

No reference counterpart: the reference names the PFEM pipeline (mesh,
free-surface detection, assembly, solution: README.md:12-18) but ships no
velocity update.  The step restated here is Chorin's projection in its
textbook P1 form:

    m_i   = sum_T |T| / 3                          (lumped mass)
    (K u)_i = sum_T |T| grad phi_i . sum_k grad phi_k u_k
    u*    = u + dt (g - nu M^-1 K u)               (explicit predictor)
    D_i(u) = sum_T (|T| / 3) div u_h|T,  G_i(p) = sum_T (|T| / 3) grad p_h|T
    K p   = -(rho / dt) D(u*),  p = 0 on free-surface nodes
    u_new = u* - (dt / rho) M^-1 G(p)

Element loops in FP64 numpy (np.add.at); the GPU sums each node's own
triangles, so the two agree to rounding, not bit for bit.  Pinned in
tests/test_pfem_flow_oracle.py by closed forms: D of a linear field is
m_i div u, G of a linear pressure is m_i grad p, K of a linear field
vanishes on interior nodes (patch test), and the mass sums to the mesh
area.
"""
from __future__ import annotations

import numpy as np


def geometry(xy, elem):
    """Areas [T] and basis gradients gx, gy [T, 3] of the triangles."""
    p = xy[elem]  # [T, 3, 2]
    x, y = p[..., 0], p[..., 1]
    a2 = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (y[:, 1] - y[:, 0]) * (x[:, 2] - x[:, 0])
    gx = np.stack([y[:, 1] - y[:, 2], y[:, 2] - y[:, 0], y[:, 0] - y[:, 1]], 1) / a2[:, None]
    gy = np.stack([x[:, 2] - x[:, 1], x[:, 0] - x[:, 2], x[:, 1] - x[:, 0]], 1) / a2[:, None]
    return 0.5 * np.abs(a2), gx, gy


def lumped_mass(xy, elem):
    area, _, _ = geometry(xy, elem)
    m = np.zeros(len(xy))
    for k in range(3):
        np.add.at(m, elem[:, k], area / 3.0)
    return m


def stiffness_apply(xy, elem, u):
    """(K u) for a nodal field u [n] or [n, c]."""
    area, gx, gy = geometry(xy, elem)
    u = np.asarray(u, np.float64)
    out = np.zeros_like(u)
    ue = u[elem]  # [T, 3, ...]
    for a in range(3):
        s = 0.0
        for k in range(3):
            kak = area * (gx[:, a] * gx[:, k] + gy[:, a] * gy[:, k])
            s = s + (kak[:, None] * ue[:, k] if u.ndim == 2 else kak * ue[:, k])
        np.add.at(out, elem[:, a], s)
    return out


def divergence(xy, elem, uv):
    """D_i(u) = sum_T (|T| / 3) div u_h|T for u [n, 2]."""
    area, gx, gy = geometry(xy, elem)
    ue = uv[elem]
    div = (ue[..., 0] * gx).sum(1) + (ue[..., 1] * gy).sum(1)
    d = np.zeros(len(xy))
    for k in range(3):
        np.add.at(d, elem[:, k], area / 3.0 * div)
    return d


def gradient(xy, elem, p):
    """G_i(p) = sum_T (|T| / 3) grad p_h|T, [n, 2]."""
    area, gx, gy = geometry(xy, elem)
    pe = p[elem]
    g = np.stack([(pe * gx).sum(1), (pe * gy).sum(1)], 1)
    out = np.zeros((len(xy), 2))
    for k in range(3):
        np.add.at(out, elem[:, k], (area / 3.0)[:, None] * g)
    return out


def predictor(xy, elem, uv, dt, nu, g):
    m = lumped_mass(xy, elem)
    us = uv + dt * np.asarray(g, np.float64)[None, :]
    has = m > 0
    us[has] -= dt * nu * stiffness_apply(xy, elem, uv)[has] / m[has, None]
    return us, m


def pressure_rhs(xy, elem, bnd, us, dt, rho):
    b = -(rho / dt) * divergence(xy, elem, us)
    b[np.asarray(bnd) != 0] = 0.0
    return b


def correct(xy, elem, us, m, p, dt, rho):
    out = us.copy()
    has = m > 0
    out[has] -= (dt / rho) * gradient(xy, elem, p)[has] / m[has, None]
    return out

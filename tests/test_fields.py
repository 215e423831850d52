"""Velocity fields of the advection sequences: divergence-free and bounded."""
import numpy as np

from stream_amd import clouds


def _u3(modes, p):
    u = np.zeros_like(p)
    for kx, ky, kz, phi, bx, by, bz in modes:
        c = np.cos(kx * p[:, 0] + ky * p[:, 1] + kz * p[:, 2] + phi)
        u += c[:, None] * np.array([bx, by, bz])
    return u


def test_potential_modes_divergence_free():
    m = clouds.potential_modes()
    assert m.shape == (16, 7)
    # b = k x a is orthogonal to k: div of cos(k.x) b is -sin(k.x) k.b = 0
    assert np.max(np.abs(np.einsum("ij,ij->i", m[:, :3], m[:, 4:]))) < 1e-12
    p = np.random.default_rng(0).random((500, 3))
    h = 1e-5
    div = np.zeros(len(p))
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        div += (_u3(m, p + e)[:, a] - _u3(m, p - e)[:, a]) / (2 * h)
    assert np.max(np.abs(div)) < 1e-7
    assert np.max(np.linalg.norm(_u3(m, p), axis=1)) <= 0.25 + 1e-12


def test_advect3_matches_field():
    m = clouds.potential_modes()
    p = np.random.default_rng(1).random((100, 3))
    q = clouds.advect3(p, m, 1e-3, 1)
    assert np.allclose(q, p + 1e-3 * _u3(m, p), rtol=0, atol=1e-15)


def test_stream_modes_divergence_free():
    m = clouds.stream_modes()
    p = np.random.default_rng(2).random((300, 2))
    h = 1e-5

    def u(q):
        ux = np.zeros(len(q))
        uy = np.zeros(len(q))
        for kx, ky, phi, a in m:
            c = a * np.cos(kx * q[:, 0] + ky * q[:, 1] + phi)
            ux += ky * c
            uy -= kx * c
        return np.stack([ux, uy], 1)

    div = (u(p + [h, 0])[:, 0] - u(p - [h, 0])[:, 0] + u(p + [0, h])[:, 1] - u(p - [0, h])[:, 1]) / (2 * h)
    assert np.max(np.abs(div)) < 1e-7

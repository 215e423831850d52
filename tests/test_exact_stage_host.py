"""The device exact stage (stream_amd/csrc/predicates.hpp: monomial
expansion, error-free products, zero-eliminating growth with periodic
renormalisation) compiled for the host and checked against exact rational
arithmetic: the sign of det | x y 1 |, | x y x^2+y^2 1 |, | x y z 1 | and
| x y z x^2+y^2+z^2 1 | on random, near-degenerate (perturbed cocircular /
cospherical / collinear / coplanar) and exactly degenerate point sets, in
FP32 and FP64.  Long expansions (FP64 inputs with wide exponent ranges)
exercise the renormalisation many times over."""
from __future__ import annotations

import ctypes
from fractions import Fraction

import numpy as np
import pytest

import stream_amd

KINDS = {"orient2d": (0, 3, 2), "incircle": (1, 4, 2), "orient3d": (2, 4, 3), "insphere": (3, 5, 3)}


def _sign_gpu_code(kind, pts):
    lib = stream_amd.lib
    fn = lib.Stream_exact_sign_host
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    a = np.ascontiguousarray(pts)
    out = ctypes.c_int32(0)
    stream_amd.check(fn(KINDS[kind][0], int(a.dtype == np.float64), a.ctypes.data_as(ctypes.c_void_p),
                        ctypes.byref(out)))
    return out.value


def _det(m):
    m = [row[:] for row in m]
    n = len(m)
    det = Fraction(1)
    for c in range(n):
        p = next((r for r in range(c, n) if m[r][c] != 0), None)
        if p is None:
            return Fraction(0)
        if p != c:
            m[c], m[p] = m[p], m[c]
            det = -det
        det *= m[c][c]
        for r in range(c + 1, n):
            f = m[r][c] / m[c][c]
            for k in range(c, n):
                m[r][k] -= f * m[c][k]
    return det


def _sign_exact(kind, pts):
    P = [[Fraction(float(v)) for v in row] for row in pts]
    if kind in ("orient2d", "orient3d"):
        m = [row + [Fraction(1)] for row in P]
    else:
        m = [row + [sum(v * v for v in row), Fraction(1)] for row in P]
    d = _det(m)
    return (d > 0) - (d < 0)


def _cases(kind, dtype, rng):
    _, npts, dim = KINDS[kind]
    out = [rng.random((npts, dim)) for _ in range(30)]
    base = {2: np.array([[1, 0], [0, 1], [-1, 0], [0, -1]], np.float64),
            3: np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, 1]], np.float64)}[dim][:npts]
    for k in range(40):
        # on a circle / sphere (or a line / plane for the orientations), then
        # moved by a few ulps, sometimes not at all
        q = base.copy() * (0.37 + k) + rng.random(dim) * 100
        if kind in ("orient2d", "orient3d"):
            q = q.copy()
            q[-1] = q[0] + 0.5 * (q[1] - q[0]) if dim == 2 else q[0] + 0.3 * (q[1] - q[0]) + 0.6 * (q[2] - q[0])
        if k % 3:
            q[-1] = q[-1] * (1 + (k % 5 - 2) * np.finfo(dtype).eps)
        out.append(q)
    # lattice points (exactly degenerate in every predicate)
    g = np.array(np.meshgrid(*[np.arange(3)] * dim, indexing="ij")).reshape(dim, -1).T.astype(np.float64)
    for _ in range(20):
        out.append(g[rng.choice(len(g), npts, replace=False)] * 0.125 + 3.0)
    # wide exponent ranges: long expansions
    for _ in range(20):
        out.append(rng.random((npts, dim)) * np.exp2(rng.integers(-40, 40, (npts, dim))))
    return [np.asarray(c, dtype) for c in out]


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_exact_signs_match_rationals(kind, dtype):
    rng = np.random.default_rng(KINDS[kind][0] + (10 if dtype == np.float64 else 0))
    zeros = 0
    for pts in _cases(kind, dtype, rng):
        want = _sign_exact(kind, pts)
        assert _sign_gpu_code(kind, pts) == want, (kind, pts)
        zeros += want == 0
    assert zeros >= 3  # some degenerate sets did reach sign 0

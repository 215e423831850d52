"""Nodal-field transfer between input order and Morton order on the GPU:
permuteForward (out[i] = in[perm[i]]) and permuteBackward
(out[perm[i]] = in[i]) of include/mesh/DirectAlphaShape2D.hpp:220-242 and
the 3D counterpart, through the C-ABI, for fields of several element
widths (1 to 24 bytes), host arrays and device buffers, against numpy
indexing with the permutation the same object reports."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from stream_amd import clouds

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _fields(n, rng):
    return [
        rng.standard_normal(n).astype(np.float32),
        rng.standard_normal((n, 2)),
        rng.standard_normal((n, 3)).astype(np.float32),
        rng.integers(0, 255, n, dtype=np.uint8),
        rng.integers(0, 60000, (n, 3), dtype=np.uint16),
        rng.standard_normal((n, 3)),
    ]


def _check_host(obj, perm, fields):
    for fld in fields:
        fwd = obj.permute_forward(fld)
        assert fwd.dtype == fld.dtype and fwd.shape == fld.shape
        assert np.array_equal(fwd, fld[perm]), fld.dtype
        back = obj.permute_backward(fld)
        want = np.empty_like(fld)
        want[perm] = fld
        assert np.array_equal(back, want), fld.dtype
        assert np.array_equal(obj.permute_backward(fwd), fld)


def test_alpha_shape_2d_permute_host_and_device():
    from stream_amd.mesh import AlphaShape2D

    _say("AlphaShape2D permute")
    pts, al = clouds.jittered_square(64, seed=4)
    a = AlphaShape2D(dtype=np.float64)
    a.set_nodes(pts, al)
    a.compute()
    perm = a.get_permutation().astype(np.int64)
    n = pts.shape[0]
    assert np.array_equal(np.sort(perm), np.arange(n))
    rng = np.random.default_rng(1)
    fields = _fields(n, rng)
    _check_host(a, perm, fields)
    # the Morton-ordered nodes are the forward permutation of the input nodes
    assert np.array_equal(a.permute_forward(pts), a.get_ordered_nodes())

    # device buffers owned by the caller
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    import stream_amd

    for fld in fields[:3]:
        nb = fld.nbytes
        din, dout = ctypes.c_void_p(), ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(din), nb) == 0 and h.hipMalloc(ctypes.byref(dout), nb) == 0
        try:
            assert h.hipMemcpy(din, fld.ctypes.data, nb, 1) == 0
            eb = fld.itemsize * int(np.prod(fld.shape[1:], dtype=np.int64))
            lib = stream_amd.libmesh
            lib.AlphaShape2D_permute_forward_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                                ctypes.c_uint32]
            lib.AlphaShape2D_permute_forward_device.restype = ctypes.c_int32
            lib.AlphaShape2D_permute_backward_device.argtypes = lib.AlphaShape2D_permute_forward_device.argtypes
            lib.AlphaShape2D_permute_backward_device.restype = ctypes.c_int32
            stream_amd.check(lib.AlphaShape2D_permute_forward_device(a.mesh, din, dout, eb))
            stream_amd.synchronize()
            got = np.empty_like(fld)
            assert h.hipMemcpy(got.ctypes.data, dout, nb, 2) == 0
            assert np.array_equal(got, fld[perm])
            stream_amd.check(lib.AlphaShape2D_permute_backward_device(a.mesh, dout, din, eb))
            stream_amd.synchronize()
            assert h.hipMemcpy(got.ctypes.data, din, nb, 2) == 0
            assert np.array_equal(got, fld)
        finally:
            h.hipFree(din)
            h.hipFree(dout)


def test_alpha_shape_3d_permute():
    from stream_amd.mesh import AlphaShape3D

    _say("AlphaShape3D permute")
    rng = np.random.default_rng(2)
    pts = rng.random((3000, 3)).astype(np.float32)
    a = AlphaShape3D(dtype=np.float32)
    a.set_nodes(pts, np.full(3000, 0.06, np.float32))
    a.compute()
    perm = a.get_permutation().astype(np.int64)
    _check_host(a, perm, _fields(3000, rng))
    assert np.array_equal(a.permute_forward(pts), a.get_ordered_nodes())


def test_steps_permute_matches_their_solution_orders():
    from stream_amd.pfem import PfemStep2D, PfemStep3D

    _say("PfemStep2D / PfemStep3D permute")
    pts, al = clouds.dam_break_scaled(20_000)
    f = clouds.divergence_source(pts, clouds.fourier_modes(11))
    st = PfemStep2D(pts, al, f)
    st.run("fp64", rtol=1e-8)
    xm = st.solution(morton=True)
    assert np.array_equal(st.permute_backward(xm), st.solution())
    assert np.array_equal(st.permute_forward(st.solution()), xm)
    rng = np.random.default_rng(3)
    _, _, _, perm = st.mesh(st.run("fp64", rtol=1e-8)["n_elems"])
    _check_host(st, perm.astype(np.int64), _fields(pts.shape[0], rng)[:3])

    p3, a3 = clouds.sloshing_tank_scaled(6000)
    f3 = np.sin(3.0 * p3[:, 0]) + p3[:, 2]
    s3 = PfemStep3D(p3, a3, f3)
    s3.run("mixed", rtol=1e-8)
    _, _, _, perm3 = s3.mesh()
    _check_host(s3, perm3.astype(np.int64), _fields(p3.shape[0], rng)[:3])

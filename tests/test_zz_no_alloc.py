"""The allocation-free scan (Stream_exclusive_scan_u32_noalloc) against the
allocating entry point Stream_exclusive_scan_u32 on the same inputs: two
calls in a row, the second with a smaller input, so that what the first
call left in the shared scratch array cannot leak into the second.  The
test owns the device buffers (hipMalloc through the HIP runtime the
library already loaded)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import stream_amd

pytestmark = pytest.mark.gpu

MAX_N = 1 << 28


def _hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMalloc.restype = ctypes.c_int
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipFree.restype = ctypes.c_int
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemcpy.restype = ctypes.c_int
    return h


H2D, D2H = 1, 2


def _alloc(h, nbytes):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), max(nbytes, 4)) == 0
    return p


def _scan_noalloc(h, x):
    lib = stream_amd.lib
    n = x.shape[0]
    din = _alloc(h, 4 * n)
    dout = _alloc(h, 4 * (n + 1))
    try:
        if n:
            assert h.hipMemcpy(din, x.ctypes.data, 4 * n, H2D) == 0
        print(f"[no_alloc] scan n={n}", flush=True)
        stream_amd.check(lib.Stream_exclusive_scan_u32_noalloc(ctypes.c_uint32(n), din, dout, None))
        out = np.empty(n + 1, np.uint32)
        assert h.hipMemcpy(out.ctypes.data, dout, 4 * (n + 1), D2H) == 0
        return out
    finally:
        h.hipFree(din)
        h.hipFree(dout)


def _scan_alloc(x):
    lib = stream_amd.lib
    n = x.shape[0]
    out = np.empty(n + 1, np.uint32)
    xin = np.ascontiguousarray(x, np.uint32)
    stream_amd.check(lib.Stream_exclusive_scan_u32(ctypes.c_uint32(n), xin.ctypes.data_as(ctypes.c_void_p),
                                                   out.ctypes.data_as(ctypes.c_void_p)))
    return out


@pytest.fixture(scope="module", autouse=True)
def _signatures():
    lib = stream_amd.lib
    lib.Stream_exclusive_scan_u32_noalloc.restype = ctypes.c_int32
    lib.Stream_exclusive_scan_u32_noalloc.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_void_p]
    lib.Stream_exclusive_scan_u32.restype = ctypes.c_int32
    lib.Stream_exclusive_scan_u32.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]


def test_noalloc_scan_matches_allocating_scan_twice():
    h = _hip()
    rng = np.random.default_rng(5)
    # first call: 3M items spanning 1465 look-back tiles; second call
    # smaller (a ragged last tile), different data
    for n in (3_000_000, 100_003):
        x = rng.integers(0, 16, n, dtype=np.uint32)
        got = _scan_noalloc(h, x)
        want = _scan_alloc(x)
        assert np.array_equal(got, want), n
        ref = np.zeros(n + 1, np.uint64)
        np.cumsum(x, dtype=np.uint64, out=ref[1:])
        assert np.array_equal(got.astype(np.uint64), ref), n


def test_noalloc_scan_edge_sizes():
    h = _hip()
    for n in (0, 1, 2047, 2048, 2049):
        x = np.arange(1, n + 1, dtype=np.uint32)
        assert np.array_equal(_scan_noalloc(h, x), _scan_alloc(x)), n


def test_noalloc_scan_refuses_sizes_beyond_the_scratch():
    # refused on the host before any launch or pointer use
    st = stream_amd.lib.Stream_exclusive_scan_u32_noalloc(ctypes.c_uint32(MAX_N + 1), None, None, None)
    assert st == 2
    assert b"2^28" in stream_amd.lib.Stream_last_error()

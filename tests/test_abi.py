"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares; the headers compile as C."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "stream_amd", "lib")


def _declared():
    names = []
    for h in sorted(os.listdir(INC)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.append(name)
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = _declared()
    assert "AlphaShape2D_create" in names and "SolverCG_jacobi_solve" in names and "LBVH2D_build_f64" in names
    assert len(names) > 50


@pytest.mark.parametrize("libname", ["libstream_amd.so", "libgeometry.so", "libmesh.so", "libnumerics.so"])
def test_library_exports_every_declared_symbol(libname):
    lib = ctypes.CDLL(os.path.join(LIBDIR, libname))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_compile_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "stream_geometry.h"\n#include "stream_mesh.h"\n#include "stream_numerics.h"\n#include "stream_learned.h"\n'
                   "int main(void){return 0;}\n")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INC, str(src), "-o", str(tmp_path / "t")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_python_package_imports_and_reports_no_device_here():
    import stream_amd

    assert stream_amd.device_count() >= 0
    from stream_amd.mesh import AlphaShape2D  # noqa: F401
    from stream_amd.numerics import DeviceCSR, HostCSR, LinSys, PrecJacobi, SolverCG  # noqa: F401


def test_wrappers_refuse_mismatched_node_arrays():
    """Node arrays of the wrong width and per-node arrays of the wrong length
    are refused in Python, before the native library reads n rows of them."""
    import numpy as np

    from stream_amd._ct import node_arrays

    xy = np.zeros((10, 2))
    a, b = node_arrays(xy, 2, np.float64, np.ones(10))
    assert a.shape == (10, 2) and b.shape == (10,)
    for bad in (np.zeros((10, 3)), np.zeros(20)):
        with pytest.raises(ValueError):
            node_arrays(bad, 2, np.float64)
    with pytest.raises(ValueError):
        node_arrays(xy, 2, np.float64, np.ones(9))
    with pytest.raises(ValueError):
        node_arrays(np.zeros((10, 2)), 3, np.float64)


def test_noalloc_scan_refuses_oversize_without_a_device():
    """The allocation-free scan checks its size limit on the host before any
    HIP call, so the refusal is observable without a GPU."""
    import stream_amd

    lib = stream_amd.lib
    lib.Stream_exclusive_scan_u32_noalloc.restype = ctypes.c_int32
    lib.Stream_exclusive_scan_u32_noalloc.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_void_p]
    assert lib.Stream_exclusive_scan_u32_noalloc((1 << 28) + 1, None, None, None) == 2

"""Build the HIP library for gfx950.

Compiles every ``csrc/*.hip`` with hipcc into objects and links
``lib/libstream_amd.so``.  The reference's three library names
(``libgeometry.so``, ``libmesh.so``, ``libnumerics.so``; reference
CMakeLists.txt:93-95, python/__init__.py:40-42) are created as symlinks to
it, so a loader written for the reference finds every symbol under the name
it expects.

Usage: ``python stream_amd/build.py`` (or ``build()`` from Python; the
package itself needs the built library, so load this file by path).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIBDIR = HERE / "lib"
OBJDIR = HERE / "lib" / "obj"
LIBNAME = "libstream_amd.so"
ALIASES = ("libgeometry.so", "libmesh.so", "libnumerics.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-exact floating point: no implicit a*b+c contraction
    "-ffp-contract=off",
    # packed FP32 instructions stay off (explicit fma only where written)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: Path) -> Path:
    obj = OBJDIR / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile and link; returns the library path.  Always rebuilds."""
    LIBDIR.mkdir(parents=True, exist_ok=True)
    OBJDIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    for alias in ALIASES:
        p = LIBDIR / alias
        if p.is_symlink() or p.exists():
            p.unlink()
        p.symlink_to(LIBNAME)
    shutil.rmtree(OBJDIR, ignore_errors=True)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)

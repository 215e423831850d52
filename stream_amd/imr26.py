"""The IMR26 benchmark workloads and point-cloud files on the MI355X build.

Counterparts of the reference's testcases/IMR26 programs:

* ``uniform2d`` / ``uniform3d``: ``uniform2D.cpp`` / ``uniform3D.cpp``
  (points drawn with srand(seed), (float)rand()/RAND_MAX per coordinate,
  one alpha for all points, FP32; the alpha-shape object is created once
  and init + compute are timed per run).  Columns as the reference prints
  them (``uniform2D.cpp:186-191``): run, alpha, npoints, ntri, tlbvh,
  tashape, tcompress, ttot = tlbvh + tashape, all in ms; tlbvh, tashape and
  tcompress are GPU times (HIP events between the launches), ``wall`` is
  the host time of the two calls.
* ``files``: ``run_on_simulation_meshes.py:28-151`` on point-cloud files in
  the dataset's format (``README.md:138-145``): a flat float32 array of
  records (x, y, alpha) in 2D or (x, y, z, alpha) in 3D.
* ``all``: ``run_all.py:24-69``, both parameter tables of the README
  (``README.md:34-41,80-86``) and, when a directory of files is given, the
  file runner.

    python -m stream_amd.imr26 uniform2d -n 1000000 -a 0.0012 -r 5
    python -m stream_amd.imr26 uniform3d -n 100000 -a 0.022 -r 5
    python -m stream_amd.imr26 files --dim 2 input_meshes/mesh_dambreak/*.bin
"""
from __future__ import annotations

import argparse
import ctypes
import pathlib
import sys
import time

import numpy as np

from . import check, lib
from ._ct import bind, u32

# parameter tables of testcases/IMR26/README.md:34-41 and :80-86
PAPER_2D = ((10_000, 0.024), (50_000, 0.0084), (100_000, 0.006), (500_000, 0.0018), (1_000_000, 0.0012),
            (10_000_000, 0.00042))
PAPER_3D = ((10_000, 0.06), (50_000, 0.036), (100_000, 0.024), (500_000, 0.0144), (1_000_000, 0.0108))

_uniform = bind(lib, "Stream_uniform_cloud_f32", ctypes.c_int32, [u32, u32, u32, ctypes.c_void_p])


def uniform(n: int, dim: int, seed: int = 42) -> np.ndarray:
    """float32 [n, dim] points as uniform2D.cpp:130-141 draws them."""
    out = np.empty((n, dim), np.float32)
    check(_uniform(n, dim, seed, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def write_cloud(path, coords, alpha) -> None:
    """Write a point cloud in the dataset's format: float32 records
    (x, y, [z,] alpha)."""
    c = np.asarray(coords, np.float32)
    if c.ndim != 2 or c.shape[1] not in (2, 3):
        raise ValueError(f"coords must have shape (n, 2) or (n, 3), got {c.shape}")
    a = np.broadcast_to(np.asarray(alpha, np.float32), (c.shape[0],))
    rec = np.empty((c.shape[0], c.shape[1] + 1), np.float32)
    rec[:, :-1] = c
    rec[:, -1] = a
    rec.tofile(str(path))


def read_cloud(path, dim: int):
    """Read a point cloud in the dataset's format (README.md:138-145:
    numpy.fromfile(float32).reshape(-1, dim + 1)); returns (coords [n, dim],
    alpha [n]) as float32."""
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    raw = np.fromfile(str(path), dtype=np.float32)
    if raw.size % (dim + 1):
        raise ValueError(f"{path}: {raw.size} float32 values is not a whole number of {dim + 1}-value records")
    rec = raw.reshape(-1, dim + 1)
    return np.ascontiguousarray(rec[:, :dim]), np.ascontiguousarray(rec[:, dim])


def _shape(dim, dtype=np.float32):
    from .mesh import AlphaShape2D, AlphaShape3D

    return AlphaShape2D(dtype=dtype) if dim == 2 else AlphaShape3D(dtype=dtype)


def run_uniform(dim: int, n: int, alpha: float, runs: int = 5, seed: int = 42):
    """uniform2D.cpp / uniform3D.cpp: one object, set_nodes once, then per
    run init + compute; returns one dict per run."""
    pts = uniform(n, dim, seed)
    a = _shape(dim)
    a.set_nodes(pts, np.full(n, alpha, np.float32))
    rows = []
    for k in range(runs):
        t = a.timed_run()
        ntri = len(a.get_elem())
        rows.append(dict(run=k, alpha=float(np.float32(alpha)), npoints=n, ntri=ntri, tlbvh=t["tlbvh"],
                         tashape=t["tashape"], tcompress=t["tcompress"], ttot=t["tlbvh"] + t["tashape"],
                         wall=t["wall"]))
    return rows


def print_rows(rows, out=sys.stdout):
    """The reference's table (uniform2D.cpp:186-191) plus the host wall time."""
    print(f"{'run':<5} {'alpha':<10} {'npoints':<10} {'ntri':<10} {'tlbvh':<10} {'tashape':<10} {'tcompress':<10} "
          f"{'ttot':<10} {'wall':<10}", file=out)
    for r in rows:
        print(f"{r['run']:<5} {r['alpha']:<10.3e} {r['npoints']:<10} {r['ntri']:<10} {r['tlbvh']:<10.3e} "
              f"{r['tashape']:<10.3e} {r['tcompress']:<10.3e} {r['ttot']:<10.3e} {r['wall']:<10.3e}", file=out)


def run_files(files, dim: int, nruns: int = 30, out=sys.stdout):
    """run_on_simulation_meshes.py: for each file, nruns of set_nodes +
    compute timed on the host (as the reference times them), then the
    element count; missing files are reported and skipped.  Returns one
    dict per file read."""
    a = None  # one object for all files, created at the first file read
    res = []
    for f in files:
        if not pathlib.Path(f).exists():
            print(f"Could not find input file '{f}'", file=out)
            continue
        coords, alpha = read_cloud(f, dim)
        if a is None:
            a = _shape(dim)
        t0 = time.perf_counter()
        for _ in range(nruns):
            a.set_nodes(coords, alpha)
            a.compute()
        per_run = (time.perf_counter() - t0) / nruns
        elem = a.get_elem()
        kind = "triangles" if dim == 2 else "tetrahedrons"
        print(f"[{dim}D] Time to mesh '{f}' ({len(coords)} nodes, {len(elem)} {kind}) : {per_run:.4f}s", file=out)
        res.append(dict(file=str(f), nodes=len(coords), elems=len(elem), seconds_per_run=per_run))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m stream_amd.imr26", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("uniform2d", "uniform3d"):
        p = sub.add_parser(name)
        p.add_argument("-n", "--npoints", type=int, default=100)
        p.add_argument("-a", "--alpha", type=float, default=1.0)
        p.add_argument("-r", "--nruns", type=int, default=30)
        p.add_argument("--seed", type=int, default=42)
    p = sub.add_parser("files")
    p.add_argument("--dim", type=int, choices=(2, 3), required=True)
    p.add_argument("--nruns", type=int, default=30)
    p.add_argument("paths", nargs="+")
    p = sub.add_parser("all")
    p.add_argument("--nruns", type=int, default=30)
    p.add_argument("--dataset", help="directory holding input_meshes/ (the unzipped dataset), optional")
    args = ap.parse_args(argv)
    if args.cmd in ("uniform2d", "uniform3d"):
        print_rows(run_uniform(2 if args.cmd == "uniform2d" else 3, args.npoints, args.alpha, args.nruns, args.seed))
    elif args.cmd == "files":
        run_files(args.paths, args.dim, args.nruns)
    else:
        for dim, table in ((2, PAPER_2D), (3, PAPER_3D)):
            for n, alpha in table:
                print("=" * 70)
                print(f"uniform{dim}D -n {n} -a {alpha}")
                print_rows(run_uniform(dim, n, alpha, args.nruns))
        if args.dataset:
            root = pathlib.Path(args.dataset) / "input_meshes"
            run_files(sorted(root.glob("mesh_waterfall/*.bin")) + sorted(root.glob("mesh_dambreak/*.bin")), 2,
                      args.nruns)
            run_files(sorted(root.glob("*.bin")), 3, args.nruns)


if __name__ == "__main__":
    main()

"""Benchmark of the PFEM mesh-and-solve step on one MI355X.

Workload: the 2D dam-break cloud, 1,000,000 FP64 points (a 704 x 1407
jittered fluid column plus 9,472 scattered fragments, alpha = 1.2h), with
the nodal source f = div u of 8 seeded Fourier modes.  Points and source
are resident in HBM before timing starts.

One step = LBVH build + 2D alpha-shape + P1 pressure assembly + CG to a
true FP64 relative residual ||b - A x|| / ||b|| <= 1e-8 (verified after
every solve).  The headline solver is the mixed-precision PCG (FP32 SpMV
and preconditioner, FP64 dots and residual replacement); the FP64
Jacobi-PCG step rate is reported beside it.

Prints one JSON line (see the keys below).  ``--gpus N`` with N > 1 exits:
this build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mesh+solve steps/s @1M 2D pts, rtol 1e-8: learned mixed PCG vs FP64 Jacobi-CG"
HBM_PEAK_GBS = 8000.0


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline(pts, al, f, fp64_iters):
    """Oracle step on the host: one mesh + assembly at full size and a bounded
    number of FP64 Jacobi-PCG iterations, the CG time scaled to the
    iteration count the GPU FP64 solve needed."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    L = O.lib()
    n = pts.shape[0]
    xyr = O.spheres(pts, al, np.float64)
    fa = np.ascontiguousarray(f, np.float64)
    times = np.zeros(3)
    sizes = np.zeros(2, np.uint64)
    sample_iters = 100
    it = L.orc_step_f64(n, xyr.ctypes.data_as(ctypes.c_void_p), fa.ctypes.data_as(ctypes.c_void_p), sample_iters,
                        times.ctypes.data_as(ctypes.c_void_p), sizes.ctypes.data_as(ctypes.c_void_p))
    per_iter = times[2] / max(1, it)
    step_ms = times[0] + times[1] + per_iter * fp64_iters
    return {
        "value": 1000.0 / step_ms,
        "unit": "steps/s",
        "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
        "kind": "port",
        "sample": (f"one full-size mesh ({times[0]:.0f} ms) + P1 assembly ({times[1]:.0f} ms) + {it} FP64 "
                   f"Jacobi-PCG iterations ({per_iter:.2f} ms/iter) on the C++ oracle (OpenMP), CG extrapolated "
                   f"to the {fp64_iters} iterations of the FP64 solve; host CPU: {_cpu_model()}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--solver", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 companion measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU decomposition is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import stream_amd
    from stream_amd import clouds
    from stream_amd.pfem import PfemStep2D

    if stream_amd.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    pts, al = clouds.dam_break()
    f = clouds.divergence_source(pts, clouds.fourier_modes(11))
    step = PfemStep2D(pts, al, f)

    def timed(solver, k, w):
        infos = []
        for _ in range(w):
            step.run(solver, rtol=1e-8)
        stream_amd.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            infos.append(step.run(solver, rtol=1e-8))
        stream_amd.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / k, infos

    ms, infos = timed(args.solver, args.steps, args.warmup)
    last = infos[-1]
    print(f"[bench] {args.solver}: {ms:.2f} ms/step  mesh {last['mesh_ms']:.2f}  asm {last['assemble_ms']:.2f}  "
          f"solve {last['solve_ms']:.2f} ms  iters {last['iters']}  rel {last['true_rel']:.2e}", file=sys.stderr,
          flush=True)
    fp64 = None
    if not args.no_fp64 and args.solver != "fp64":
        ms64, inf64 = timed("fp64", max(1, args.steps // 2), 1)
        fp64 = dict(ms_per_step=ms64, iters=inf64[-1]["iters"], true_rel=inf64[-1]["true_rel"],
                    solve_ms=inf64[-1]["solve_ms"])
        print(f"[bench] fp64: {ms64:.2f} ms/step iters {fp64['iters']}", file=sys.stderr, flush=True)

    # dominant kernel: the CG direction kernel (d update + SpMV + d.q)
    fp32 = args.solver != "fp64"
    kms = step.time_dir_kernel(fp32=fp32, reps=200)
    n = pts.shape[0]
    nnz = last["nnz"]
    vb = 4 if fp32 else 8
    # algorithmic bytes per launch: row pointers, (col, val) per entry,
    # read z and d_old, write d_new and q
    alg_bytes = 4 * (n + 1) + nnz * (4 + vb) + 4 * vb * n
    achieved = alg_bytes / (kms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"cg_dir_kernel<{'float' if fp32 else 'double'}>", "kernel_ms": round(kms, 5),
                "alg_bytes_per_launch": int(alg_bytes)}

    cpu = None
    if not args.no_cpu:
        iters64 = fp64["iters"] if fp64 else last["iters"]
        cpu = cpu_baseline(pts, al, f, iters64)

    out = {
        "metric": METRIC,
        "value": round(1000.0 / ms, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-residual" if fp32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": "2D dam-break cloud, 1,000,000 FP64 points, alpha=1.2h, LBVH + alpha-shape + P1 pressure "
                        "assembly + CG to true FP64 rtol 1e-8",
            "points": int(n),
            "solver": "mixed FP32/FP64 Jacobi-PCG" if fp32 else "FP64 Jacobi-PCG",
            "nnz": int(nnz),
            "triangles": int(last["n_elems"]),
        },
        "stages_ms": {"mesh": round(last["mesh_ms"], 3), "assembly": round(last["assemble_ms"], 3),
                      "solve": round(last["solve_ms"], 3)},
        "iters": last["iters"],
        "true_rel": last["true_rel"],
        "fp64_jacobi": None if fp64 is None else {
            "steps_per_s": round(1000.0 / fp64["ms_per_step"], 4), **fp64},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

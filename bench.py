"""Benchmark of the PFEM mesh-and-solve step on one MI355X.

Workload: the 2D dam-break cloud, 1,000,000 FP64 points (a 704 x 1407
jittered fluid column plus 9,472 scattered fragments, alpha = 1.2h), with
the nodal source f = div u of 8 seeded Fourier modes.  Points and source
are resident in HBM before timing starts.

One step = LBVH build + 2D alpha-shape + P1 pressure assembly +
preconditioner set-up + CG to a true FP64 relative residual
||b - A x|| / ||b|| <= 1e-8 (verified after every solve).  The headline
solver is the learned mixed PCG: M^-1 = L L^T from the trained
message-passing network (weights in stream_amd/weights), FP32 SpMV and
preconditioner, FP64 dots and residual replacement.  The FP64 Jacobi-PCG
step rate is measured beside it in the same run.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the step runs on one
GPU.
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
BF16_PEAK_TFLOPS = 2500.0
WEIGHTS = os.path.join(ROOT, "stream_amd", "weights", "mpnn_dam2d")
WEIGHTS3D = os.path.join(ROOT, "stream_amd", "weights", "mpnn_tank3d")
SOLVER_LABEL = {
    "learned": "learned mixed PCG (MPNN L L^T, FP32 SpMV/apply, FP64 dots + residual replacement)",
    "mixed": "mixed FP32/FP64 Jacobi-PCG",
    "fsai": "mixed PCG with the classical FSAI L L^T",
    "fp64": "FP64 Jacobi-PCG",
}


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline(pts, al, f):
    """Oracle step on the host cores (the C++ restatement with OpenMP, one
    parallel loop per device launch of the reference): one full-size mesh,
    P1 assembly and FP64 Jacobi-PCG from x = 0 to a true relative residual
    of 1e-8, all timed; then, untimed, the same PCG continued to 1e-11,
    whose solution is the FP64 reference the GPU solutions are compared
    with.  Threads: OMP_NUM_THREADS (the GPU box allots 16 host cores per
    GPU and sets it to 16), bound close."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle_lib as O

    r = O.step_solve(pts, al, f, rtol=1e-8, ref_rtol=1e-11)
    mesh_ms, asm_ms, cg_ms = r["times_ms"]
    step_ms = mesh_ms + asm_ms + cg_ms
    base = {
        "value": round(1000.0 / step_ms, 4),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"one full step on the C++ oracle (OpenMP, {threads} threads, OMP_PROC_BIND=close): mesh "
                   f"{mesh_ms:.0f} ms + P1 assembly {asm_ms:.0f} ms + FP64 Jacobi-PCG to a true rtol of 1e-8 "
                   f"({r['iters']} iterations, {cg_ms:.0f} ms, rel {r['rel']:.2e}); host CPU: {_cpu_model()}, "
                   f"{os.cpu_count()} logical CPUs visible"),
        "ms_per_step": round(step_ms, 1),
        "cpu_model": _cpu_model(),
    }
    ref = {"x": r["x_ref"], "rel": r["ref_rel"], "iters": r["ref_iters"]}
    return base, ref


def sell_dir_bytes(n, nnz, value_bytes, llt=False):
    """Algorithmic bytes of one CG direction launch on the SELL-64 matrix.
    Jacobi path (d = z + beta d_old, q = A d, d.q): per row the length (2 B),
    the diagonal, the own {z, d_old} pair, d and q written.  L L^T path
    (q = A d on a ready-made d, d.q): per row the length, the diagonal, d
    read and q written.  Per off-diagonal entry its 16-bit column offset
    and value (the 4-byte columns of the entries whose offset does not fit,
    one in a few hundred, are not counted); per 64-row slice its offset and
    width (6 B).  Each vector entry is counted once (neighbour re-reads are
    cache traffic)."""
    vb = value_bytes
    nslices = (n + 63) // 64
    row = 2 + vb + vb + vb if llt else 2 + vb + 2 * vb + vb + vb
    return n * row + (nnz - n) * (2 + vb) + nslices * 6


def sell_update_bytes(n, nnz):
    """Algorithmic bytes of one L L^T update launch (x += alpha d, r' = r -
    alpha q, w = L^T r', partials of w.w and r'.r') on the upper SELL of the
    factor: per row the length (2 B), L_ii (4 B), the own {r, q} pair (8 B),
    d (4 B), the FP64 x read and written (16 B), r' and w written (8 B); per
    strict-upper entry (nnz - n) / 2 its 16-bit column offset and FP32 value
    (6 B); per 64-row slice its offset and width (6 B).  Each vector entry
    is counted once (neighbour re-reads are cache traffic)."""
    nslices = (n + 63) // 64
    return n * 42 + (nnz - n) // 2 * 6 + nslices * 6


def sell_lo_bytes(n, nnz):
    """Algorithmic bytes of one L L^T lower-apply launch (z = L w, d_new =
    z + beta d_old) on the lower SELL of the factor: per row the length
    (2 B), L_ii (4 B), the own w (4 B), d_old (4 B), d_new written (4 B);
    per strict-lower entry (nnz - n) / 2 its 16-bit column offset and FP32
    value (6 B); per 64-row slice its offset and width (6 B)."""
    nslices = (n + 63) // 64
    return n * 18 + (nnz - n) // 2 * 6 + nslices * 6


def section_bytes(n, nnz, tris, edges):
    """Algorithmic HBM bytes of the mesh-and-assembly stages on n FP64
    points (32-byte padded records), per stage, each array counted once:
      morton: box partials read the records (32 n), the Morton pass reads
              them again and writes codes + indices (32 n + 12 n);
      sort: 8 digit passes, each reading and writing (u64 key, u32 index)
            pairs (8 x 24 n);
      gather: index 4 n, record read 32 n and written 32 n;
      karras: codes 8 n, node words and parent links 17 n;
      refit: record 32 n, parent links and flags 12 n, child boxes read
             and box written 2 x 32 n;
      fill: record 32 n, candidate ids written 4 x 17 n (K = 17 in the
            2 delta alpha disc at alpha = 1.2 h);
      delaunay: candidate ids 4 K n + record 32 n + local triangles
                (2 B) and counts written, ~128 n in all;
      scans: u8 and u32 counts read, two offset arrays written (13 n);
      compress: per-candidate flags and ids read (5 K n), edges written
                (5 B each), triangles written (12 B each);
      p1: per node its local triangles (2 B each, ~6), the candidate ids
          they name (4 B, ~6), counts and flags (2 B), edge offsets (4 B),
          its record (32 B) and source (8 B); per edge its id (4 B); per
          non-zero column and value written (12 B); per row the row offset,
          FP64 inverse diagonal, FP32 inverse diagonal and load written
          (24 B)."""
    K = 17
    return {
        "morton": 76 * n,
        "sort": 192 * n,
        "gather": 68 * n,
        "karras": 25 * n,
        "refit": 108 * n,
        "fill": (32 + 4 * K) * n,
        "delaunay": 128 * n,
        "scans": 13 * n,
        "compress": 5 * K * n + 5 * edges + 12 * tris,
        "p1": (12 + 24 + 2 + 4 + 32 + 8 + 24) * n + 4 * edges + 12 * nnz,
    }


def imr26_leg(runs=5):
    """The reference's published alpha-shape timings (testcases/IMR26/
    README.md:49-54, 94-99: 2D 1M uniform points at alpha 0.0012 and 3D
    100k at alpha 0.022, FP32, seed 42, ttot = tlbvh + tashape) measured the
    same way here (stream_amd.imr26): the median over runs after the first,
    GPU ms, with the host wall time of init + compute beside it."""
    import numpy as np

    from stream_amd import imr26

    out = {}
    for dim, n, alpha, ref_ms, ref_count in ((2, 1_000_000, 0.0012, 22.35, 1_877_465),
                                              (3, 100_000, 0.022, 28.0, 526_040)):
        rows = imr26.run_uniform(dim, n, alpha, runs)[1:]
        med = {k: float(np.median([r[k] for r in rows])) for k in ("tlbvh", "tashape", "tcompress", "ttot", "wall")}
        out[f"uniform{dim}D"] = {"npoints": n, "alpha": alpha, "elements": rows[-1]["ntri"],
                                 "reference_elements": ref_count,
                                 **{k: round(v, 4) for k, v in med.items()},
                                 "reference_ttot_ms": ref_ms,
                                 "reference_hardware": "RTX 4060 (testcases/IMR26/README.md:8)",
                                 "speedup_ttot": round(ref_ms / med["ttot"], 2)}
        print(f"[bench] imr26 uniform{dim}D n={n}: ttot {med['ttot']:.3f} ms (reference {ref_ms} ms), "
              f"wall {med['wall']:.3f} ms, elements {rows[-1]['ntri']}", file=sys.stderr, flush=True)
    return out


def tank3d_leg(steps=2, weights3d=None):
    """The 3D step on the sloshing-tank cloud (1,996,920 FP64 points,
    alpha = 1.3h): LBVH + alpha-shape tetrahedra + P1 + CG to rtol 1e-8,
    mixed and FP64 Jacobi-PCG, the mixed PCG with the classical FSAI factor
    (and the learned mixed PCG when 3D weights are given); steps/s and stage
    times of the last step."""
    import numpy as np

    import stream_amd
    from stream_amd import clouds
    from stream_amd.pfem import PfemStep3D

    pts, al = clouds.sloshing_tank()
    f = np.sin(3.0 * pts[:, 0]) * np.cos(2.0 * pts[:, 1]) + pts[:, 2]
    st = PfemStep3D(pts, al, f)
    out = {"points": int(len(pts))}
    solvers = ("mixed", "fp64", "fsai")
    if weights3d:
        from stream_amd.learned import PrecLearned

        st.set_learned(PrecLearned(PrecLearned.load_params(weights3d)))
        solvers += ("learned",)
    for solver in solvers:
        st.run(solver, rtol=1e-8)
        stream_amd.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = st.run(solver, rtol=1e-8)
        stream_amd.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        out[solver] = {"steps_per_s": round(1000.0 / ms, 4), "ms_per_step": round(ms, 2),
                       "mesh_ms": round(r["mesh_ms"], 2), "assembly_ms": round(r["assemble_ms"], 2),
                       "prep_ms": round(r["prep_ms"], 2), "solve_ms": round(r["solve_ms"], 2),
                       "iters": r["iters"], "true_rel": r["true_rel"], "tets": r["n_elems"], "nnz": r["nnz"]}
        print(f"[bench] tank3d {solver}: {ms:.1f} ms/step {out[solver]}", file=sys.stderr, flush=True)
    del st
    return out


def components(step, pts, al, net):
    """GPU times of the path's components on the 1M dam-break cloud (the
    step's own stage timers, the TRBVH and Mesh2D phase timers, and the CG
    kernels timed alone), for the per-component record."""
    import numpy as np

    import stream_amd
    from stream_amd import geometry
    from stream_amd.mesh import Mesh2D

    out = {"workload": "2D dam-break cloud, 1,000,000 FP64 points (FP32 copies for TRBVH and Mesh2D)"}
    for solver in ("fp64", "mixed", "fsai") + (("learned",) if net is not None else ()):
        step.run(solver, rtol=1e-8)
        r = step.run(solver, rtol=1e-8)
        out[f"step_{solver}"] = {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in r.items()}
    out["dir_kernel_fp32_ms"] = round(step.time_dir_kernel(fp32=True, reps=200), 5)
    out["dir_kernel_llt_ms"] = round(step.time_dir_kernel(fp32=True, reps=200, llt=True), 5)
    out["update_kernel_llt_ms"] = round(step.time_update_kernel(reps=200), 5)
    out["dir_kernel_fp64_ms"] = round(step.time_dir_kernel(fp32=False, reps=200), 5)
    xy = np.ascontiguousarray(pts, np.float32)
    rad = np.full(len(xy), np.float32(0.5) * np.float32(al[0]), np.float32)
    geometry.build_trbvh(xy, rad, 3)
    t = geometry.build_trbvh(xy, rad, 3)
    out["trbvh_ms"] = {"build": round(float(t["ms"][0]), 4), "3_rounds": round(float(t["ms"][1]), 4)}
    m = Mesh2D()
    m.set_nodes(xy)
    m.insert_BVH_neighbors()
    m.insert_iterative()
    m.remove_super_nodes()
    ne = len(m.get_elem())
    st = m.get_stats()
    out["mesh2d"] = {"triangles": ne, **st}
    stream_amd.synchronize()
    return out


def sequence3d(steps, weights):
    """The 3D advect-remesh-solve sequence on the 8,012,034-point tank
    (342 x 171 x 137 jittered grid, alpha = 1.3h): every step meshes,
    assembles and solves with the learned mixed PCG to a verified FP64
    rtol of 1e-8, then advects the points (dt = 1e-3, seed-44 field).
    Reports steps/s over the timed steps, per-stage means and the peak
    device memory (the library's allocator high-water mark and the
    device's used memory)."""
    import numpy as np

    import stream_amd
    from stream_amd import clouds
    from stream_amd.learned import PrecLearned
    from stream_amd.pfem import PfemStep3D

    if not os.path.exists(weights + ".bin"):
        print(f"3D learned solver: weight file {weights}.bin not found "
              "(python -m stream_amd.train_learned --dim 3 --out <prefix>)", file=sys.stderr)
        sys.exit(1)
    pts, al = clouds.sloshing_tank(*clouds.TANK8M)
    f = np.sin(3.0 * pts[:, 0]) * np.cos(2.0 * pts[:, 1]) + pts[:, 2]
    st = PfemStep3D(pts, al, f)
    net = PrecLearned(PrecLearned.load_params(weights))
    st.set_learned(net)
    modes = clouds.potential_modes()
    stream_amd.device_bytes(reset_peak=True)
    used_peak = 0
    rows = []
    stream_amd.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        r = st.run("learned", rtol=1e-8)
        mem = stream_amd.device_bytes()
        used_peak = max(used_peak, mem["total"] - mem["free"])
        rows.append(r)
        print(f"[seq3d] step {k}: mesh {r['mesh_ms']:.1f} asm {r['assemble_ms']:.1f} prep {r['prep_ms']:.1f} "
              f"solve {r['solve_ms']:.1f} ms, {r['iters']} it, rel {r['true_rel']:.2e}, tets {r['n_elems']}",
              file=sys.stderr, flush=True)
        st.advect(modes, 1e-3, 1)
    stream_amd.synchronize()
    wall = time.perf_counter() - t0
    mem = stream_amd.device_bytes()
    mean = lambda key: round(float(np.mean([r[key] for r in rows])), 3)
    return {
        "metric": "3D advect-remesh-solve steps/s, 8,012,034-point tank, learned mixed PCG rtol 1e-8",
        "value": round(steps / wall, 4), "unit": "steps/s", "steps": steps, "ms_per_step": round(1000 * wall / steps, 2),
        "points": int(len(pts)),
        "stages_ms_mean": {"mesh": mean("mesh_ms"), "assembly": mean("assemble_ms"), "precond_setup": mean("prep_ms"),
                           "solve": mean("solve_ms")},
        "iters_mean": mean("iters"), "iters_max": int(max(r["iters"] for r in rows)),
        "true_rel_max": float(max(r["true_rel"] for r in rows)), "fallbacks": int(sum(r["fallback"] for r in rows)),
        "tets_step0": int(rows[0]["n_elems"]), "nnz_step0": int(rows[0]["nnz"]),
        "peak_library_bytes": int(mem["peak"]), "peak_device_used_bytes": int(used_peak),
        "device_total_bytes": int(mem["total"]),
        "weights": os.path.relpath(weights, ROOT),
    }


COPY_BYTES = 256 << 20
def traffic_file(kernel):
    return os.path.join(ROOT, "profiles", f"{kernel}_dam1M_traffic.txt")


def measured_traffic(kernel, n, nnz, fp32):
    """HBM bytes per direction-kernel launch from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    reading of a copy kernel of known size), when it was taken on this
    workload; else None."""
    try:
        kv = {}
        with open(traffic_file(kernel)) as fh:
            for line in fh:
                if "=" in line and not line.startswith("#"):
                    k, v = line.split("=", 1)
                    kv[k.strip()] = v.strip()
        if int(kv["n"]) != n or int(kv["nnz"]) != nnz or kv["precision"] != ("fp32" if fp32 else "fp64"):
            return None
        return int(float(kv["traffic_bytes_per_launch"]))
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--solver", default="learned", choices=["learned", "mixed", "fsai", "fp64"])
    ap.add_argument("--weights", default=WEIGHTS, help="weight file prefix (.bin + .json) of the learned solver")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 companion measurement")
    ap.add_argument("--tank3d", action="store_true",
                    help="also time the 3D step: 1,996,920-point sloshing tank, alpha-shape tetrahedra, "
                         "P1 assembly, mixed FP32/FP64 Jacobi-PCG, FP64 Jacobi-PCG and FSAI mixed PCG to rtol 1e-8")
    ap.add_argument("--sequence3d", action="store_true",
                    help="the 3D advect-remesh-solve sequence instead of the headline: 8,012,034-point tank, "
                         "--seq-steps steps of LBVH + alpha-shape tetrahedra + P1 + learned mixed PCG (rtol 1e-8), "
                         "advected in the seeded 3D divergence-free field; prints one JSON line")
    ap.add_argument("--seq-steps", type=int, default=50)
    ap.add_argument("--weights3d", default=WEIGHTS3D, help="weight prefix of the 3D learned solver")
    ap.add_argument("--only-tank3d", action="store_true",
                    help="time only the 2M 3D tank step (mixed, FP64 Jacobi, FSAI and learned PCG) and print its JSON")
    ap.add_argument("--components", action="store_true",
                    help="time the path's components at 1M points (LBVH step, TRBVH, Mesh2D phases, P1, CG "
                         "kernels) and print them as one JSON line instead of the headline")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="counter runs: one step, then 20 launches of the direction kernel (and, on the L L^T "
                         "path, of the update kernel and the lower apply) and 5 copies of 256 MiB (the "
                         "byte-counter calibration); prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the solver runs on a single "
              "device: no row partition, halo exchange or collective reductions are implemented)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import stream_amd
    from stream_amd import clouds
    from stream_amd.pfem import PfemStep2D

    if stream_amd.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.sequence3d:
        print(json.dumps(sequence3d(args.seq_steps, args.weights3d)), flush=True)
        return
    if args.only_tank3d:
        print(json.dumps(tank3d_leg(args.steps, args.weights3d)), flush=True)
        return
    pts, al = clouds.dam_break()
    f = clouds.divergence_source(pts, clouds.fourier_modes(11))
    step = PfemStep2D(pts, al, f)
    net = None
    if args.solver == "learned":
        from stream_amd.learned import PrecLearned

        if not os.path.exists(args.weights + ".bin"):
            print(f"learned solver: weight file {args.weights}.bin not found "
                  "(train with python -m stream_amd.train_learned --out <prefix>)", file=sys.stderr)
            sys.exit(1)
        net = PrecLearned(PrecLearned.load_params(args.weights))
        step.set_learned(net)

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

    if args.components:
        print(json.dumps(components(step, pts, al, net)), flush=True)
        return

    if args.profile_kernels:
        import ctypes

        r = step.run(args.solver, rtol=1e-8)
        llt = args.solver in ("learned", "fsai")
        kms = step.time_dir_kernel(fp32=args.solver != "fp64", reps=20, llt=llt)
        if llt:
            step.time_update_kernel(reps=20)
            step.time_update_kernel(reps=20, lower=True)
        cms = ctypes.c_double(0)
        stream_amd.check(stream_amd.lib.Stream_copy_calibration(ctypes.c_uint64(COPY_BYTES), 5, ctypes.byref(cms)))
        print(f"[profile] n={pts.shape[0]} nnz={r['nnz']} dir kernel {kms * 1000:.2f} us, "
              f"copy {COPY_BYTES} B {cms.value * 1000:.1f} us", flush=True)
        return

    ms, infos = timed(args.solver, args.steps, args.warmup)
    last = infos[-1]
    if any(i["fallback"] for i in infos):
        print("[bench] the learned factor was rejected (Jacobi fallback used)", file=sys.stderr)
    print(f"[bench] {args.solver}: {ms:.2f} ms/step  mesh {last['mesh_ms']:.2f}  asm {last['assemble_ms']:.2f}  "
          f"prep {last['prep_ms']:.2f}  solve {last['solve_ms']:.2f} ms  iters {last['iters']}  "
          f"rel {last['true_rel']:.2e}", file=sys.stderr, flush=True)
    fp64 = None
    if not args.no_fp64 and args.solver != "fp64":
        ms64, inf64 = timed("fp64", max(1, args.steps // 2), 1)
        fp64 = dict(ms_per_step=ms64, iters=inf64[-1]["iters"], true_rel=inf64[-1]["true_rel"],
                    solve_ms=inf64[-1]["solve_ms"])
        print(f"[bench] fp64: {ms64:.2f} ms/step iters {fp64['iters']}", file=sys.stderr, flush=True)
    fsai = None
    if not args.no_fp64 and args.solver == "learned":
        # the classical FSAI factor on the same pattern (the minimiser of the
        # network's training loss), timed the same way
        msf, inff = timed("fsai", max(1, args.steps // 2), 1)
        fsai = dict(ms_per_step=msf, steps_per_s=1000.0 / msf, iters=inff[-1]["iters"],
                    true_rel=inff[-1]["true_rel"], solve_ms=inff[-1]["solve_ms"], prep_ms=inff[-1]["prep_ms"])
        print(f"[bench] fsai: {msf:.2f} ms/step iters {fsai['iters']}", file=sys.stderr, flush=True)

    # PCIe-inclusive rate: the points, alpha and source copied from host
    # memory into the step before every mesh-and-solve
    k_pcie = max(2, args.steps // 4)
    stream_amd.synchronize()
    t0 = time.perf_counter()
    for _ in range(k_pcie):
        step.set_points(pts, al, f)
        step.run(args.solver, rtol=1e-8)
    stream_amd.synchronize()
    ms_pcie = (time.perf_counter() - t0) * 1000.0 / k_pcie
    print(f"[bench] {args.solver} with host-to-device points: {ms_pcie:.2f} ms/step", file=sys.stderr, flush=True)

    # dominant kernel: on the L L^T path the longest of the direction
    # kernel, the update kernel (x, r' and the L^T apply) and the lower apply
    # (L apply, next direction), the other two reported beside it; on the
    # Jacobi paths the direction kernel (d = z + beta d_old, SpMV, d.q)
    fp32 = args.solver != "fp64"
    llt = args.solver in ("learned", "fsai")
    n = pts.shape[0]
    nnz = last["nnz"]

    def roof(kname, kms, alg_bytes, traffic):
        achieved = alg_bytes / (kms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": None if traffic is None else
                "committed counter summary (profiles/, rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE of "
                "bench.py --profile-kernels), not measured in this run",
                "kernel": kname, "kernel_ms": round(kms, 5), "alg_bytes_per_launch": int(alg_bytes),
                "copy_same_bytes": copy_ceiling(alg_bytes, achieved)}

    def copy_ceiling(alg_bytes, achieved):
        # a plain device copy (16-byte loads and stores) that moves the same
        # bytes in one launch, timed the same way (back-to-back launches
        # between HIP events): what the HBM sustains at this transfer size
        import ctypes

        half = int(alg_bytes) // 2 // 16 * 16
        cms = ctypes.c_double(0)
        stream_amd.check(stream_amd.lib.Stream_copy_calibration(ctypes.c_uint64(half), 200, ctypes.byref(cms)))
        gbs = 2 * half / (cms.value * 1e-3) / 1e9
        return {"bytes_moved": 2 * half, "kernel_ms": round(cms.value, 5), "achieved": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel_vs_copy": round(achieved / gbs, 4)}

    kms = step.time_dir_kernel(fp32=fp32, reps=200, llt=llt)
    dname = "llt_dir_kernel" if llt else "cg_dir_kernel"
    dir_roof = roof(f"{dname}<{'float, true' if fp32 else 'double, false'}>", kms,
                    sell_dir_bytes(n, nnz, 4 if fp32 else 8, llt=llt), measured_traffic(dname, n, nnz, fp32))
    if llt:
        ums = step.time_update_kernel(reps=200)
        lms = step.time_update_kernel(reps=200, lower=True)
        upd_roof = roof("llt_update_kernel<float, double>", ums, sell_update_bytes(n, nnz),
                        measured_traffic("llt_update_kernel", n, nnz, fp32))
        lo_roof = roof("llt_lo_kernel<float>", lms, sell_lo_bytes(n, nnz),
                       measured_traffic("llt_lo_kernel", n, nnz, fp32))
        three = sorted([("direction_kernel", dir_roof), ("update_kernel", upd_roof), ("lower_kernel", lo_roof)],
                       key=lambda kv: -kv[1]["kernel_ms"])
        roofline = dict(three[0][1])
        roofline["role"] = three[0][0]
        for key, rf in three[1:]:
            roofline[key] = rf
    else:
        roofline = dir_roof

    mfma = None
    if net is not None:
        step.run(args.solver, rtol=1e-8)  # leaves the system assembled
        net.bind_step(step)
        gms, gflops = net.time_gemm(reps=100)
        tf = gflops / (gms * 1e-3) / 1e12
        # the layer GEMM streams FP32 activations: H [n][64] in, [P|Q] [n][128] out
        gbytes = n * (64 + 128) * 4
        gbs = gbytes / (gms * 1e-3) / 1e9
        mfma = {"kernel": "gemm_rows_kernel<64, 64, 128, 0> (layer-0 [P|Q] = H W_pq, bf16 MFMA, fp32 accumulate)",
                "achieved": round(tf, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tf / BF16_PEAK_TFLOPS, 5), "kernel_ms": round(gms, 5), "flops_per_launch": gflops,
                "arith_intensity": round(gflops / gbytes, 2),
                "bound": "hbm (arithmetic intensity below the bf16 ridge of 312 FLOP/B)",
                "bytes_per_launch": int(gbytes), "achieved_gbs": round(gbs, 1),
                "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}

    # mesh-and-assembly stages (HIP events between the stages' launches)
    step.run(args.solver, rtol=1e-8)
    sec = step.time_sections(reps=5)
    sb = section_bytes(n, nnz, int(last["n_elems"]), int(nnz - n))
    stages = {}
    for k, b in sb.items():
        gbs = b / (sec[k] * 1e-3) / 1e9
        stages[k] = {"ms": round(sec[k], 5), "alg_bytes": int(b), "achieved_gbs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    stages["mesh_and_assembly_ms"] = round(sec["mesh_and_assembly"], 4)
    stages["sort_per_pass_ms"] = round(sec["sort"] / 8, 5)

    # the solutions of every solver, for the error against the reference
    sols = {}
    for solver in ("fp64", "mixed", "fsai") + (("learned",) if net is not None else ()):
        step.run(solver, rtol=1e-8)
        sols[solver] = step.solution(morton=True)

    imr = imr26_leg()

    tank = None
    if args.tank3d:
        tank = tank3d_leg()

    cpu = None
    err = None
    if not args.no_cpu:
        cpu, ref = cpu_baseline(pts, al, f)
        xr = ref["x"]
        nr = float(np.linalg.norm(xr))
        err = {"reference": (f"oracle FP64 Jacobi-PCG on the host, continued from its 1e-8 solution to a true "
                             f"relative residual of {ref['rel']:.1e} (FP64 rounding floor ~3e-12; a sparse direct "
                             f"solve does not finish in 15 min at 1M)"),
               **{k: float(np.linalg.norm(v - xr) / nr) for k, v in sols.items()}}

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
            "solver": SOLVER_LABEL[args.solver],
            "weights": os.path.relpath(args.weights, ROOT) if net is not None else None,
            "nnz": int(nnz),
            "triangles": int(last["n_elems"]),
        },
        "stages_ms": {"mesh": round(last["mesh_ms"], 3), "assembly": round(last["assemble_ms"], 3),
                      "precond_setup": round(last["prep_ms"], 3), "solve": round(last["solve_ms"], 3)},
        "iters": last["iters"],
        "true_rel": last["true_rel"],
        "fp64_jacobi": None if fp64 is None else {
            "steps_per_s": round(1000.0 / fp64["ms_per_step"], 4), **fp64},
        "speedup_vs_fp64_jacobi": None if fp64 is None else round(fp64["ms_per_step"] / ms, 3),
        "fsai": fsai,
        "faster_llt_preconditioner": None if fsai is None else ("learned" if ms <= fsai["ms_per_step"] else "fsai"),
        "pcie_inclusive": {"steps_per_s": round(1000.0 / ms_pcie, 4), "ms_per_step": round(ms_pcie, 3),
                           "steps": k_pcie, "note": "points, alpha and source copied host-to-device every step"},
        "err_vs_fp64_reference": err,
        "mesh_stages": stages,
        "imr26": imr,
        "roofline": roofline,
        "mfma": mfma,
        "tank3d": tank,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

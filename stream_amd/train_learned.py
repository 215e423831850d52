"""Train the learned L L^T preconditioner on the advected dam-break sequence
and evaluate it against Jacobi-PCG.

Sequence: the 1,000,000-point dam-break cloud (704 x 1407 jittered column +
9,472 scattered points, alpha = 1.2h), advected by forward Euler (dt = 1e-3)
in the divergence-free velocity of a 16-mode stream function (seed 33), and
re-meshed and re-assembled on the GPU at every step; the source is the
divergence of the 8-mode velocity (seed 11) at the current points.  The
network trains on steps 0-149 (every snapshot's own matrix, Adam on the
GPU through the library's forward and backward kernels) and is evaluated
on steps 150-199: PCG iterations and times of the learned mixed PCG, the
mixed Jacobi-PCG and the FP64 Jacobi-PCG, each verified to a true FP64
relative residual of 1e-8.

    python -m stream_amd.train_learned --out stream_amd/weights/mpnn_dam2d

With --dim 3 the sequence is the sloshing-tank cloud (jittered cubic grid,
alpha = 1.3h; --points picks the size) advected in the divergence-free 3D
field of clouds.potential_modes (seed 44), source z + sin(3x) cos(2y):

    python -m stream_amd.train_learned --dim 3 --points 2000000 --out stream_amd/weights/mpnn_tank3d
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import clouds, synchronize
from .learned import PrecLearned
from .pfem import PfemStep2D, PfemStep3D


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _source3(pts):
    return np.sin(3.0 * pts[:, 0]) * np.cos(2.0 * pts[:, 1]) + pts[:, 2]


def make_step(n_points: int, dim: int = 2):
    if dim == 3:
        pts, al = clouds.sloshing_tank() if n_points == 1_996_920 else clouds.sloshing_tank_scaled(n_points)
        return PfemStep3D(pts, al, _source3(pts)), pts, al
    if n_points == 1_000_000:
        pts, al = clouds.dam_break()
    else:
        pts, al = clouds.dam_break_scaled(n_points)
    step = PfemStep2D(pts, al, np.zeros(len(pts)))
    step.set_source_modes(clouds.fourier_modes(11))
    return step, pts, al


def reset(step, pts, al):
    if isinstance(step, PfemStep3D):
        step.set_points(pts, al, _source3(pts))
        return
    step.set_points(pts, al, np.zeros(len(pts)))
    step.set_source_modes(clouds.fourier_modes(11))


def field(step):
    return clouds.potential_modes() if isinstance(step, PfemStep3D) else clouds.stream_modes()


def train(step, pts, al, net, epochs=4, per_snapshot=2, train_steps=150, lr=2e-3, lr_min=2e-4, dt=1e-3,
          log=_log):
    modes = field(step)
    total = epochs * train_steps * per_snapshot
    k = 0
    history = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        reset(step, pts, al)
        losses = []
        for s in range(train_steps):
            step.assemble()
            net.bind_step(step)
            for it in range(per_snapshot):
                a = 0.5 * (1.0 + math.cos(math.pi * k / max(1, total - 1)))
                rate = lr_min + (lr - lr_min) * a
                last = it == per_snapshot - 1
                loss = net.train_step(rate, read_loss=last)
                k += 1
            losses.append(loss)
            if s % 50 == 0:
                log(f"[train] epoch {ep} step {s}: loss {loss:.5f} {time.perf_counter() - t0:.1f}s")
            step.advect(modes, dt, 1)
        history.append(float(np.mean(losses)))
        log(f"[train] epoch {ep}: mean loss {history[-1]:.5f} (first {losses[0]:.5f}, last {losses[-1]:.5f}) "
            f"{time.perf_counter() - t0:.1f}s")
    return history, time.perf_counter() - t0


def evaluate(step, pts, al, net, start=150, stop=200, dt=1e-3, solvers=("learned", "mixed", "fp64", "fsai"),
             log=_log):
    modes = field(step)
    reset(step, pts, al)
    step.advect(modes, dt, start)
    step.set_learned(net)
    rows = []
    for s in range(start, stop):
        row = {"step": s}
        for sol in solvers:
            synchronize()
            t = time.perf_counter()
            r = step.run(sol, rtol=1e-8)
            synchronize()
            r["wall_ms"] = (time.perf_counter() - t) * 1000.0
            row[sol] = r
        rows.append(row)
        if s % 10 == 0:
            log(f"[eval] step {s}: " + ", ".join(f"{sol} {row[sol]['iters']} it {row[sol]['wall_ms']:.1f} ms"
                                              for sol in solvers))
        step.advect(modes, dt, 1)
    summary = {}
    for sol in solvers:
        it = [r[sol]["iters"] for r in rows]
        summary[sol] = dict(mean_iters=float(np.mean(it)), max_iters=int(np.max(it)),
                            mean_solve_ms=float(np.mean([r[sol]["solve_ms"] for r in rows])),
                            mean_step_ms=float(np.mean([r[sol]["wall_ms"] for r in rows])),
                            max_true_rel=float(np.max([r[sol]["true_rel"] for r in rows])),
                            fallbacks=int(sum(bool(r[sol]["fallback"]) for r in rows)))
    return summary, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=2, choices=[2, 3])
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--per-snapshot", type=int, default=2)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--train-steps", type=int, default=150)
    ap.add_argument("--eval-stop", type=int, default=200)
    ap.add_argument("--out", default=None, help="weight file prefix (.bin + .json)")
    ap.add_argument("--summary", default=None, help="JSON summary path")
    args = ap.parse_args()
    step, pts, al = make_step(args.points, args.dim)
    net = PrecLearned()
    hist, secs = train(step, pts, al, net, epochs=args.epochs, per_snapshot=args.per_snapshot,
                       train_steps=args.train_steps, lr=args.lr, lr_min=args.lr / 10)
    summary, rows = evaluate(step, pts, al, net, start=args.train_steps, stop=args.eval_stop)
    meta = dict(points=args.points, epochs=args.epochs, per_snapshot=args.per_snapshot, lr=args.lr,
                train_steps=args.train_steps, train_seconds=secs, loss_per_epoch=hist, eval=summary,
                dim=args.dim,
                sequence=("sloshing-tank cloud advected by the 16-mode divergence-free potential field (seed 44), "
                          "dt 1e-3" if args.dim == 3 else
                          "dam-break cloud advected by the 16-mode stream function (seed 33), dt 1e-3"))
    if args.out:
        net.save(args.out, meta)
    out = dict(meta, per_step=[{k: (v if k == "step" else {kk: vv for kk, vv in v.items()
                                                          if kk in ("iters", "solve_ms", "wall_ms", "true_rel")})
                                for k, v in r.items()} for r in rows])
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(out, f, indent=1)
    _log(json.dumps(dict(meta)))


if __name__ == "__main__":
    main()

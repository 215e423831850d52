"""The benchmark configurations at full size on the GPU.

* 3D sloshing tank, 1,996,920 FP64 points (alpha = 1.3h, 63-bit Morton):
  every alpha-shape output array bit-exact against the oracle, the P1
  tetrahedral matrix and load bit-exact against the oracle's assembly of
  the same mesh, and the mixed FP32/FP64 PCG to a true FP64 relative
  residual <= 1e-8 recomputed here.
* Learned sequence: the 1M dam-break cloud advected on the device (16-mode
  stream function, seed 33, dt = 1e-3) to held-out steps of the training
  sequence (150, 175, 199); the shipped 2D network (trained on steps 0-149)
  against FP64 Jacobi-PCG, both verified at 1e-8 here, and the iteration
  gain.
* 3D tank at 8,012,034 points: step 0 of the sequence bit-exact against the
  oracle (every alpha-shape array), then advect-remesh-solve steps with the
  shipped 3D network, each verified at 1e-8, with the peak device memory.
The oracle runs on the host cores (OpenMP, OMP_NUM_THREADS)."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as ol
from stream_amd import clouds

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "stream_amd", "weights")


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _source3(xyz):
    return np.sin(3.0 * xyz[:, 0]) * np.cos(2.0 * xyz[:, 1]) + xyz[:, 2]


def _alpha3d_gpu(pts, al):
    from stream_amd.mesh import AlphaShape3D

    a = AlphaShape3D(dtype=np.float64, morton_bits=21)
    a.set_nodes(pts, al)
    a.compute()
    row, neig, eb = a.get_edges()
    out = dict(elem=a.get_elem(), ordered=a.get_ordered_nodes(), perm=a.get_permutation(), row=row, neig=neig,
               edge_is_bnd=eb, node_is_bnd=a.get_boundary_nodes(), nclamped=a.get_nclamped())
    del a
    return out


def _compare_alpha3d(g, o, label):
    print(f"{label}: {len(g['elem'])} tets, {len(g['neig'])} directed edges, clamped {g['nclamped']}", flush=True)
    for k in ("perm", "ordered", "row", "neig", "edge_is_bnd", "node_is_bnd", "elem"):
        assert np.array_equal(g[k], o[k]), k


def _true_rel(row, col, val, b, x):
    n = len(row) - 1
    A = sp.csr_matrix((val, col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def test_tank_2m_mesh_assembly_bit_exact_and_mixed_solve():
    from stream_amd.pfem import PfemStep3D

    _say("3D tank 1,996,920 points: oracle alpha-shape")
    pts, al = clouds.sloshing_tank()
    assert len(pts) == 1_996_920
    o = ol.ashape3d(pts, al, bits=21, dtype=np.float64)
    _say("3D tank 2M: GPU alpha-shape")
    g = _alpha3d_gpu(pts, al)
    _compare_alpha3d(g, o, "tank 2M")
    del g
    _say("3D tank 2M: step (P1 + mixed PCG)")
    f = _source3(pts)
    st = PfemStep3D(pts, al, f)
    r = st.run("mixed", rtol=1e-8)
    elem, bnd, xyz, perm = st.mesh()
    assert np.array_equal(elem, o["elem"]) and np.array_equal(perm, o["perm"]) and np.array_equal(xyz, o["ordered"])
    row, col, val, b = st.system()
    _say("3D tank 2M: oracle P1 assembly")
    o_row, o_col, o_val, o_b = ol.assemble_p1_tet(o["ordered"], o["elem"], o["node_is_bnd"], f[o["perm"]])
    assert np.array_equal(row, o_row) and np.array_equal(col, o_col)
    assert np.array_equal(val, o_val) and np.array_equal(b, o_b)
    p = st.solution()
    rel = _true_rel(row, col, val, b, p[perm])
    print(f"tank 2M mixed: {r['iters']} iterations, true rel {rel:.3e}, nnz {len(col)}, "
          f"mesh {r['mesh_ms']:.1f} ms, assembly {r['assemble_ms']:.1f} ms, solve {r['solve_ms']:.1f} ms", flush=True)
    assert rel <= 1e-8


def test_dam_break_heldout_steps_learned_vs_jacobi():
    from stream_amd.learned import PrecLearned
    from stream_amd.pfem import PfemStep2D

    _say("1M dam-break sequence: held-out steps 150, 175, 199")
    pts, al = clouds.dam_break()
    st = PfemStep2D(pts, al, np.zeros(len(pts)))
    st.set_source_modes(clouds.fourier_modes(11))
    st.set_learned(PrecLearned(PrecLearned.load_params(os.path.join(WEIGHTS, "mpnn_dam2d"))))
    modes = clouds.stream_modes()
    done = 0
    for target in (150, 175, 199):
        st.advect(modes, 1e-3, target - done)
        done = target
        info = {}
        for solver in ("learned", "fp64"):
            r = st.run(solver, rtol=1e-8)
            row, col, val, b = st.system()
            rel = _true_rel(row, col, val, b, st.solution(morton=True))
            assert rel <= 1e-8 and not r["fallback"], (target, solver, rel)
            info[solver] = r
        gain = info["fp64"]["iters"] / info["learned"]["iters"]
        print(f"step {target}: learned {info['learned']['iters']} it ({info['learned']['solve_ms']:.1f} ms solve), "
              f"FP64 Jacobi {info['fp64']['iters']} it ({info['fp64']['solve_ms']:.1f} ms solve), "
              f"iteration gain {gain:.2f}x", flush=True)
        assert gain > 1.6
        assert info["learned"]["solve_ms"] < info["fp64"]["solve_ms"]


def test_tank_8m_step0_bit_exact_and_learned_sequence():
    import stream_amd
    from stream_amd.learned import PrecLearned
    from stream_amd.pfem import PfemStep3D

    _say("3D tank 8,012,034 points: GPU alpha-shape, step 0")
    pts, al = clouds.sloshing_tank(*clouds.TANK8M)
    assert len(pts) == 8_012_034
    g = _alpha3d_gpu(pts, al)
    _say("3D tank 8M: oracle alpha-shape, step 0")
    o = ol.ashape3d(pts, al, bits=21, dtype=np.float64)
    _compare_alpha3d(g, o, "tank 8M step 0")
    del g, o
    _say("3D tank 8M: learned advect-remesh-solve steps")
    f = _source3(pts)
    st = PfemStep3D(pts, al, f)
    st.set_learned(PrecLearned(PrecLearned.load_params(os.path.join(WEIGHTS, "mpnn_tank3d"))))
    modes = clouds.potential_modes()
    stream_amd.device_bytes(reset_peak=True)
    used_peak = 0
    for k in range(3):
        r = st.run("learned", rtol=1e-8)
        mem = stream_amd.device_bytes()
        used_peak = max(used_peak, mem["total"] - mem["free"])
        print(f"8M step {k}: {r['n_elems']} tets, nnz {r['nnz']}, {r['iters']} it, true rel {r['true_rel']:.2e}, "
              f"mesh {r['mesh_ms']:.0f} ms, assembly {r['assemble_ms']:.0f} ms, set-up {r['prep_ms']:.0f} ms, "
              f"solve {r['solve_ms']:.0f} ms", flush=True)
        assert r["true_rel"] <= 1e-8 and not r["fallback"]
        if k < 2:
            st.advect(modes, 1e-3, 1)
    # the last solve's residual recomputed on the host
    row, col, val, b = st.system()
    _, _, _, perm = st.mesh()
    rel = _true_rel(row, col, val, b, st.solution()[perm])
    mem = stream_amd.device_bytes()
    print(f"8M: host-checked true rel {rel:.3e}; peak device memory {used_peak / 1e9:.1f} GB used of "
          f"{mem['total'] / 1e9:.0f} GB (library high-water mark {mem['peak'] / 1e9:.1f} GB)", flush=True)
    assert rel <= 1e-8

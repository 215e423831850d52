"""The IMR26 runner on the GPU: the README's two timed workloads give the
golden element counts (1,877,465 triangles for 1M points at alpha = 0.0012;
526,042 tetrahedra for 100k points at alpha = 0.022, the exact count --
the README's FP32 build reports 526,040, see tests/test_oracle3d.py), and
the file runner meshes clouds written in the dataset's format to the same
elements as the alpha-shape classes and the oracle."""
import io

import numpy as np
import pytest

import oracle_lib as O
from stream_amd import clouds, imr26

pytestmark = pytest.mark.gpu


def test_uniform2d_readme_workload():
    rows = imr26.run_uniform(2, 1_000_000, 0.0012, runs=3)
    imr26.print_rows(rows)
    assert all(r["ntri"] == 1_877_465 for r in rows)
    assert all(r["tlbvh"] > 0 and r["tashape"] > 0 and r["tcompress"] > 0 for r in rows)
    assert all(abs(r["ttot"] - (r["tlbvh"] + r["tashape"])) < 1e-9 for r in rows)


def test_uniform3d_readme_workload():
    rows = imr26.run_uniform(3, 100_000, 0.022, runs=2)
    imr26.print_rows(rows)
    assert all(r["ntri"] == 526_042 for r in rows)


def test_file_runner_on_written_clouds(tmp_path):
    pts, al = clouds.dam_break_scaled(20_000)
    f2 = tmp_path / "dam.bin"
    imr26.write_cloud(f2, pts, al)
    rng = np.random.default_rng(5)
    p3 = rng.random((8000, 3)).astype(np.float32)
    f3 = tmp_path / "cube.bin"
    imr26.write_cloud(f3, p3, 0.06)
    out = io.StringIO()
    r2 = imr26.run_files([f2], 2, nruns=2, out=out)
    r3 = imr26.run_files([f3], 3, nruns=2, out=out)
    print(out.getvalue())
    c2, a2 = imr26.read_cloud(f2, 2)
    o2 = O.ashape2d(c2, a2, bits=32, dtype=np.float32)
    assert r2[0]["nodes"] == len(pts) and r2[0]["elems"] == len(o2["elem"])
    c3, a3 = imr26.read_cloud(f3, 3)
    o3 = O.ashape3d(c3, a3, bits=21, dtype=np.float32)
    assert r3[0]["elems"] == len(o3["elem"])
    assert "[2D] Time to mesh" in out.getvalue() and "tetrahedrons" in out.getvalue()

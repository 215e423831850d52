"""The IMR26 point-cloud file format and generator (host side): records of
float32 (x, y, [z,] alpha) as testcases/IMR26/README.md:138-145 reads them,
and the benchmarks' srand/rand generator, which is the one the README's
golden counts were drawn with (the oracle reproduces 1,877,465 triangles
from it: tests/test_oracle.py)."""
import io

import numpy as np
import pytest

import oracle_lib as O
from stream_amd import imr26


@pytest.mark.parametrize("dim", [2, 3])
def test_cloud_file_round_trip(tmp_path, dim):
    rng = np.random.default_rng(dim)
    c = rng.random((1000, dim)).astype(np.float32)
    a = rng.random(1000).astype(np.float32)
    f = tmp_path / f"cloud{dim}.bin"
    imr26.write_cloud(f, c, a)
    raw = np.fromfile(f, dtype=np.float32).reshape((-1, dim + 1))  # the README's own reader
    assert np.array_equal(raw[:, :dim], c) and np.array_equal(raw[:, dim], a)
    c2, a2 = imr26.read_cloud(f, dim)
    assert np.array_equal(c2, c) and np.array_equal(a2, a)
    assert c2.flags.c_contiguous and a2.flags.c_contiguous


def test_cloud_file_rejects_partial_records(tmp_path):
    f = tmp_path / "bad.bin"
    np.arange(10, dtype=np.float32).tofile(f)
    with pytest.raises(ValueError):
        imr26.read_cloud(f, 2)  # 10 values are not whole (x, y, alpha) records
    with pytest.raises(ValueError):
        imr26.read_cloud(f, 4)


def test_uniform_generator_is_the_benchmarks():
    for dim in (2, 3):
        assert np.array_equal(imr26.uniform(2000, dim, 42), O.glibc_uniform(2000, dim, 42))
    assert not np.array_equal(imr26.uniform(10, 2, 42), imr26.uniform(10, 2, 43))


def test_file_runner_reports_missing_files(tmp_path):
    out = io.StringIO()
    res = imr26.run_files([tmp_path / "nope.bin"], 2, nruns=1, out=out)
    assert res == [] and "Could not find input file" in out.getvalue()


def test_paper_tables():
    # testcases/IMR26/README.md:34-41, 80-86
    assert imr26.PAPER_2D[4] == (1_000_000, 0.0012) and len(imr26.PAPER_2D) == 6
    assert imr26.PAPER_3D[2] == (100_000, 0.024) and len(imr26.PAPER_3D) == 5

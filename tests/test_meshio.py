"""The library's text-table reader (mesh and stimulus files) against
np.loadtxt, CPU only: identical arrays (values parse correctly rounded,
like Python's float()), the reference's count checks, and errors on ragged
or non-numeric rows."""
import numpy as np
import pytest

from torchcor_amd import _lib as L
from torchcor_amd import meshes
from torchcor_amd.core.mesh import MeshReader, read_table
from torchcor_amd.core.stimulation import load_stimulus_region


def _write(path, header, rows):
    with open(path, "w") as f:
        f.write(header)
        for r in rows:
            f.write(r + "\n")


def test_floats_match_loadtxt_across_thread_splits(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.normal(size=(120_000, 3)) * 10.0 ** rng.integers(-8, 9, size=(120_000, 3))
    p = tmp_path / "x.pts"
    np.savetxt(p, a, header=str(a.shape[0]), comments="", fmt="%.17g")   # > 1 MiB: several threads
    got = read_table(p, 1)
    assert np.array_equal(got, np.loadtxt(p, skiprows=1))
    assert np.array_equal(got, a)                                           # round trip is exact


def test_ints_tags_comments_blank_lines(tmp_path):
    p = tmp_path / "m.elem"
    rows = [f"Tt {i} {i + 1} {i + 2} {i + 3} {i % 4}" for i in range(1000)]
    rows.insert(10, "")
    rows.insert(20, "   # a comment line")
    rows[30] += "   # trailing comment"
    _write(p, "1000\n", rows)
    got = read_table(p, 1, col0=1, dtype=np.int64)
    want = np.loadtxt(p, skiprows=1, usecols=range(1, 6), dtype=int)
    assert got.shape == (1000, 5) and np.array_equal(got, want)


def test_signs_exponents_and_tabs(tmp_path):
    p = tmp_path / "t.lon"
    _write(p, "1\n", ["+1.5e-3\t-2\t3.25E+2", "  -0.0  +7 1e308", "4.9406564584124654e-324 2 3"])
    assert np.array_equal(read_table(p, 1), np.loadtxt(p, skiprows=1))


def test_ragged_and_non_numeric_rows_raise(tmp_path):
    p = tmp_path / "r.pts"
    _write(p, "3\n", ["1 2 3", "4 5", "7 8 9"])
    with pytest.raises(L.TorchcorError):
        read_table(p, 1)
    q = tmp_path / "n.pts"
    _write(q, "2\n", ["1 2 3", "4 x 6"])
    with pytest.raises(L.TorchcorError):
        read_table(q, 1)
    with pytest.raises(L.TorchcorError):
        read_table(tmp_path / "missing.pts", 1)


def test_mesh_reader_and_vtx(tmp_path):
    nodes, elems, regions, fib = meshes.cube_workload(9)
    meshes.write_mesh(tmp_path, nodes, elems, regions, fib)
    n2, e2, r2, f2 = MeshReader(tmp_path).read()
    assert np.allclose(n2, nodes, rtol=0, atol=1e-9)
    assert np.array_equal(e2, elems) and np.array_equal(r2, regions)
    assert np.allclose(f2, fib, rtol=0, atol=1e-12)
    ids = meshes.corner_nodes(9, 3)
    v = meshes.write_vtx(tmp_path / "s.vtx", ids)
    assert np.array_equal(load_stimulus_region(v).numpy(), ids)


def test_count_mismatch_raises_like_the_reference(tmp_path):
    nodes, elems, regions, fib = meshes.cube_workload(5)
    meshes.write_mesh(tmp_path, nodes, elems, regions, fib)
    pts = next(tmp_path.glob("*.pts"))
    lines = pts.read_text().splitlines()
    pts.write_text(str(int(lines[0]) + 1) + "\n" + "\n".join(lines[1:]) + "\n")
    with pytest.raises(ValueError):
        MeshReader(tmp_path).read_nodes()
    v = tmp_path / "bad.vtx"
    _write(v, "5\nextra\n", ["1", "2", "3"])
    with pytest.raises(Exception):
        load_stimulus_region(v)

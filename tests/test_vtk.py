"""Legacy VTK output (visualize.py VTK3D / VTK3DSurface, pyvista's legacy
binary files), CPU only: the file is parsed back by a reader written from
the legacy-format specification and must give the same points, cells,
types and point scalars."""
import numpy as np

from torchcor_amd import meshes
from torchcor_amd.core.vtk import VTK3D, VTK3DSurface


def _read_legacy(path):
    data = path.read_bytes()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    assert line().startswith("# vtk DataFile Version")
    line()
    assert line() == "BINARY" and line() == "DATASET UNSTRUCTURED_GRID"
    _, n, t = line().split()
    dt = np.dtype(">f8" if t == "double" else ">f4")
    pts = np.frombuffer(data, dt, int(n) * 3, pos).reshape(-1, 3)
    pos += pts.nbytes + 1
    _, E, size = line().split()
    conn = np.frombuffer(data, ">i4", int(size), pos)
    pos += conn.nbytes + 1
    _, E2 = line().split()
    types = np.frombuffer(data, ">i4", int(E2), pos)
    pos += types.nbytes + 1
    assert line() == f"POINT_DATA {n}"
    _, name, t2, _ = line().split()
    assert line() == "LOOKUP_TABLE default"
    sc = np.frombuffer(data, np.dtype(">f8" if t2 == "double" else ">f4"), int(n), pos)
    return pts, conn, types, name, sc


def test_tetra_frame_round_trip(tmp_path):
    nodes, elems = meshes.kuhn_cube(4)
    vals = np.linspace(-80, 20, len(nodes)).astype(np.float32)
    p = VTK3D(nodes, elems).save_frame(vals, tmp_path / "Vm_vtk" / "frame_0.vtk")
    pts, conn, types, name, sc = _read_legacy(p)
    assert np.array_equal(pts, nodes) and name == "colors"
    c = conn.reshape(-1, 5)
    assert np.all(c[:, 0] == 4) and np.array_equal(c[:, 1:], elems)
    assert np.all(types == 10) and np.array_equal(sc, vals)


def test_surface_frame_round_trip(tmp_path):
    nodes, elems = meshes.kuhn_cube(4)
    s_nodes, tris = meshes.surface_of(nodes, elems)
    vals = np.arange(len(s_nodes), dtype=np.float64)
    p = VTK3DSurface(s_nodes, tris).save_frame(vals, tmp_path / "s.vtk")
    pts, conn, types, name, sc = _read_legacy(p)
    c = conn.reshape(-1, 4)
    assert np.array_equal(pts, s_nodes) and np.all(c[:, 0] == 3)
    assert np.array_equal(c[:, 1:], tris) and np.all(types == 5) and np.array_equal(sc, vals)

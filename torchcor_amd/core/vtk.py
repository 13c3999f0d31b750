"""Legacy-VTK output of the solution frames (torchcor/core/visualize.py:93-175,
torchcor/simulator/monodomain.py:229-251).

The reference hands its frames to pyvista (``UnstructuredGrid`` with point
data "colors", ``grid.save``), which writes legacy binary VTK.  The same
files are written here directly: header, POINTS, CELLS (count + node ids per
cell), CELL_TYPES (10 tetrahedron, 5 triangle) and POINT_DATA "colors",
binary big-endian as the legacy format requires.  Points and scalars keep
their dtype (float or double).
"""
from pathlib import Path

import numpy as np
import torch

VTK_TRIANGLE = 5
VTK_TETRA = 10


def _np(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _vtk_type(a):
    return "double" if a.dtype == np.float64 else "float"


def write_unstructured(path, points, cells, cell_type, scalars, name="colors"):
    """One legacy binary VTK unstructured grid with point scalars."""
    points = _np(points)
    if points.dtype not in (np.float32, np.float64):
        points = points.astype(np.float64)
    if points.shape[1] == 2:
        points = np.hstack((points, np.zeros((points.shape[0], 1), dtype=points.dtype)))
    cells = _np(cells).astype(np.int64)
    scalars = _np(scalars)
    if scalars.dtype not in (np.float32, np.float64):
        scalars = scalars.astype(np.float64)
    n, (E, k) = points.shape[0], cells.shape
    conn = np.empty((E, k + 1), dtype=">i4")
    conn[:, 0] = k
    conn[:, 1:] = cells
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("wb") as f:
        f.write(b"# vtk DataFile Version 4.2\nvtk output\nBINARY\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {n} {_vtk_type(points)}\n".encode())
        f.write(points.astype(points.dtype.newbyteorder(">")).tobytes())
        f.write(f"\nCELLS {E} {E * (k + 1)}\n".encode())
        f.write(conn.tobytes())
        f.write(f"\nCELL_TYPES {E}\n".encode())
        f.write(np.full(E, cell_type, dtype=">i4").tobytes())
        f.write(f"\nPOINT_DATA {n}\nSCALARS {name} {_vtk_type(scalars)} 1\n"
                "LOOKUP_TABLE default\n".encode())
        f.write(scalars.reshape(-1).astype(scalars.dtype.newbyteorder(">")).tobytes())
        f.write(b"\n")
    return path


class VTK3D:
    """Tetrahedral frames (visualize.py:140-162)."""

    def __init__(self, vertices, tetrahedrons):
        self.vertices = _np(vertices)
        self.tetrahedrons = _np(tetrahedrons)
        self.n_tetrahedrons = self.tetrahedrons.shape[0]

    def save_frame(self, color_values, frame_path):
        return write_unstructured(frame_path, self.vertices, self.tetrahedrons, VTK_TETRA,
                                  color_values)


class VTK3DSurface:
    """Triangle frames in 3-D (visualize.py:118-137)."""

    def __init__(self, vertices, triangles):
        self.vertices = _np(vertices)
        self.triangles = _np(triangles)
        self.n_triangles = self.triangles.shape[0]

    def save_frame(self, color_values, frame_path):
        return write_unstructured(frame_path, self.vertices, self.triangles, VTK_TRIANGLE,
                                  color_values)

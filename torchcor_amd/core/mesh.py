"""openCARP-style mesh reader (torchcor/core/mesh.py:5-63).

``.pts``: count line, then x y z per node (divided by ``unit_conversion``);
``.elem``: count line, then ``<tag> n0 n1 ... region`` per element (the tag is
dropped, the last column is the region); ``.lon``: a header line, then one
fibre vector per element.  A count line that disagrees with the data raises
``ValueError`` as in the reference.  The tables are parsed by the library's
multi-threaded reader (``tc_table_*``) instead of np.loadtxt.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from .. import _lib as L


def read_table(path, skip, col0=0, ncols=None, dtype=np.float64):
    """Rows of the numeric table in ``path`` after ``skip`` header lines, as
    a (rows, ncols) array of columns col0.. (np.loadtxt semantics: blank
    lines and '#' comments ignored, rows of equal width)."""
    p = str(path).encode()
    rows, cols = C.c_int64(0), C.c_int32(0)
    L.call("tc_table_shape", p, int(skip), C.byref(rows), C.byref(cols))
    if ncols is None:
        ncols = cols.value - col0
    out = np.empty((rows.value, max(ncols, 0)), dtype=dtype)
    if rows.value and ncols > 0:
        fn = "tc_table_read_f64" if dtype == np.float64 else "tc_table_read_i64"
        L.call(fn, p, int(skip), int(col0), int(ncols), rows.value, out.ctypes.data)
    return out


def header_count(path):
    with Path(path).open("r") as f:
        return int(f.readline().strip())


class MeshReader:
    def __init__(self, mesh_dir):
        self.mesh_dir = Path(mesh_dir)
        self.node_file = self.elem_file = self.fibre_file = None
        for f in self.mesh_dir.iterdir():
            if f.suffix == ".pts":
                self.node_file = f
            elif f.suffix == ".elem":
                self.elem_file = f
            elif f.suffix == ".lon":
                self.fibre_file = f
        self.nodes = self.elems = self.regions = self.fibres = None

    def read_nodes(self, unit_conversion=1000):
        expected = header_count(self.node_file)
        nodes = read_table(self.node_file, 1)
        if nodes.shape[0] != expected:
            raise ValueError(f"Mismatch in number of nodes: expected {expected}, "
                             f"but found {nodes.shape[0]}")
        self.nodes = nodes / unit_conversion

    def read_elems(self):
        expected = header_count(self.elem_file)
        data = read_table(self.elem_file, 1, col0=1, dtype=np.int64)
        if data.shape[0] != expected:
            raise ValueError(f"Mismatch: expected {expected}, but found {data.shape[0]}")
        self.elems = data[:, :-1]
        self.regions = data[:, -1]

    def read_fibres(self):
        self.fibres = read_table(self.fibre_file, 1)

    def read(self, unit_conversion=1000):
        self.read_nodes(unit_conversion)
        self.read_elems()
        self.read_fibres()
        return self.nodes, self.elems, self.regions, self.fibres

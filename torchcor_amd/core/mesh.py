"""openCARP-style mesh reader (torchcor/core/mesh.py:5-63).

``.pts``: count line, then x y z per node (divided by ``unit_conversion``);
``.elem``: count line, then ``<tag> n0 n1 ... region`` per element (the tag is
dropped, the last column is the region); ``.lon``: a header line, then one
fibre vector per element.  A count line that disagrees with the data raises
``ValueError`` as in the reference.
"""
from pathlib import Path

import numpy as np


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
        with self.node_file.open("r") as f:
            expected = int(f.readline().strip())
        nodes = np.loadtxt(self.node_file, dtype=float, skiprows=1, ndmin=2)
        if nodes.shape[0] != expected:
            raise ValueError(f"Mismatch in number of nodes: expected {expected}, "
                             f"but found {nodes.shape[0]}")
        self.nodes = nodes / unit_conversion

    def read_elems(self):
        with self.elem_file.open("r") as f:
            expected = int(f.readline().strip())
            first = f.readline().strip().split()
        usecols = list(range(1, len(first)))
        data = np.loadtxt(self.elem_file, dtype=int, skiprows=1, usecols=usecols, ndmin=2)
        if data.shape[0] != expected:
            raise ValueError(f"Mismatch: expected {expected}, but found {data.shape[0]}")
        self.elems = data[:, :-1]
        self.regions = data[:, -1]

    def read_fibres(self):
        self.fibres = np.loadtxt(self.fibre_file, dtype=np.float64, skiprows=1, ndmin=2)

    def read(self, unit_conversion=1000):
        self.read_nodes(unit_conversion)
        self.read_elems()
        self.read_fibres()
        return self.nodes, self.elems, self.regions, self.fibres

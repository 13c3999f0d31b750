"""P1 mass/stiffness assembly classes (torchcor/core/assemble.py:99-248).

Both classes build the pattern and values on the GPU through ``P1System``
and return coalesced sparse COO tensors (K, M) in ``dtype`` like the
reference.  ``assemble_matrices`` takes the per-element conductivity tensor
(E,3,3) as in the reference.
"""
import torch

from .system import P1System


class Matrices3D:
    """P1 tetrahedra (assemble.py:148-248)."""

    def __init__(self, vertices, tetrahedrons, device, dtype):
        self.vertices = vertices
        self.tetrahedrons = tetrahedrons
        self.n_vertices = vertices.shape[0]
        self.device = torch.device(device)
        self.dtype = dtype

    def system(self):
        return P1System(self.vertices, self.tetrahedrons, self.device)

    def assemble_matrices(self, sigma):
        sys_ = self.system()
        M, K = sys_.assemble(sigma=sigma)
        K = sys_.csr(K.to(self.dtype)).to_sparse_coo().coalesce()
        M = sys_.csr(M.to(self.dtype)).to_sparse_coo().coalesce()
        return K, M


class Matrices3DSurface(Matrices3D):
    """P1 surface triangles embedded in 3-D (assemble.py:99-145)."""

    def __init__(self, vertices, triangles, device, dtype):
        super().__init__(vertices, triangles, device, dtype)
        self.triangles = triangles

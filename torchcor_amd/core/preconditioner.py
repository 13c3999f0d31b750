"""Jacobi preconditioner (torchcor/core/preconditioner.py:4-43).

Inside the time step the scaling by 1/diag(A) is fused into the PCG kernels;
this class keeps the reference's stand-alone API.
"""
import torch


class Preconditioner:
    def __init__(self):
        self.PC = torch.tensor([1.0])

    def create_Jocobi(self, A):
        A = A.to_sparse_coo().coalesce()
        idx, val = A.indices(), A.values()
        m = idx[0] == idx[1]
        d = torch.zeros(A.shape[0], device=val.device, dtype=val.dtype)
        d[idx[0][m]] = val[m]
        self.PC = 1.0 / d

    def apply(self, r):
        return self.PC * r

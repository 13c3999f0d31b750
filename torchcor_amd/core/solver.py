"""Jacobi-PCG with the reference's API (torchcor/core/solver.py:5-76).

``solve`` runs entirely in the HIP library (SELL-64 SpMV, fused vector
passes, float64 reductions, device-side stopping test).  The matrix is
stored in ``dtype`` (float64 by default, as in the reference).
"""
import torch

from .system import csr_system


class ConjugateGradient:
    def __init__(self, preconditioner, A, dtype=torch.float64):
        self.preconditioner = preconditioner
        self.dtype = dtype
        A = A.to_sparse_csr()
        n = A.shape[0]
        self.sys = csr_system(A.crow_indices(), A.col_indices(), n, A.device)
        vals = A.values().to(torch.float64)
        zero = torch.zeros_like(vals)
        # Av = 1*A + 0*0, Dinv = 1/diag(A)
        self.sys.set_operators(vals, zero, 1.0, 1.0, 0.0, 0.0, dtype)
        self.x = self.x_prev = self.x_next = None
        self.linear_guess = False

    def initialize(self, x, linear_guess=True):
        self.x = x.to(self.dtype).clone()
        self.x_prev = x.to(self.dtype).clone()
        self.x_next = torch.empty_like(self.x)
        self.linear_guess = linear_guess

    def solve(self, b, a_tol=1e-6, r_tol=1e-6, max_iter=100):
        b_dtype = b.dtype
        bb = b.to(self.dtype).contiguous()
        n_iter = self.sys.pcg_solve(bb, self.x, self.x_prev, self.x_next, a_tol, r_tol, max_iter,
                                    self.linear_guess)
        if self.linear_guess:
            self.x_prev, self.x, self.x_next = self.x, self.x_next, self.x_prev
        return self.x.to(b_dtype), n_iter

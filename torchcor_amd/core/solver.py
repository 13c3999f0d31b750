"""Jacobi-PCG with the reference's API (torchcor/core/solver.py:5-76).

``solve`` runs entirely in the HIP library (SELL-64 SpMV, fused vector
passes, float64 reductions, device-side stopping test).  As in the
reference, the matrix and the vectors are held in ``dtype`` (float64 by
default), the Jacobi scaling is the preconditioner's own (computed in A's
dtype, preconditioner.py:9-27) and x_prev keeps the dtype of the initial x
(solver.py:26): it is stored in ``dtype`` with its values rounded to that
dtype after every linear-guess update.
"""
import torch

from .system import csr_system


class ConjugateGradient:
    def __init__(self, preconditioner, A, dtype=torch.float64):
        self.preconditioner = preconditioner
        self.dtype = dtype
        A = A.to_sparse_csr()
        n = A.shape[0]
        self.n = n
        self.sys = csr_system(A.crow_indices(), A.col_indices(), n, A.device)
        vals = A.values().to(torch.float64)
        zero = torch.zeros_like(vals)
        # Av = 1*A + 0*0 in dtype; Dinv is replaced by the preconditioner's
        # scaling before every solve
        self.sys.set_operators(vals, zero, 1.0, 1.0, 0.0, 0.0, dtype)
        self.x = self.x_prev = self.x_next = None
        self.x_prev_dtype = None
        self.linear_guess = False

    def _load_preconditioner(self):
        """Dinv <- the preconditioner's PC, as Preconditioner.apply multiplies
        r by it (a one-element PC, the value before create_Jocobi,
        broadcasts)."""
        pc = torch.as_tensor(self.preconditioner.PC)
        D = self.sys.Dinv
        if pc.numel() == self.n:
            D.copy_(pc.reshape(-1).to(device=D.device, dtype=D.dtype))
        elif pc.numel() == 1:
            D.fill_(float(pc.reshape(-1)[0]))
        else:
            raise ValueError(f"preconditioner has {pc.numel()} entries for a system of {self.n}")

    def initialize(self, x, linear_guess=True):
        self.x = x.to(self.dtype).clone()
        self.x_prev = x.to(self.dtype).clone()
        self.x_prev_dtype = x.dtype
        self.x_next = torch.empty_like(self.x)
        self.linear_guess = linear_guess

    def solve(self, b, a_tol=1e-6, r_tol=1e-6, max_iter=100):
        b_dtype = b.dtype
        bb = b.to(self.dtype).contiguous()
        self._load_preconditioner()
        n_iter = self.sys.pcg_solve(bb, self.x, self.x_prev, self.x_next, a_tol, r_tol, max_iter,
                                    self.linear_guess)
        if self.linear_guess:
            self.x_prev, self.x, self.x_next = self.x, self.x_next, self.x_prev
            if self.x_prev_dtype != self.dtype:
                self.x_prev.copy_(self.x_prev.to(self.x_prev_dtype))
        return self.x.to(b_dtype), n_iter

"""Jacobi-preconditioned CG of the reference, restated in float64.

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
import scipy.sparse as sp


def jacobi(row_ptr, col, val):
    """PC = 1 / diag(A) (preconditioner.py:9-27)."""
    n = row_ptr.shape[0] - 1
    d = np.zeros(n)
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    m = rows == col
    d[rows[m]] = val[m]
    return 1.0 / d


class PCG:
    """State and solve of ``ConjugateGradient`` (solver.py:5-76).

    ``x`` persists between solves (warm start); with ``linear_guess`` each
    solve starts from ``2x - x_prev`` and then ``x_prev`` takes the old ``x``
    (solver.py:35-38).  ``x_prev_dtype`` is the dtype of the initial ``u``
    (solver.py:26): float64 for a float64 run, float32 for a float32 run.
    """

    def __init__(self, row_ptr, col, val, pc):
        n = row_ptr.shape[0] - 1
        self.A = sp.csr_matrix((np.asarray(val, dtype=np.float64), col, row_ptr), shape=(n, n))
        self.pc = np.asarray(pc, dtype=np.float64)

    def initialize(self, x, linear_guess=True, x_prev_dtype=np.float64):
        self.x = np.array(x, dtype=np.float64)
        self.x_prev = np.array(x, dtype=x_prev_dtype)
        self.linear_guess = linear_guess

    def solve(self, b, a_tol, r_tol, max_iter):
        """One PCG solve; returns (x, n_iter).

        Stops after an update when ||PC r|| < a_tol or ||PC r||/||PC b|| < r_tol
        (solver.py:68); ``n_iter`` counts only the iterations that did not
        stop (solver.py:74), so a solve that converges on its first
        iteration reports 0.  ``max_iter`` is reached silently.
        """
        b = np.asarray(b, dtype=np.float64)
        if self.linear_guess:
            old = self.x.copy()
            self.x = 2.0 * self.x - self.x_prev
            self.x_prev = old.astype(self.x_prev.dtype)
        A, pc = self.A, self.pc
        r = b - A @ self.x
        z = pc * r
        p = z.copy()
        b_norm = np.linalg.norm(pc * b)
        n_iter = 0
        for _ in range(max_iter):
            Ap = A @ p
            rz = r @ z
            alpha = rz / (p @ Ap)
            self.x = self.x + alpha * p
            r = r - alpha * Ap
            z = pc * r
            zn = np.linalg.norm(z)
            if zn < a_tol or zn / b_norm < r_tol:
                break
            beta = (r @ z) / rz
            p = beta * p + z
            n_iter += 1
        return self.x, n_iter

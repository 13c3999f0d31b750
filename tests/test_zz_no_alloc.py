"""tc_pcg_solve_static (no allocation, static device scratch) against
tc_pcg_solve (plan with a caller workspace): same kernels, so the solution
and the iteration count must agree bit for bit, on two calls in a row, the
second on a smaller system, and for both storage types.  The size limit is
checked on the host before any device work."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402

DEV = "cuda:0"


def _say(msg):
    print(msg, flush=True)


def _system(N, dtype):
    from torchcor_amd.core.system import P1System
    nodes, elems = meshes.kuhn_cube(N)
    E = len(elems)
    S = P1System(torch.from_numpy(nodes), torch.from_numpy(elems), DEV)
    rng = np.random.default_rng(N)
    f = rng.normal(size=(E, 3))
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    M, K = S.assemble(sigma_l=torch.full((E,), 0.3, dtype=torch.float64),
                      sigma_t=torch.full((E,), 0.1, dtype=torch.float64),
                      fibres=torch.from_numpy(f))
    S.set_operators(M, K, 140, 0.01, 0.5, 0.01, dtype)
    return S


def _inputs(S, dtype, seed):
    rng = np.random.default_rng(seed)
    n = S.n
    u0 = -80 + rng.normal(size=n)
    x = torch.tensor(u0, dtype=dtype, device=DEV)
    xp = torch.tensor(u0 + 0.1 * rng.normal(size=n), dtype=dtype, device=DEV)
    b = S.spmv("M", torch.tensor(0.01 * (u0 + 5 * np.sin(np.arange(n))), dtype=dtype,
                                  device=DEV))
    return b, x, xp


def _static_solve(S, b, x, xp, xn, tol, max_iter, lg):
    import torchcor_amd._lib as L
    it = C.c_int(-1)
    L.call("tc_pcg_solve_static", C.byref(S.sell), L.ptr(S.sA), L.ptr(S.Dinv), S.code, L.ptr(b),
           L.ptr(x), L.ptr(xp), L.ptr(xn), float(tol), float(tol), int(max_iter),
           1 if lg else 0, C.byref(it), L.stream_handle(S.device))
    return it.value


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_static_solve_equals_plan_solve(dtype):
    for N, seed, tol, max_iter, lg in ((21, 1, 1e-8, 100, True), (9, 2, 1e-6, 100, True),
                                      (13, 3, 1e-7, 100, False), (9, 4, 1e-30, 3, True)):
        _say(f"[gpu] static vs plan PCG, {N}^3, {dtype}, tol {tol}, max_iter {max_iter}, "
             f"linear guess {lg}")
        S = _system(N, dtype)
        b, x0, xp0 = _inputs(S, dtype, seed)
        xa, xpa, xna = x0.clone(), xp0.clone(), torch.full_like(x0, float("nan"))
        ita = S.pcg_solve(b, xa, xpa, xna, tol, tol, max_iter, lg)
        xb, xpb, xnb = x0.clone(), xp0.clone(), torch.full_like(x0, float("nan"))
        itb = _static_solve(S, b, xb, xpb, xnb, tol, max_iter, lg)
        assert ita == itb, (N, ita, itb)
        if lg:
            assert torch.equal(xna, xnb), N
            assert torch.equal(xa, xb) and torch.equal(xpa, xpb)
        else:
            assert torch.equal(xa, xb), N
        if max_iter == 100:
            assert 0 < itb < max_iter


@pytest.mark.gpu
def test_static_solve_rejects_oversize():
    import torchcor_amd._lib as L
    _say("[gpu] static PCG size limit")
    t = torch.zeros(64, dtype=torch.float64, device=DEV)
    big = L.Sell(L.TC_PCG_STATIC_MAX_ROWS + 1, (L.TC_PCG_STATIC_MAX_ROWS + 64) // 64,
                 L.ptr(t), L.ptr(t), None, None, None)
    it = C.c_int(-1)
    rc = L.lib.tc_pcg_solve_static(C.byref(big), L.ptr(t), L.ptr(t), L.TC_F64, L.ptr(t), L.ptr(t),
                                   L.ptr(t), L.ptr(t), 1e-5, 1e-5, 10, 1, C.byref(it),
                                   L.stream_handle(DEV))
    assert rc == -3
    assert b"static scratch maximum" in L.lib.tc_last_error()
    assert it.value == -1

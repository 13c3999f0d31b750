"""P1 finite-element pieces of the reference, restated in float64 numpy.

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np

# Reference-element shape-function gradients (assemble.py:156-161, :13-17).
_DN_TET = np.array([[-1.0, -1.0, -1.0],
                    [1.0, 0.0, 0.0],
                    [0.0, 1.0, 0.0],
                    [0.0, 0.0, 1.0]])
_DN_TRI = np.array([[-1.0, -1.0],
                    [1.0, 0.0],
                    [0.0, 1.0]])


def conductivity_fields(regions, calls):
    """Per-element (sigma_l, sigma_t) from a list of ``add_conductivity`` calls.

    Follows conductivity.py:18-47: with ``el``/``et`` both None the
    monodomain values are ``il``/``it``; otherwise they are the harmonic means
    ``il*el*(1/(il+el))``.  Elements in no listed region keep 0.
    """
    regions = np.asarray(regions)
    sl = np.zeros(regions.shape[0])
    st = np.zeros(regions.shape[0])
    for (ids, il, it, el, et) in calls:
        if el is None and et is None:
            l, t = il, it
        else:
            l = il * el * (1 / (il + el))
            t = it * et * (1 / (it + et))
        for rid in ids:
            m = regions == rid
            sl[m] = l
            st[m] = t
    return sl, st


def bidomain_fields(regions, calls):
    """Per-element (il, it, el, et) from ``add_conductivity`` calls, or None
    when any call leaves el/et out (conductivity.py:18-47 then drops the
    intra/extracellular fields)."""
    regions = np.asarray(regions)
    f = [np.zeros(regions.shape[0]) for _ in range(4)]
    for (ids, il, it, el, et) in calls:
        if el is None and et is None:
            return None
        for rid in ids:
            m = regions == rid
            for arr, v in zip(f, (il, it, el, et)):
                arr[m] = v
    return tuple(f)


def sigma_tensor(sl, st, fibres):
    """sigma = st*I + (sl-st) f f^T per element (conductivity.py:64-67)."""
    f = np.asarray(fibres, dtype=np.float64)
    out = st[:, None, None] * np.eye(3)[None]
    out = out + (sl - st)[:, None, None] * (f[:, :, None] @ f[:, None, :])
    return out


def tet_local(coords, sigma):
    """Local mass and stiffness of P1 tetrahedra (assemble.py:163-215).

    coords: (E,4,3); sigma: (E,3,3).  Returns (Me, Ke), each (E,4,4).
    """
    X = np.asarray(coords, dtype=np.float64)
    E = X.shape[0]
    aug = np.concatenate([np.ones((E, 4, 1)), X], axis=2)
    vol = np.abs(np.linalg.det(aug)) / 6.0
    Me = (vol / 20.0)[:, None, None] * (np.ones((4, 4)) + np.eye(4))[None]
    J = np.transpose(X, (0, 2, 1)) @ _DN_TET
    detJ = np.abs(np.linalg.det(J))
    G = _DN_TET[None] @ np.linalg.inv(J)
    Ke = (detJ / 6.0)[:, None, None] * (G @ (sigma @ np.transpose(G, (0, 2, 1))))
    return Me, Ke


def tri_local(coords, sigma):
    """Local mass and stiffness of P1 surface triangles in 3-D (assemble.py:99-145).

    Each triangle gets an in-plane orthonormal frame R (2x3); the conductivity
    is rotated to R sigma R^T and the 2-D P1 formulas are applied.
    """
    X = np.asarray(coords, dtype=np.float64)
    u1 = X[:, 1] - X[:, 0]
    u2 = X[:, 2] - X[:, 0]
    a = u1 / np.linalg.norm(u1, axis=1, keepdims=True)
    b = u2 - np.sum(u2 * a, axis=1, keepdims=True) * a
    b = b / np.linalg.norm(b, axis=1, keepdims=True)
    R = np.stack([a, b], axis=1)                       # (E,2,3)
    area = 0.5 * np.linalg.norm(np.cross(u1, u2), axis=1)
    Me = (area / 12.0)[:, None, None] * (np.ones((3, 3)) + np.eye(3))[None]
    s2 = R @ sigma @ np.transpose(R, (0, 2, 1))
    J = R @ np.transpose(X, (0, 2, 1)) @ _DN_TRI
    detJ = np.linalg.det(J)
    G = _DN_TRI[None] @ np.linalg.inv(J)
    Ke = (0.5 * detJ)[:, None, None] * (G @ (s2 @ np.transpose(G, (0, 2, 1))))
    return Me, Ke


def assemble_csr(elems, local, n, round_f32=True):
    """Global CSR from local matrices, as the reference's COO path does it.

    Contributions are listed in the order (i, j, element) (assemble.py:225-231),
    each value is rounded to float32 because ``sparse_coo_tensor`` infers
    float32 from a Python float list (assemble.py:233-243), and duplicates are
    summed in float64 in contribution order (``coalesce``, assemble.py:248).
    Explicit zeros are kept.  Returns (row_ptr int64, col int64, val float64).
    ``round_f32=False`` skips the float32 rounding (pure float64 assembly).
    """
    elems = np.asarray(elems, dtype=np.int64)
    E, k = elems.shape
    rows = np.concatenate([elems[:, i] for i in range(k) for j in range(k)])
    cols = np.concatenate([elems[:, j] for i in range(k) for j in range(k)])
    vals = np.concatenate([local[:, i, j] for i in range(k) for j in range(k)])
    if round_f32:
        vals = vals.astype(np.float32).astype(np.float64)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    ks = key[order]
    starts = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]]))
    val = np.add.reduceat(vals[order], starts) if len(starts) else np.zeros(0)
    ukey = ks[starts]
    r = ukey // n
    c = ukey % n
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row_ptr, r + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    return row_ptr, c, val


def spmv(row_ptr, col, val, x):
    """y = A x for CSR arrays, summing each row in column order."""
    import scipy.sparse as sp
    n = row_ptr.shape[0] - 1
    A = sp.csr_matrix((val, col, row_ptr), shape=(n, n))
    return A @ x

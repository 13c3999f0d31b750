"""Node reordering (torchcor/core/reorder.py:7-48).

``RCM`` keeps the reference's interface (``calculate_rcm_order``, ``reorder``,
``map``, ``inverse``); the order itself comes from the library's native
reverse Cuthill-McKee (``tc_rcm_order``, host code), not from scipy.
``rcm_order`` orders an assembled pattern by reverse Cuthill-McKee;
``level_order`` keeps the Cuthill-McKee level structure and orders each
level along a Z-order curve of the node coordinates (``tc_level_order``);
``slice_sort`` then groups rows of similar length inside windows of that
order (the simulator's default applies both).
"""
import ctypes as C
from itertools import combinations

import numpy as np
import torch

from .. import _lib as L


def rcm_order(row_ptr, col, n):
    """Reverse Cuthill-McKee order of a symmetric CSR pattern.

    Returns (order, bandwidth_before, bandwidth_after): ``order[i]`` is the
    old index of the node placed at position i (int64 numpy array)."""
    rp = np.ascontiguousarray(torch.as_tensor(row_ptr).cpu().numpy(), dtype=np.int32)
    cl = np.ascontiguousarray(torch.as_tensor(col).cpu().numpy(), dtype=np.int32)
    n = int(n)
    if rp.shape[0] != n + 1:
        raise ValueError("row_ptr must have n + 1 entries")
    order = np.empty(n, dtype=np.int32)
    b0, b1 = C.c_int64(0), C.c_int64(0)
    L.call("tc_rcm_order", n, rp.ctypes.data, cl.ctypes.data, order.ctypes.data, C.byref(b0),
           C.byref(b1))
    return order.astype(np.int64), int(b0.value), int(b1.value)


def level_order(row_ptr, col, n, xyz):
    """Cuthill-McKee levels, each along a Z-order curve of ``xyz`` (n, 3).

    Returns (order, bandwidth_before, bandwidth_after) as ``rcm_order``."""
    rp = np.ascontiguousarray(torch.as_tensor(row_ptr).cpu().numpy(), dtype=np.int32)
    cl = np.ascontiguousarray(torch.as_tensor(col).cpu().numpy(), dtype=np.int32)
    p = np.ascontiguousarray(torch.as_tensor(xyz).cpu().numpy(), dtype=np.float64)
    n = int(n)
    if rp.shape[0] != n + 1:
        raise ValueError("row_ptr must have n + 1 entries")
    if p.shape != (n, 3):
        raise ValueError("xyz must be (n, 3)")
    order = np.empty(n, dtype=np.int32)
    b0, b1 = C.c_int64(0), C.c_int64(0)
    L.call("tc_level_order", n, rp.ctypes.data, cl.ctypes.data, p.ctypes.data, order.ctypes.data,
           C.byref(b0), C.byref(b1))
    return order.astype(np.int64), int(b0.value), int(b1.value)


def slice_sort(order, row_ptr, window):
    """``order`` (position -> old node index) with the nodes of each window
    of ``window`` consecutive positions sorted by their pattern row length,
    longest first; equal lengths keep their order, and no node leaves its
    window.

    A SELL slice of 64 rows is as wide as its longest row, so ragged rows
    (an unstructured mesh: 5..25 entries a row) pad every slice to its
    longest; grouping rows of similar length cuts that padding (on the
    640k-node shell from 37 % of the nonzeros to 6 % with 640-row windows).
    With windows equal to the PCG blocks' row ranges each block keeps the
    same nodes, so its staged column set is unchanged.  ``row_ptr`` is the
    pattern in the old numbering."""
    order = np.asarray(order, dtype=np.int64)
    window = int(window)
    if window < 1:
        raise ValueError("window must be at least 1")
    deg = np.diff(np.asarray(torch.as_tensor(row_ptr).cpu().numpy(), dtype=np.int64))
    if deg.shape[0] != order.shape[0]:
        raise ValueError("row_ptr must have n + 1 entries")
    d = deg[order]
    top = int(d.max()) if d.size else 0
    key = (np.arange(order.shape[0], dtype=np.int64) // window) * (top + 1) + (top - d)
    return order[np.argsort(key, kind="stable")]


def pick_order(row_ptr, col, n, xyz, mode, window=0):
    """The internal node order that ``mode`` asks for, as (order,
    bandwidth_before, bandwidth_after), or None to keep the mesh's numbering.

    ``row_ptr``/``col``: the symmetric pattern (tensors; its device does the
    bandwidth arithmetic); ``xyz`` (n, 3) the node coordinates.  "levels":
    ``level_order``; "rcm": ``rcm_order``; "auto": the level order when the
    mesh's own numbering scatters neighbours (bandwidth above n/8) and the
    level order at least halves that bandwidth, else None; None: None.
    With ``window`` > 0 the order is then ``slice_sort``-ed in windows of
    that many rows, and bandwidth_after is the sorted order's."""
    if mode not in ("auto", "levels", "rcm", None):
        raise ValueError("mode must be 'auto', 'levels', 'rcm' or None")
    if mode is None:
        return None
    n = int(n)
    rp = torch.as_tensor(row_ptr)
    cl = torch.as_tensor(col).to(torch.int64)
    rows = torch.repeat_interleave(torch.arange(n, device=rp.device), rp.diff().to(torch.int64))
    if mode == "auto":
        bw0 = int((cl - rows).abs().max().item()) if cl.numel() else 0
        if 8 * bw0 <= n:
            return None
    if mode == "rcm":
        order, bw0, bw1 = rcm_order(row_ptr, col, n)
    else:
        order, bw0, bw1 = level_order(row_ptr, col, n, xyz)
    if mode == "auto" and 2 * bw1 > bw0:
        return None
    if window:
        order = slice_sort(order, row_ptr, window)
        bw1 = order_bandwidth(row_ptr, col, order)
    return order, bw0, bw1


def order_bandwidth(row_ptr, col, order):
    """Bandwidth (largest |new row - new column|) of the pattern in ``order``
    (position -> old node index); tensors, computed on their device."""
    rp = torch.as_tensor(row_ptr)
    dev = rp.device
    n = rp.shape[0] - 1
    cl = torch.as_tensor(col, device=dev).to(torch.int64)
    if cl.numel() == 0:
        return 0
    pos = torch.empty(n, dtype=torch.int64, device=dev)
    pos[torch.as_tensor(order, device=dev)] = torch.arange(n, device=dev)
    rows = torch.repeat_interleave(torch.arange(n, device=dev), rp.diff().to(torch.int64))
    return int((pos[cl] - pos[rows]).abs().max().item())


def window_columns(row_ptr, col, order, window):
    """The most distinct columns that the rows of one window of ``window``
    consecutive positions of ``order`` reference (the pattern in the old
    numbering; tensors, computed on their device).  A slice sort inside
    those windows leaves each window's node set, and so this count, as it
    is; a staged split over the same windows needs a set this large."""
    rp = torch.as_tensor(row_ptr)
    dev = rp.device
    n = rp.shape[0] - 1
    pos = torch.empty(n, dtype=torch.int64, device=dev)
    pos[torch.as_tensor(order, device=dev)] = torch.arange(n, device=dev)
    rows = torch.repeat_interleave(torch.arange(n, device=dev), rp.diff().to(torch.int64))
    w = pos[rows] // int(window)
    key = torch.unique(w * n + pos[torch.as_tensor(col, device=dev).to(torch.int64)])
    return int(torch.bincount(key // n).max().item()) if key.numel() else 0


def element_pattern(elems, n):
    """Symmetric CSR pattern (row_ptr, col) of the node pairs of the elements."""
    e = np.asarray(elems, dtype=np.int64)
    k = e.shape[1]
    pairs = [e[:, [a, b]] for a, b in combinations(range(k), 2)]
    pr = np.concatenate(pairs + [p[:, ::-1] for p in pairs] + [np.stack([np.arange(n)] * 2, 1)])
    key = np.unique(pr[:, 0] * n + pr[:, 1])
    rows, cols = key // n, key % n
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    return np.cumsum(row_ptr), cols


class RCM:
    def __init__(self, device, dtype):
        self.rcm_order = None
        self.rcm_map = None
        self.inverse_rcm_order = None
        self.device = device
        self.dtype = dtype

    def calculate_rcm_order(self, vertices, triangles):
        n = int(vertices.shape[0])
        rp, cl = element_pattern(torch.as_tensor(triangles).cpu().numpy(), n)
        order, _, _ = rcm_order(rp, cl, n)
        self.rcm_order = torch.as_tensor(order, dtype=torch.long, device=self.device)
        self.rcm_map = torch.empty_like(self.rcm_order)
        self.rcm_map[self.rcm_order] = torch.arange(n, dtype=torch.long, device=self.device)
        self.inverse_rcm_order = self.rcm_map

    def reorder(self, tensor_any):
        return tensor_any[self.rcm_order]

    def map(self, tensor_long):
        return self.rcm_map[tensor_long]

    def inverse(self, tensor_any):
        return tensor_any[self.inverse_rcm_order]

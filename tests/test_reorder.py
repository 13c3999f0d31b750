"""Node reordering (native reverse Cuthill-McKee, host code) and the
irregular shell mesh; CPU only.

The RCM order is checked by what fixes it independently of its code: a
randomly labelled path must come back with bandwidth 1, a randomly labelled
grid with bandwidth at most its short side + 1, disjoint components stay
contiguous, and on the shell mesh the bandwidth and profile are no worse
than scipy's reverse_cuthill_mckee (the routine the reference's RCM class
calls, torchcor/core/reorder.py:32).

The level order (Cuthill-McKee levels, each along a Z-order curve of the
coordinates) is checked against its definition computed independently:
breadth-first levels from its first node (scipy), positions sorted by level,
and inside a level by a Morton key formed here by plain bit loops; its
bandwidth is then at most two level widths.

The slice sort (rows grouped by length inside fixed windows of an order) is
checked against its definition by plain loops, with the properties that make
it safe for the staged split: no node leaves its window, and the bandwidth
moves by at most twice the window.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402
from torchcor_amd.core.reorder import (RCM, element_pattern, level_order,  # noqa: E402
                                       order_bandwidth, pick_order, rcm_order, slice_sort,
                                       window_columns)


def _csr(n, edges):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    a = sp.coo_matrix((np.ones(2 * len(e) + n), (np.r_[e[:, 0], e[:, 1], np.arange(n)],
                                                 np.r_[e[:, 1], e[:, 0], np.arange(n)])),
                      shape=(n, n)).tocsr()
    a.sum_duplicates()
    a.sort_indices()
    return a.indptr.astype(np.int64), a.indices.astype(np.int64)


def _band_profile(rp, col, order):
    n = rp.shape[0] - 1
    pos = np.empty(n, dtype=np.int64)
    pos[order] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(rp))
    d = np.abs(pos[rows] - pos[col])
    return int(d.max()), int(np.maximum.reduceat(d, rp[:-1]).sum())


def _relabel(n, edges, rng):
    lab = rng.permutation(n)
    return [(lab[a], lab[b]) for a, b in edges]


def test_rcm_path_graph_bandwidth_one():
    rng = np.random.default_rng(0)
    n = 200
    rp, col = _csr(n, _relabel(n, [(i, i + 1) for i in range(n - 1)], rng))
    order, bw0, bw1 = rcm_order(rp, col, n)
    assert sorted(order.tolist()) == list(range(n))
    assert bw0 > 50 and bw1 == 1
    assert _band_profile(rp, col, order)[0] == 1


def test_rcm_grid_bandwidth():
    rng = np.random.default_rng(1)
    nx, ny = 17, 40
    idx = np.arange(nx * ny).reshape(nx, ny)
    edges = [(idx[i, j], idx[i + 1, j]) for i in range(nx - 1) for j in range(ny)]
    edges += [(idx[i, j], idx[i, j + 1]) for i in range(nx) for j in range(ny - 1)]
    n = nx * ny
    rp, col = _csr(n, _relabel(n, edges, rng))
    order, bw0, bw1 = rcm_order(rp, col, n)
    assert sorted(order.tolist()) == list(range(n))
    assert bw1 <= nx + 1
    assert _band_profile(rp, col, order)[0] == bw1


def test_rcm_disconnected_components_contiguous():
    rng = np.random.default_rng(2)
    n1, n2 = 60, 45
    e1 = [(i, i + 1) for i in range(n1 - 1)]
    e2 = [(n1 + i, n1 + (i + 1) % n2) for i in range(n2)]        # a ring
    n = n1 + n2 + 3                                              # + 3 isolated nodes
    lab = rng.permutation(n)
    rp, col = _csr(n, [(lab[a], lab[b]) for a, b in e1 + e2])
    order, _, bw1 = rcm_order(rp, col, n)
    assert sorted(order.tolist()) == list(range(n))
    comp = np.zeros(n, dtype=np.int64)
    comp[lab[n1:n1 + n2]] = 1
    comp[lab[n1 + n2:]] = 2 + np.arange(3)
    seq = comp[order]
    runs = 1 + np.count_nonzero(seq[1:] != seq[:-1])
    assert runs == 5                                             # each component one run
    assert bw1 <= 2                                              # path 1, ring 2


def test_rcm_no_worse_than_scipy_on_shell():
    nodes, elems, *_ = meshes.ellipsoid_shell(15000, seed=3)
    n = nodes.shape[0]
    rp, col = element_pattern(elems, n)
    order, bw0, bw1 = rcm_order(rp, col, n)
    assert sorted(order.tolist()) == list(range(n))
    ours = _band_profile(rp, col, order)
    assert ours[0] == bw1 and bw0 == _band_profile(rp, col, np.arange(n))[0]
    a = sp.csr_matrix((np.ones(col.shape[0]), col, rp), shape=(n, n))
    ref = _band_profile(rp, col, reverse_cuthill_mckee(a, symmetric_mode=True))
    assert ours[0] <= ref[0] and ours[1] <= 1.02 * ref[1]
    assert bw1 < bw0 / 10


def test_rcm_rejects_bad_pattern():
    from torchcor_amd._lib import TorchcorError
    rp = np.array([0, 2, 3])
    col = np.array([0, 5, 1])
    with pytest.raises(TorchcorError):
        rcm_order(rp, col, 2)


def test_rcm_class_interface():
    """reorder/map/inverse as the reference's RCM class defines them."""
    nodes, elems, *_ = meshes.ellipsoid_shell(4000, seed=4)
    v = torch.from_numpy(nodes)
    t = torch.from_numpy(elems.astype(np.int64))
    r = RCM(device="cpu", dtype=torch.float64)
    r.calculate_rcm_order(v, t)
    rv, rt = r.reorder(v), r.map(t)
    assert torch.equal(rv[rt], v[t])                 # same elements, new labels
    assert torch.equal(r.inverse(rv), v)
    x = torch.randn(v.shape[0], dtype=torch.float64)
    assert torch.equal(r.reorder(r.inverse(x)), x)


def test_shell_mesh_properties():
    target = 20000
    nodes, elems, regions, fib, cond, stims = meshes.shell_workload(target, seed=0)
    n, E = nodes.shape[0], elems.shape[0]
    assert 0.8 * target < n < 1.4 * target
    assert np.array_equal(np.unique(elems), np.arange(n))           # no unused nodes
    X = nodes[elems]
    vol = np.abs(np.einsum("ij,ij->i", X[:, 1] - X[:, 0],
                           np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))) / 6
    assert vol.min() > 0
    # total volume close to the shell's (tets by centroid cover it)
    A = np.array(meshes.SHELL_OUTER_MM)
    shell = 4 / 3 * np.pi * (np.prod(A) - np.prod(A - meshes.SHELL_WALL_MM))
    assert abs(vol.sum() / shell - 1) < 0.08
    rp, col = element_pattern(elems, n)
    assert 12 <= np.diff(rp).mean() <= 20
    assert set(np.unique(regions)) == {0, 1, 2, 3}
    c = X.mean(axis=1)
    assert np.all((regions & 1) == (c[:, 0] >= 0)) and np.all((regions >> 1) == (c[:, 2] >= 0))
    # fibres: unit, tangent to the layer through the centroid, rotating with depth
    s = np.clip(meshes.shell_depth(c), 0, 1)
    nrm = c / (A - meshes.SHELL_WALL_MM * (1 - s)[:, None]) ** 2
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    assert np.abs(np.linalg.norm(fib, axis=1) - 1).max() < 1e-12
    assert np.abs((fib * nrm).sum(1)).max() < 1e-9
    ez = np.cross([0.0, 0.0, 1.0], nrm)
    ok = np.linalg.norm(ez, axis=1) > 0.3
    ez = ez[ok] / np.linalg.norm(ez[ok], axis=1, keepdims=True)
    ang = np.rad2deg(np.arccos(np.clip(np.abs((fib[ok] * ez).sum(1)), 0, 1)))
    want = np.abs(-60 + 120 * s[ok])
    assert np.abs(ang - want).max() < 1e-6
    # the node order is scattered (random permutation)
    assert _band_profile(rp, col, np.arange(n))[0] > n // 2
    # five patches, apical ones at 0 ms and basal ones at 5 ms
    assert [st for _, st in stims] == [0.0, 5.0, 0.0, 5.0, 0.0]
    for k, (ids, st) in enumerate(stims):
        assert len(ids) > 0
        z = nodes[ids, 2]
        assert np.all(z < -20) if st == 0.0 else np.all(z > 20)
    assert len(cond) == 4


def _morton_loop(xyz):
    """30-bit Z-order keys on a 1024^3 grid over the bounding box, one bit at a time."""
    lo = xyz.min(0)
    ext = (xyz.max(0) - lo).max()
    q = np.clip(((xyz - lo) * (1023.0 / ext)).astype(np.int64), 0, 1023)
    key = np.zeros(len(xyz), dtype=np.int64)
    for b in range(10):
        for a in range(3):
            key |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return key


def test_level_order_definition_on_mesh():
    from scipy.sparse.csgraph import breadth_first_order
    nodes, elems = meshes.jittered_delaunay(9, seed=2)
    n = nodes.shape[0]
    rp, col = element_pattern(elems, n)
    order, bw0, bw1 = level_order(rp, col, n, nodes)
    assert sorted(order.tolist()) == list(range(n))
    A = sp.csr_matrix((np.ones(col.shape[0]), col, rp), shape=(n, n))
    bfs, pred = breadth_first_order(A, order[0], directed=False, return_predecessors=True)
    assert len(bfs) == n                       # one component
    lev = np.zeros(n, dtype=np.int64)
    for v in bfs[1:]:
        lev[v] = lev[pred[v]] + 1
    key = _morton_loop(np.asarray(nodes, dtype=np.float64))
    expect = np.lexsort((np.arange(n), key, lev))
    assert np.array_equal(order, expect)
    width = np.bincount(lev).max()
    b, _ = _band_profile(rp, col, order)
    assert b == bw1 <= 2 * width
    assert bw0 == _band_profile(rp, col, np.arange(n))[0]


def test_level_order_path_and_components():
    rng = np.random.default_rng(4)
    n = 120
    lab = rng.permutation(n)
    edges = [(lab[i], lab[i + 1]) for i in range(59)] + [(lab[i], lab[i + 1]) for i in range(60, 119)]
    rp, col = _csr(n, edges)
    xyz = np.zeros((n, 3))
    xyz[lab, 0] = np.arange(n)
    order, _, bw1 = level_order(rp, col, n, xyz)
    assert bw1 == 1
    first = set(lab[:60].tolist())
    k = sum(1 for v in order[:60] if v in first)
    assert k in (0, 60)                         # each component contiguous


def test_level_order_rejects_bad_coordinates():
    rp, col = _csr(3, [(0, 1), (1, 2)])
    with pytest.raises(ValueError):
        level_order(rp, col, 3, np.zeros((3, 2)))
    bad = np.zeros((3, 3))
    bad[1, 2] = np.nan
    with pytest.raises(Exception):
        level_order(rp, col, 3, bad)


def _sell_slots(rp, order):
    """Entry slots of the SELL layout of ``order``: 64 x the longest row of
    every slice of 64 positions."""
    deg = np.diff(rp)[order]
    n = order.shape[0]
    d = np.zeros(-(-n // 64) * 64, dtype=np.int64)
    d[:n] = deg
    return int(d.reshape(-1, 64).max(axis=1).sum() * 64)


def test_slice_sort_definition_and_padding():
    nodes, elems = meshes.jittered_delaunay(16, seed=3)
    n = nodes.shape[0]
    rp, col = element_pattern(elems, n)
    lo, _, bw1 = level_order(rp, col, n, nodes)
    deg = np.diff(rp)
    for window in (1, 7, 64, 640, n + 5):
        got = slice_sort(lo, rp, window)
        expect = []
        for w0 in range(0, n, window):
            seg = lo[w0:w0 + window].tolist()
            ks = sorted(range(len(seg)), key=lambda k, seg=seg: (-deg[seg[k]], k))
            expect += [seg[k] for k in ks]
            assert sorted(got[w0:w0 + window].tolist()) == sorted(seg)   # stays in its window
        assert np.array_equal(got, np.asarray(expect))
        b, _ = _band_profile(rp, col, got)
        assert b <= bw1 + 2 * (window - 1)
        if window == 1:
            assert np.array_equal(got, lo)
    # grouping rows of similar length: fewer padded slots than the level order
    assert _sell_slots(rp, slice_sort(lo, rp, 640)) < 0.9 * _sell_slots(rp, lo)
    assert _sell_slots(rp, lo) > rp[-1]


def test_slice_sort_rejects_bad_arguments():
    rp, col = _csr(3, [(0, 1), (1, 2)])
    with pytest.raises(ValueError):
        slice_sort(np.arange(3), rp, 0)
    with pytest.raises(ValueError):
        slice_sort(np.arange(4), rp, 2)


def test_pick_order_modes():
    nodes, elems = meshes.jittered_delaunay(12, seed=5)           # randomly numbered
    n = nodes.shape[0]
    rp, col = element_pattern(elems, n)
    rpt, colt = torch.as_tensor(rp), torch.as_tensor(col)
    lo, b0, b1 = level_order(rp, col, n, nodes)
    ro, _, r1 = rcm_order(rp, col, n)
    assert pick_order(rpt, colt, n, nodes, None) is None
    o, a0, a1 = pick_order(rpt, colt, n, nodes, "auto")
    assert np.array_equal(o, lo) and (a0, a1) == (b0, b1) and 2 * a1 <= a0 and 8 * a0 > n
    o, _, a1 = pick_order(rpt, colt, n, nodes, "rcm")
    assert np.array_equal(o, ro) and a1 == r1
    o, _, a1 = pick_order(rpt, colt, n, nodes, "auto", window=64)
    assert np.array_equal(o, slice_sort(lo, rp, 64))
    assert a1 == _band_profile(rp, col, o)[0]
    # a mesh numbered along its lattice is already banded: "auto" keeps it,
    # "levels" still reorders
    nodes2, elems2 = meshes.jittered_delaunay(12, seed=5, permute=False)
    rp2, col2 = element_pattern(elems2, n)
    args = (torch.as_tensor(rp2), torch.as_tensor(col2), n, nodes2)
    assert pick_order(*args, "auto") is None
    assert pick_order(*args, "levels") is not None
    with pytest.raises(ValueError):
        pick_order(rpt, colt, n, nodes, "bfs")


def test_window_columns_and_bandwidth_by_brute_force():
    nodes, elems = meshes.jittered_delaunay(12, seed=8)
    n = nodes.shape[0]
    rp, col = element_pattern(elems, n)
    lo, _, b1 = level_order(rp, col, n, nodes)
    rpt, colt = torch.as_tensor(rp), torch.as_tensor(col)
    pos = np.empty(n, dtype=np.int64)
    pos[lo] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(rp))
    for window in (1, 64, 300, n):
        best = 0
        for w0 in range(0, n, window):
            sel = (pos[rows] >= w0) & (pos[rows] < w0 + window)
            best = max(best, len(set(pos[col[sel]].tolist())))
        assert window_columns(rpt, colt, lo, window) == best
        # the slice sort keeps every window's node set
        assert window_columns(rpt, colt, slice_sort(lo, rp, window), window) == best
    assert order_bandwidth(rpt, colt, lo) == b1 == _band_profile(rp, col, lo)[0]
    assert order_bandwidth(rpt, colt, np.arange(n)) == _band_profile(rp, col, np.arange(n))[0]

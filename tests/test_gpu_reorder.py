"""Internal node orders on the irregular shell (run with -m gpu).

The shell mesh comes with randomly permuted node numbers.  With
reorder="levels", "auto" (here the level order with rows grouped by length
inside the staged split's block windows) or "rcm" the simulator runs in
that order and maps every input and output through the permutation; with
reorder=None it runs in the mesh's own order.  Tolerances:
  * M, K, A returned in mesh numbering: bit-identical between the two orders
    (same pattern; each entry sums the same contributions in the same order).
  * float64 runs, reordered vs mesh order (only the summation order inside the
    products and reductions differs): TTP, whose right-hand side is
    continuous in V, max |dVm| < 1e-8 mV over the first millisecond and
    < 1e-6 mV over the run, the same per-step PCG iteration counts; MS2v, whose gate switches at
    V_gate, lets a roundoff-level difference flip that switch for one step
    at a node crossing it (0.16 mV after 10 ms on a coarser shell), so it is
    held to max |dVm| <= 0.5 mV.  Both:
    |dLAT|, |dRT| <= dt where both runs record one.
  * float32 in the "auto" order vs the float64 oracle in mesh order: max |dVm| <= 0.5 mV,
    |dLAT| <= dt.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import ionic as oion  # noqa: E402
from oracle import monodomain as omono  # noqa: E402
from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _say(msg):
    print(msg, flush=True)


# The shell scaled down by 4 (semi-axes 10/7.5/15 mm, wall 2.5 mm) at the
# full workload's element size (edges ~0.8 mm): small enough for the oracle,
# with the same resolution of the wavefront.
SMALL = dict(outer=(10.0, 7.5, 15.0), wall=2.5)


def _shell(n_target=9000, seed=11):
    nodes, elems, regions, fib = meshes.ellipsoid_shell(n_target, seed, **SMALL)
    stims = meshes.shell_stimuli(nodes, seed + 1, radius_mm=1.5, **SMALL)
    return nodes, elems, regions, fib, stims


def _gpu_model(name, dtype):
    from torchcor_amd.ionic import ModifiedMS2v, TenTusscherPanfilov
    if name == "ttp":
        return TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype)
    return ModifiedMS2v(0.01, device=DEV, dtype=dtype)


def _sim(mesh, model, dtype, T, reorder):
    from torchcor_amd.simulator import Monodomain
    nodes, elems, regions, fib, stims = mesh
    sim = Monodomain(_gpu_model(model, dtype), T=T, dt=0.01, device=DEV, dtype=dtype,
                     reorder=reorder)
    sim.load_arrays(nodes, elems, regions, fib)
    for c in meshes.SHELL_CONDUCTIVITIES:
        sim.add_conductivity(*c)
    for ids, start in stims:
        sim.add_stimulus(ids, start, 1.0, 100)
    return sim


@pytest.mark.parametrize("order", ["levels", "rcm"])
def test_shell_operators_identical_in_mesh_numbering(order):
    _say(f"[gpu] shell M/K/A {order} vs mesh order")
    mesh = _shell()
    a = _sim(mesh, "ms2v", torch.float64, 1.0, order)
    b = _sim(mesh, "ms2v", torch.float64, 1.0, None)
    a.assemble()
    b.assemble()
    assert a.node_order is not None and b.node_order is None
    assert a.system.bandwidth < 0.1 * mesh[0].shape[0] < b.system.bandwidth
    assert a.system.bandwidth == a.bandwidth_reordered
    for name in ("M", "K", "A"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x.crow_indices(), y.crow_indices())
        assert torch.equal(x.col_indices(), y.col_indices())
        assert torch.equal(x.values(), y.values()), name


def test_auto_reorder_choice():
    _say("[gpu] reorder auto")
    from torchcor_amd.simulator import Monodomain
    mesh = _shell()
    s = _sim(mesh, "ms2v", torch.float64, 1.0, "auto")
    s.assemble()
    assert s.node_order is not None and s.bandwidth_reordered * 2 < s.bandwidth_natural
    from torchcor_amd import autotune
    from torchcor_amd.core.reorder import element_pattern, level_order, slice_sort
    from torchcor_amd.core.system import P1System
    n = mesh[0].shape[0]
    rp, col = element_pattern(mesh[1], n)
    lo, _, _ = level_order(rp, col, n, np.asarray(mesh[0], dtype=np.float64))
    # "auto" applies the level order, then groups rows by length inside the
    # windows of the staged split, which the stage then uses
    w = P1System.sort_window(n)
    assert s.row_window == w
    assert np.array_equal(s.node_order.cpu().numpy(), slice_sort(lo, rp, w))
    assert s.system.staged and s.system.pcg_blocks == autotune.pcg_blocks(n, 8)
    assert s.system.bandwidth == s.bandwidth_reordered
    lv = _sim(mesh, "ms2v", torch.float64, 1.0, "levels")
    lv.assemble()
    assert lv.row_window == 0 and np.array_equal(lv.node_order.cpu().numpy(), lo)
    assert s.system.padded < 0.9 * lv.system.padded      # less SELL padding
    nodes, elems, regions, fib = meshes.cube_workload(21)
    c = Monodomain(_gpu_model("ms2v", torch.float64), T=1.0, dt=0.01, device=DEV,
                   dtype=torch.float64)
    c.load_arrays(nodes, elems, regions, fib)
    c.add_conductivity([0], 0.1, 0.1)
    c.add_conductivity([1], 0.1, 0.1)
    c.assemble()
    assert c.node_order is None            # the cube's own numbering is already banded


def test_auto_order_without_slice_sort_when_sets_do_not_fit(monkeypatch):
    """When the staged split's column sets would not fit a stage, "auto"
    keeps the plain level order (the per-entry gathers would be slowed by
    the row grouping)."""
    _say("[gpu] reorder auto, stage limit below the sets")
    import torchcor_amd._lib as L
    from torchcor_amd.core.reorder import element_pattern, level_order
    monkeypatch.setattr(L, "STAGE_MAX_COLS", 64)
    mesh = _shell()
    s = _sim(mesh, "ms2v", torch.float64, 1.0, "auto")
    s.assemble()
    n = mesh[0].shape[0]
    rp, col = element_pattern(mesh[1], n)
    lo, _, bw1 = level_order(rp, col, n, np.asarray(mesh[0], dtype=np.float64))
    assert s.row_window == 0
    assert np.array_equal(s.node_order.cpu().numpy(), lo)
    assert s.bandwidth_reordered == bw1 == s.system.bandwidth


@pytest.mark.parametrize("model,tol,order", [("ttp", 1e-6, "levels"), ("ms2v", 0.5, "levels"),
                                             ("ttp", 1e-6, "rcm"), ("ttp", 1e-6, "auto")])
def test_shell_reordered_matches_mesh_order_float64(model, tol, order):
    _say(f"[gpu] shell f64 {model} {order} vs mesh order")
    mesh = _shell()
    T = 10.0
    runs = []
    for reorder in (order, None):
        s = _sim(mesh, model, torch.float64, T, reorder)
        s.solve(1e-5, 1e-5, 100, snapshot_interval=1.0, verbose=False)
        runs.append(s)
    a, b = runs
    assert (a.frames - b.frames).abs().max().item() < tol
    if model == "ttp":
        assert (a.frames[:2] - b.frames[:2]).abs().max().item() < 1e-8
        assert a.iterations == b.iterations
    else:
        assert abs(sum(a.iterations) - sum(b.iterations)) <= 0.01 * sum(b.iterations)
    for x, y in ((a.activation_time, b.activation_time), (a.repolarization_time, b.repolarization_time)):
        both = (x > 0) & (y > 0)
        assert (x > 0).sum().item() == pytest.approx((y > 0).sum().item(), rel=0.01, abs=2)
        if both.any():
            assert (x[both] - y[both]).abs().max().item() <= 0.01 + 1e-9
    assert (a.activation_time > 0).float().mean().item() > 0.01


def test_shell_step_api_in_mesh_numbering():
    """step() takes and returns u in mesh numbering under the internal order."""
    _say("[gpu] shell step() levels")
    mesh = _shell(5000, seed=5)
    a = _sim(mesh, "ms2v", torch.float64, 0.5, "levels")
    a.solve(1e-5, 1e-5, 100, snapshot_interval=1.0, verbose=False)
    b = _sim(mesh, "ms2v", torch.float64, 0.5, "levels")
    b.assemble()
    assert b.node_order is not None
    u = b.ionic_model.initialize(b.n_nodes)
    t, its = 0.0, []
    for _ in range(b.nt):
        t += 0.01
        u, it, _, _ = b.step(u, t, 1e-5, 1e-5, 100)
        its.append(it)
    assert its == a.iterations
    assert (u - a.u).abs().max().item() < 1e-10


def test_shell_float32_reordered_vs_float64_oracle():
    _say("[gpu] shell f32 ttp auto order vs f64 oracle")
    mesh = _shell()
    nodes, elems, regions, fib, stims = mesh
    T = 12.0
    om = omono.Monodomain(oion.TenTusscherPanfilov("ENDO", 0.01), T=T, dt=0.01)
    om.load_arrays(nodes, elems, regions, fib)
    for c in meshes.SHELL_CONDUCTIVITIES:
        om.add_conductivity(*c)
    for ids, start in stims:
        om.add_stimulus(ids, start, 1.0, 100)
    ro = om.solve(1e-5, 1e-5, 100, snapshot_interval=1.0)
    g = _sim(mesh, "ttp", torch.float32, T, "auto")
    g.solve(1e-5, 1e-5, 100, snapshot_interval=1.0, verbose=False)
    assert g.node_order is not None
    Vg = g.frames.cpu().numpy()
    _say(f"max |dVm| per ms: {np.abs(Vg - ro['Vm']).max(axis=1).round(4).tolist()}")
    assert np.abs(Vg - ro["Vm"]).max() <= 0.5
    ATg = g.activation_time.cpu().numpy()
    assert np.array_equal(ATg > 0, ro["AT"] > 0)
    assert np.abs(ATg - ro["AT"]).max() <= 0.01 + 1e-6
    assert (ro["AT"] > 0).mean() > 0.01

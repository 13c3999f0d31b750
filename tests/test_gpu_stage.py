"""Staged gathers in the time-step runtime's spmv pass (operators without
shared slices: the irregular shell in its locality order).

Tolerance: bit-identical to the per-entry gathers of the same run (solution,
frames, per-step iteration counts, AT/RT), float32 and float64, TTP and
MS2v, over runs whose iteration counts vary (stalls).  The staged path is
also the one the shell's oracle comparisons in test_gpu_reorder.py run.
"""
import pytest

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SMALL = dict(outer=(10.0, 7.5, 15.0), wall=2.5)


def _say(msg):
    print(msg, flush=True)


def _run(mesh, model, dtype, T, reorder, stage, override=None, blocks=None):
    import torchcor_amd._lib as L
    from torchcor_amd.core.system import P1System
    from torchcor_amd.ionic import ModifiedMS2v, TenTusscherPanfilov
    from torchcor_amd.simulator import Monodomain
    nodes, elems, regions, fib, stims = mesh
    ion = (TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype) if model == "ttp"
           else ModifiedMS2v(0.01, device=DEV, dtype=dtype))
    old, old_ov = P1System.stage_gathers, P1System.stage_override
    P1System.stage_gathers = stage      # read when start() assembles the operators
    P1System.stage_override = override
    try:
        sim = Monodomain(ion, T=T, dt=0.01, device=DEV, dtype=dtype, reorder=reorder)
        sim.load_arrays(nodes, elems, regions, fib)
        for c in meshes.SHELL_CONDUCTIVITIES:
            sim.add_conductivity(*c)
        for ids, start in stims:
            sim.add_stimulus(ids, start, 1.0, 100)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=0.5)
    finally:
        P1System.stage_gathers, P1System.stage_override = old, old_ov
    if blocks is not None:              # the per-entry passes over the same block split
        L.call("tc_pcg_set_blocks", sim.system._pcg, blocks)
        sim.system.pcg_blocks = blocks
    sim.advance(sim.nt // 3)
    sim.advance(sim.nt - sim.nt // 3)
    sim.finish()
    return sim


@pytest.mark.parametrize("model,dtype", [("ttp", torch.float32), ("ttp", torch.float64),
                                         ("ms2v", torch.float32)])
def test_staged_spmv_identical_on_shell(model, dtype):
    nodes, elems, regions, fib = meshes.ellipsoid_shell(20000, 7, **SMALL)
    stims = meshes.shell_stimuli(nodes, 8, radius_mm=1.5, **SMALL)
    mesh = (nodes, elems, regions, fib, stims)
    runs = []
    for stage in (False, True):
        _say(f"[gpu] shell {model} {dtype} staged gathers {stage}")
        runs.append(_run(mesh, model, dtype, 8.0, "levels", stage))
    a, b = runs
    assert not a.system.staged and b.system.staged
    assert b.system.stage_max_cols > 0
    assert torch.equal(a.u, b.u)
    assert a.iterations == b.iterations
    assert torch.equal(a.frames, b.frames)
    assert torch.equal(a.activation_time, b.activation_time)
    assert torch.equal(a.repolarization_time, b.repolarization_time)
    assert len(set(a.iterations)) > 1
    assert (a.activation_time > 0).any()


@pytest.mark.parametrize("model,dtype,split", [("ttp", torch.float32, (2, 8)),
                                               ("ttp", torch.float64, (3, 16)),
                                               ("ms2v", torch.float32, (8, 16))])
def test_staged_units_identical_on_shell(model, dtype, split):
    """Blocks of more slices than one stage takes, staged in units of 8 or
    16 slices one after the other: bit-identical to the per-entry passes
    over the same block split."""
    nodes, elems, regions, fib = meshes.ellipsoid_shell(20000, 7, **SMALL)
    stims = meshes.shell_stimuli(nodes, 8, radius_mm=1.5, **SMALL)
    mesh = (nodes, elems, regions, fib, stims)
    G, chunk = split
    _say(f"[gpu] shell {model} {dtype} staged units: {G} blocks, {chunk} slices a unit")
    a = _run(mesh, model, dtype, 6.0, "levels", False, blocks=G)
    b = _run(mesh, model, dtype, 6.0, "levels", True, override=split)
    assert not a.system.staged and b.system.staged
    assert b.system.pcg_blocks == G and b.system.stage_chunk == chunk
    ns = -(-nodes.shape[0] // 64)
    assert -(-ns // G) > chunk                  # several units a block
    assert torch.equal(a.u, b.u)
    assert a.iterations == b.iterations
    assert torch.equal(a.frames, b.frames)
    assert torch.equal(a.activation_time, b.activation_time)
    assert torch.equal(a.repolarization_time, b.repolarization_time)
    assert len(set(a.iterations)) > 1


def test_stage_not_used_on_shared_slices_or_scattered_orders():
    """The cube's slices share one stencil (no stage); the shell in its own
    randomly permuted order references too many columns per block (no
    stage); both still run."""
    from torchcor_amd.ionic import ModifiedMS2v
    from torchcor_amd.simulator import Monodomain
    nodes, elems, regions, fib = meshes.cube_workload(21)
    c = Monodomain(ModifiedMS2v(0.01, device=DEV, dtype=torch.float32), T=0.2, dt=0.01,
                   device=DEV, dtype=torch.float32)
    c.load_arrays(nodes, elems, regions, fib)
    c.add_conductivity([0], 0.5272, 0.2076)
    c.add_conductivity([1], 0.9074, 0.3332)
    c.assemble()
    assert c.system.n_shared > 0 and not c.system.staged
    nodes, elems, regions, fib = meshes.ellipsoid_shell(9000, 11, **SMALL)
    stims = meshes.shell_stimuli(nodes, 12, radius_mm=1.5, **SMALL)
    s = _run((nodes, elems, regions, fib, stims), "ms2v", torch.float32, 0.5, None, True)
    assert not s.system.staged
    assert torch.isfinite(s.u).all()

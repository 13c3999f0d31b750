"""Ionic states of the time-step runtime in blocks of 64 nodes
(Monodomain.block_states, tc_sim_desc.state_block) against states in rows.

Tolerance: bit-identical (solution, per-step iteration counts, AT/RT, and
the model's states after the run, which the runtime copies back into its
(S, n) rows), for the three models, float32 and float64, on a node count
that is not a multiple of 64 (the last block is partial).
"""
import pytest

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _model(name, dtype):
    from torchcor_amd.ionic import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov
    if name == "ttp":
        return TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype)
    if name == "crn":
        return CourtemancheRamirezNattel(0.01, device=DEV, dtype=dtype)
    return ModifiedMS2v(0.01, device=DEV, dtype=dtype)


@pytest.mark.parametrize("name,dtype", [("ttp", torch.float32), ("crn", torch.float64),
                                        ("ms2v", torch.float32), ("ttp", torch.float64)])
def test_state_blocks_identical_to_rows(name, dtype):
    from torchcor_amd.simulator import Monodomain
    N = 13                                       # 2197 nodes: 34 full blocks and one of 21
    nodes, elems, regions, fib = meshes.cube_workload(N)
    print(f"[gpu] state blocks vs rows {name} {dtype}", flush=True)
    out = []
    old = Monodomain.block_states
    try:
        for blocks in (False, True):
            Monodomain.block_states = blocks
            ion = _model(name, dtype)
            sim = Monodomain(ion, T=6.0, dt=0.01, device=DEV, dtype=dtype)
            sim.load_arrays(nodes, elems, regions, fib)
            sim.add_conductivity([0], 0.5272, 0.2076)
            sim.add_conductivity([1], 0.9074, 0.3332)
            sim.add_stimulus(meshes.corner_nodes(N, 3), 0.0, 1.0, 100)
            sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
            assert (sim._state_blocks is not None) == blocks
            sim.advance(sim.nt // 2)
            sim.advance(sim.nt - sim.nt // 2)
            sim.finish()
            out.append((sim.u.clone(), list(sim.iterations), sim.activation_time.clone(),
                        sim.repolarization_time.clone(), ion.states.clone()))
    finally:
        Monodomain.block_states = old
    a, b = out
    assert torch.equal(a[0], b[0])
    assert a[1] == b[1]
    assert torch.equal(a[2], b[2]) and torch.equal(a[3], b[3])
    assert torch.equal(a[4], b[4])
    assert (a[2] > 0).any()

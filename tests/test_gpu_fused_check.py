"""The PCG stopping test run in each step's last update pass (the default)
against the same test in a one-block pass of its own: bit-identical
solutions, frames, iteration counts and AT/RT, with steps that need more
iterations than the pipeline launched (stalls) exercised."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _say(msg):
    print(msg, flush=True)


def _sim(N, model, dtype, T):
    from torchcor_amd.ionic import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov
    from torchcor_amd.simulator import Monodomain
    make = {"ttp": lambda: TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype),
            "crn": lambda: CourtemancheRamirezNattel(0.01, device=DEV, dtype=dtype),
            "ms2v": lambda: ModifiedMS2v(0.01, device=DEV, dtype=dtype)}[model]
    nodes, elems, regions, fib = meshes.cube_workload(N)
    sim = Monodomain(make(), T=T, dt=0.01, device=DEV, dtype=dtype)
    sim.load_arrays(nodes, elems, regions, fib)
    sim.add_conductivity([0], 0.5272, 0.2076)
    sim.add_conductivity([1], 0.9074, 0.3332)
    sim.add_stimulus(meshes.corner_nodes(N, 3), 0.0, 1.0, 100)
    sim.add_stimulus(meshes.corner_nodes(N, 2), 3.0, 0.5, 60, period=4.0, count=3)
    return sim


@pytest.mark.parametrize("model,dtype,N", [("ttp", torch.float32, 21), ("crn", torch.float64, 13),
                                           ("ms2v", torch.float32, 33)])
def test_fused_check_identical_to_check_pass(model, dtype, N):
    T = 15.0
    runs = []
    for separate in (True, False):
        _say(f"[gpu] fused check {model} {dtype} {N}^3, separate check pass {separate}")
        sim = _sim(N, model, dtype, T)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=0.5, depth=4, separate_check=separate)
        sim.advance(sim.nt // 2)
        sim.advance(sim.nt - sim.nt // 2)
        sim.finish()
        runs.append((sim.u.clone(), list(sim.iterations), sim.activation_time.clone(),
                     sim.repolarization_time.clone(), sim.frames.clone()))
    a, b = runs
    assert torch.equal(a[0], b[0])
    assert a[1] == b[1]
    assert torch.equal(a[2], b[2]) and torch.equal(a[3], b[3])
    assert torch.equal(a[4], b[4])
    assert len(set(a[1])) > 1          # iteration counts vary, so stalls happened


def test_fused_check_kernel_timing_counts():
    """Timing every step with the fused check: every step has one ionic and
    one rhs_init launch; the last update is timed as an update or, when the
    solve had stopped before it, as the check."""
    sim = _sim(21, "ttp", torch.float32, 5.0)
    sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
    sim.set_kernel_timing(True, 1)
    sim.advance(200)
    kt = sim.kernel_times()
    sim.finish()
    assert kt["ionic"][1] == 200 and kt["rhs_init"][1] == 200
    assert kt["pcg_spmv"][1] >= sum(sim.iterations) + 200 - 1 - 10
    assert kt["pcg_update"][1] + kt["pcg_check"][1] >= 200


@pytest.mark.parametrize("model,dtype,N", [("ttp", torch.float32, 21), ("crn", torch.float64, 13)])
def test_k_history_identical(model, dtype, N):
    """The number of iterations launched ahead of each status read (set by
    k_history) changes how often a step stalls or runs halted passes, never
    the results."""
    runs = []
    for H in (1, 2, 16):
        _say(f"[gpu] k_history {H} {model} {dtype} {N}^3")
        sim = _sim(N, model, dtype, 15.0)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=0.5, depth=4, k_history=H)
        sim.advance(sim.nt)
        sim.finish()
        runs.append((sim.u.clone(), list(sim.iterations), sim.activation_time.clone(),
                     sim.frames.clone()))
    for other in runs[1:]:
        assert torch.equal(runs[0][0], other[0])
        assert runs[0][1] == other[1]
        assert torch.equal(runs[0][2], other[2]) and torch.equal(runs[0][3], other[3])


def test_k_history_out_of_range():
    sim = _sim(9, "ms2v", torch.float32, 1.0)
    with pytest.raises(RuntimeError, match="k_history"):
        sim.start(1e-5, 1e-5, 100, k_history=257)

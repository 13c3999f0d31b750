"""Graph mode of the three engines (single run, shared-mesh batch, served
pool): each step's fixed launch sequence replayed as a captured HIP graph.

Tolerance: bit-identical to the eager launches of the same engine (solution,
per-step iteration counts, AT/RT, frames), float32 and float64, over runs
whose iteration counts vary (so steps stall and the pipeline re-issues the
later steps) and, for the served pool, across admissions and retirements
(which record the graphs anew).  The eager paths are pinned against the
float64 oracle by the other GPU tests.
"""
import pytest

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
COND = [([0], 0.5272, 0.2076), ([1], 0.9074, 0.3332)]


def _say(msg):
    print(msg, flush=True)


def _model(name, dtype):
    from torchcor_amd.ionic import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov
    if name == "ttp":
        return TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype)
    if name == "crn":
        return CourtemancheRamirezNattel(0.01, device=DEV, dtype=dtype)
    return ModifiedMS2v(0.01, device=DEV, dtype=dtype)


def _sim(N, model, dtype, T):
    from torchcor_amd.simulator import Monodomain
    nodes, elems, regions, fib = meshes.cube_workload(N)
    sim = Monodomain(_model(model, dtype), T=T, dt=0.01, device=DEV, dtype=dtype)
    sim.load_arrays(nodes, elems, regions, fib)
    for c in COND:
        sim.add_conductivity(*c)
    sim.add_stimulus(meshes.corner_nodes(N, 3), 0.0, 1.0, 100)
    sim.add_stimulus(meshes.corner_nodes(N, 2), 3.0, 0.5, 60, period=4.0, count=3)
    return sim


@pytest.mark.parametrize("model,dtype,N,separate", [("ttp", torch.float32, 21, False),
                                                    ("crn", torch.float64, 13, False),
                                                    ("ms2v", torch.float32, 25, True)])
def test_single_run_graph_identical(model, dtype, N, separate):
    runs = []
    for graph in (False, True):
        _say(f"[gpu] single-run graph={graph} {model} {dtype} {N}^3 separate check {separate}")
        sim = _sim(N, model, dtype, 12.0)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=0.5, separate_check=separate, graph=graph)
        sim.advance(sim.nt // 3)
        sim.advance(sim.nt - sim.nt // 3)
        cap, rep = sim.graph_stats()
        sim.finish()
        runs.append((sim.u.clone(), list(sim.iterations), sim.activation_time.clone(),
                     sim.repolarization_time.clone(), sim.frames.clone(), cap, rep))
    a, b = runs
    assert a[5] == a[6] == 0
    assert b[5] >= 2 and b[6] >= sim.nt          # every step replayed (re-issued ones again)
    assert torch.equal(a[0], b[0])
    assert a[1] == b[1]
    assert torch.equal(a[2], b[2]) and torch.equal(a[3], b[3])
    assert torch.equal(a[4], b[4])
    assert len(set(a[1])) > 1                    # iteration counts vary: stalls happened


def test_single_run_graph_with_timed_steps():
    """Steps that carry timing events launch eagerly; the others replay."""
    runs = []
    for graph in (False, True):
        sim = _sim(17, "ttp", torch.float32, 6.0)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9, graph=graph)
        sim.set_kernel_timing(True, 7)
        sim.advance(sim.nt)
        kt = sim.kernel_times()
        cap, rep = sim.graph_stats()
        sim.finish()
        runs.append((sim.u.clone(), list(sim.iterations), kt["ionic"][1], rep))
    assert torch.equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]
    assert runs[0][2] == runs[1][2] > 0
    assert runs[1][3] > 0


def _batch(dtype, T, graph):
    from torchcor_amd.batch import SharedMeshBatch
    N = 11
    nodes, elems, regions, fib = meshes.cube_workload(N)
    c3, far = meshes.corner_nodes(N, 3), meshes.face_nodes(N, 0, N - 1)
    specs = [("ms2v", 1.0, 0.0, [(c3, 0.0, 1.0, 100.0)]),
             ("ttp", 1.3, -15.0, [(c3, 0.0, 1.0, 100.0)]),
             ("ms2v", 0.7, 25.0, [(far, 2.0, 1.0, 80.0)]),
             ("ttp", 0.9, 30.0, [(far, 0.0, 1.0, 100.0), (c3, 3.0, 0.5, 60.0, 4.0, 2)])]
    b = SharedMeshBatch(T=T, dt=0.01, device=DEV, dtype=dtype)
    b.load_arrays(nodes, elems, regions, fib)
    for c in COND:
        b.add_conductivity(*c)
    for name, s, rot, st in specs * 16:               # 64 instances: the 64-lane kernels
        b.add_instance(_model(name, dtype), s, rot, st)
    b.start(1e-5, 1e-5, 100, graph=graph)
    b.advance(b.nt // 2)
    b.advance(b.nt - b.nt // 2)
    cap, rep = b.graph_stats()
    tot = b.finish()
    return b, tot, cap, rep


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_graph_identical(dtype):
    _say(f"[gpu] batch graph vs eager {dtype}")
    a, ta, ca, ra = _batch(dtype, 8.0, False)
    b, tb, cb, rb = _batch(dtype, 8.0, True)
    assert ca == ra == 0 and cb >= 2 and rb >= b.nt
    assert torch.equal(a.u, b.u)
    assert ta == tb and a.iterations == b.iterations
    assert torch.equal(a.activation_time, b.activation_time)
    assert torch.equal(a.repolarization_time, b.repolarization_time)
    assert len(set(a.iterations)) > 1


def _requests():
    from torchcor_amd.serve import Request
    out = []
    for k, (N, model, T, extra) in enumerate([(7, "ms2v", 6.0, False), (9, "ttp", 8.0, False),
                                               (11, "crn", 5.0, False), (8, "ms2v", 7.0, True),
                                               (10, "ttp", 4.0, True)]):
        nodes, elems, regions, fib = meshes.cube_workload(N)
        st = [(meshes.corner_nodes(N, 3), 0.0, 1.0, 100.0)]
        if extra:
            st.append((meshes.face_nodes(N, 0, N - 1), 2.0, 0.5, 80.0, 3.0, 2))
        out.append(Request(nodes, elems, regions, fib, COND, model, st, T, tag=k))
    return out


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_serve_graph_identical(dtype):
    from torchcor_amd.serve import ServeEngine
    res = []
    for graph in (False, True):
        _say(f"[gpu] served pool graph={graph} {dtype}")
        eng = ServeEngine(cap_nodes=2200, cap_entries=2200 * 16, max_sys=16, dtype=dtype,
                          device=DEV, graph=graph)   # two or three runs at a time
        out = eng.run_until_done([eng.prepare(r) for r in _requests()], max_chunk=37)
        res.append((out, list(eng.step_iters), eng.graph_stats()))
    (a, ia, ga), (b, ib, gb) = res
    assert ga == (0, 0)
    assert gb[0] >= 4 and gb[1] >= len(ib)      # recorded anew after pool changes
    assert ia == ib
    for ra, rb in zip(a, b):
        assert ra["steps"] == rb["steps"]
        assert torch.equal(ra["u"], rb["u"])
        assert torch.equal(ra["AT"], rb["AT"]) and torch.equal(ra["RT"], rb["RT"])

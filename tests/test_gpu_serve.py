"""Served stream (ServeEngine): continuous batching of runs on different
meshes.

Tolerances:
  * a run alone in the pool and the same run among others, admitted at other
    times and pool places (a pool that holds only some of the runs, so later
    ones wait and reuse freed room): bit-identical V, AT, RT and iteration
    sums (a run's arithmetic never mixes with another run's).
  * against the single-run path (Monodomain, same assembly and operators),
    float64: max |dVm| < 1e-9 mV, LAT equal, n_iter sums equal (only the
    order of the PCG partial sums differs).
  * a mesh with scattered node numbers (a small ellipsoidal shell) runs in
    the level order inside the pool: float64 against Monodomain in the same
    order as above; against the pool in the mesh's own order max |dVm| <
    1e-6 mV with the same iteration sum (summation order only).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
COND = [([0], 0.5272, 0.2076), ([1], 0.9074, 0.3332)]


def _say(msg):
    print(msg, flush=True)


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


def _engine(dtype, cap_nodes=20000, **kw):
    from torchcor_amd.serve import ServeEngine
    return ServeEngine(cap_nodes=cap_nodes, cap_entries=cap_nodes * 16, max_sys=16,
                       dtype=dtype, device=DEV, **kw)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_serve_alone_equals_in_mix(dtype):
    _say(f"[gpu] serve alone vs mix {dtype}")
    reqs = _requests()
    eng = _engine(dtype, cap_nodes=2200)      # holds two or three runs at a time
    mix = eng.run_until_done([eng.prepare(r) for r in reqs], max_chunk=37)
    for k, r in enumerate(reqs):
        alone_eng = _engine(dtype)
        alone = alone_eng.run_until_done([alone_eng.prepare(r)])[0]
        assert alone["steps"] == mix[k]["steps"] == int(r.T / 0.01)
        assert torch.equal(alone["u"], mix[k]["u"]), k
        assert torch.equal(alone["AT"], mix[k]["AT"]) and torch.equal(alone["RT"], mix[k]["RT"])
        assert alone["iterations"] == mix[k]["iterations"]


def test_serve_matches_single_run_path():
    from torchcor_amd.simulator import Monodomain
    from torchcor_amd.serve import MODEL_CLASSES
    _say("[gpu] serve vs Monodomain f64")
    reqs = _requests()
    eng = _engine(torch.float64)
    res = eng.run_until_done([eng.prepare(r) for r in reqs])
    for k, r in enumerate(reqs):
        model = MODEL_CLASSES[r.model](0.01, DEV, torch.float64)
        sim = Monodomain(model, T=r.T, dt=0.01, device=DEV, dtype=torch.float64)
        sim.load_arrays(r.nodes, r.elems, r.regions, r.fibres)
        for c in r.conductivities:
            sim.add_conductivity(*c)
        for s in r.stimuli:
            sim.add_stimulus(s[0], *s[1:])
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
        sim.advance(sim.nt)
        sim.finish()
        err = (sim.u - res[k]["u"]).abs().max().item()
        assert err < 1e-9, (k, err)
        assert torch.equal(sim.activation_time, res[k]["AT"]), k
        assert sum(sim.iterations) == res[k]["iterations"], k


def test_serve_int32_columns_identical():
    """The pool's int16 column offsets and int32 columns give bit-identical
    runs."""
    _say("[gpu] serve int16 vs int32 columns")
    reqs = _requests()[:3]
    out = []
    for flag in (True, False):
        eng = _engine(torch.float32, int16_columns=flag)
        out.append(eng.run_until_done([eng.prepare(r) for r in reqs]))
    for a, b in zip(*out):
        assert torch.equal(a["u"], b["u"]) and a["iterations"] == b["iterations"]


def test_serve_shared_slices_identical(monkeypatch):
    """Systems whose slices share one stencil bring their patterns into the
    pool, whose passes then read them (clamped to the system's own rows):
    bit-identical to the same systems on per-row columns."""
    from torchcor_amd.core.system import P1System
    _say("[gpu] serve shared vs per-row slices")
    reqs = _requests()
    out = []
    for flag in (True, False):
        monkeypatch.setattr(P1System, "share_slices", flag)
        eng = _engine(torch.float32)
        prs = [eng.prepare(r) for r in reqs]
        assert all((p.system.slice_tag is not None) == flag for p in prs)
        out.append(eng.run_until_done(prs))
    for a, b in zip(*out):
        assert torch.equal(a["u"], b["u"]) and a["iterations"] == b["iterations"]
        assert torch.equal(a["AT"], b["AT"]) and torch.equal(a["RT"], b["RT"])


def test_serve_loop_with_arrivals_matches_batch_admission():
    """The real-time loop (arrivals, assembly on a worker thread and its own
    stream, admission as room frees up) gives each run bit-identical
    results to admitting everything up front."""
    _say("[gpu] serve loop with arrivals")
    reqs = _requests()
    for k, r in enumerate(reqs):
        r.arrival = 0.02 * k
    eng = _engine(torch.float32, cap_nodes=2200)
    records, summary = eng.serve(reqs, max_chunk=40, admit_chunk=7)
    assert summary["requests"] == len(reqs)
    ref_eng = _engine(torch.float32)
    ref = ref_eng.run_until_done([ref_eng.prepare(r) for r in reqs])
    got = {rec["tag"]: rec["result"] for rec in records}
    for k in range(len(reqs)):
        assert got[k]["steps"] == ref[k]["steps"]
        assert torch.equal(got[k]["u"], ref[k]["u"]), k
        assert torch.equal(got[k]["AT"], ref[k]["AT"]) and torch.equal(got[k]["RT"], ref[k]["RT"])
        assert got[k]["iterations"] == ref[k]["iterations"]
    for rec in records:
        assert rec["admitted"] >= rec["arrival"] and rec["done"] > rec["admitted"]


@pytest.mark.parametrize("nreq", [24, 40])
def test_serve_many_systems_alone_equals_in_mix(nreq):
    """24 live systems: one check block takes them in two rounds; 40: the
    check pass runs over 3 blocks (16 systems a block) and the block that
    finishes last commits every system's clock and counts.  Runs alone give
    the same bits."""
    from torchcor_amd.serve import Request, ServeEngine
    _say(f"[gpu] serve {nreq} systems alone vs mix f32")
    rng = np.random.default_rng(11)
    reqs = []
    for k in range(nreq):
        N = int(rng.integers(6, 10))
        model = ("ms2v", "ttp", "crn")[k % 3]
        nodes, elems, regions, fib = meshes.cube_workload(N)
        st = [(meshes.corner_nodes(N, 2 + k % 2), 0.0, 1.0, 100.0)]
        reqs.append(Request(nodes, elems, regions, fib, COND, model, st, float(rng.uniform(1.0, 4.0)),
                            tag=k))
    eng = ServeEngine(cap_nodes=40000, cap_entries=640000, max_sys=64, dtype=torch.float32,
                      device=DEV)
    mix = eng.run_until_done([eng.prepare(r) for r in reqs], max_chunk=29)
    assert eng.stats()["steps"] >= 300
    for k in (0, 17, nreq - 1):
        alone_eng = _engine(torch.float32)
        alone = alone_eng.run_until_done([alone_eng.prepare(reqs[k])])[0]
        assert alone["steps"] == mix[k]["steps"] == int(reqs[k].T / 0.01)
        assert torch.equal(alone["u"], mix[k]["u"]), k
        assert torch.equal(alone["AT"], mix[k]["AT"]) and torch.equal(alone["RT"], mix[k]["RT"])
        assert alone["iterations"] == mix[k]["iterations"]
    for k, r in enumerate(reqs):
        assert mix[k]["steps"] == int(r.T / 0.01), k


SMALL_SHELL = dict(outer=(10.0, 7.5, 15.0), wall=2.5)


def _shell_request(T, model="ttp", tag=0):
    from torchcor_amd.serve import Request
    nodes, elems, regions, fib = meshes.ellipsoid_shell(3000, 11, **SMALL_SHELL)
    by_start = {}
    for ids, start in meshes.shell_stimuli(nodes, 12, radius_mm=1.5, **SMALL_SHELL):
        by_start.setdefault(start, []).append(np.asarray(ids))
    st = [(np.concatenate(v), start, 1.0, 100.0) for start, v in sorted(by_start.items())]
    return Request(nodes, elems, regions, fib, meshes.SHELL_CONDUCTIVITIES, model, st, T, tag=tag)


def test_serve_scattered_mesh_in_locality_order():
    from torchcor_amd.serve import MODEL_CLASSES
    from torchcor_amd.simulator import Monodomain
    _say("[gpu] serve scattered shell: level order vs Monodomain and mesh order (f64)")
    r = _shell_request(8.0)
    runs = {}
    for mode in ("auto", None):
        eng = _engine(torch.float64, cap_nodes=12000, reorder=mode)
        pr = eng.prepare(r)
        assert (pr.node_map is not None) == (mode == "auto")
        assert (pr.system.bandwidth < 0.1 * pr.n) == (mode == "auto")
        runs[mode] = eng.run_until_done([pr])[0]
        eng.close()
    a, b = runs["auto"], runs[None]
    assert a["iterations"] == b["iterations"]
    assert (a["u"] - b["u"]).abs().max().item() < 1e-6
    both = (a["AT"] > 0) & (b["AT"] > 0)
    assert both.sum().item() > 50
    assert (a["AT"][both] - b["AT"][both]).abs().max().item() <= 0.01 + 1e-9
    model = MODEL_CLASSES[r.model](0.01, DEV, torch.float64)
    sim = Monodomain(model, T=r.T, dt=0.01, device=DEV, dtype=torch.float64, reorder="levels")
    sim.load_arrays(r.nodes, r.elems, r.regions, r.fibres)
    for c in r.conductivities:
        sim.add_conductivity(*c)
    for st in r.stimuli:
        sim.add_stimulus(st[0], *st[1:])
    sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
    sim.advance(sim.nt)
    sim.finish()
    assert sim.node_order is not None
    assert (sim.u - a["u"]).abs().max().item() < 1e-9
    assert torch.equal(sim.activation_time, a["AT"])
    assert sum(sim.iterations) == a["iterations"]


def test_serve_scattered_mesh_alone_equals_in_mix():
    _say("[gpu] serve scattered shell alone vs among cube runs (f32)")
    reqs = _requests()[:3] + [_shell_request(6.0, "ms2v", tag=9)]
    eng = _engine(torch.float32, cap_nodes=12000)
    mix = eng.run_until_done([eng.prepare(q) for q in reqs], max_chunk=29)
    alone_eng = _engine(torch.float32, cap_nodes=12000)
    alone = alone_eng.run_until_done([alone_eng.prepare(reqs[-1])])[0]
    assert torch.equal(alone["u"], mix[-1]["u"])
    assert torch.equal(alone["AT"], mix[-1]["AT"]) and torch.equal(alone["RT"], mix[-1]["RT"])
    assert alone["iterations"] == mix[-1]["iterations"]

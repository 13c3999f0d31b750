"""Tuned block shapes keep parity: a run with every candidate shape against
the float64 oracle (single run: each PCG block count; served pool: each
slices-per-block; batch: each block cap), and the served pool's bitwise
alone-vs-mix property with a shape per system.

Tolerance: float64 against the float64 oracle, max |dVm| < 1e-6 mV, LAT
equal, n_iter sums equal (the shapes only regroup the PCG partial sums)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import ionic as oion  # noqa: E402
from oracle import monodomain as omono  # noqa: E402
from torchcor_amd import autotune, meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
COND = [([0], 0.5272, 0.2076), ([1], 0.9074, 0.3332)]


def _say(msg):
    print(msg, flush=True)


def _oracle(N, T, model="ttp", round_f32=True):
    nodes, elems, regions, fib = meshes.cube_workload(N)
    m = oion.TenTusscherPanfilov("ENDO", 0.01) if model == "ttp" else oion.ModifiedMS2v(0.01)
    om = omono.Monodomain(m, T=T, dt=0.01)
    om.round_f32 = round_f32
    om.load_arrays(nodes, elems, regions, fib)
    for c in COND:
        om.add_conductivity(*c)
    om.add_stimulus(meshes.corner_nodes(N, 4), 0.0, 1.0, 100.0)
    return om.solve(1e-5, 1e-5, 100)


@pytest.mark.parametrize("spb", autotune.CANDIDATES["pcg_slices_per_block"])
def test_single_run_block_counts_match_oracle(spb):
    import torchcor_amd._lib as L
    from torchcor_amd.ionic import TenTusscherPanfilov
    from torchcor_amd.simulator import Monodomain
    N, T = 17, 5.0
    _say(f"[gpu] single run, PCG slices per block {spb}")
    nodes, elems, regions, fib = meshes.cube_workload(N)
    sim = Monodomain(TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=torch.float64), T=T,
                     dt=0.01, device=DEV, dtype=torch.float64)
    sim.load_arrays(nodes, elems, regions, fib)
    for c in COND:
        sim.add_conductivity(*c)
    sim.add_stimulus(meshes.corner_nodes(N, 4), 0.0, 1.0, 100.0)
    sim.start(1e-5, 1e-5, 100, snapshot_interval=1e9)
    G = autotune.pcg_blocks(N ** 3, spb)
    L.call("tc_pcg_set_blocks", sim.system._pcg, G)
    sim.advance(sim.nt)
    sim.finish()
    ro = _oracle(N, T)
    assert np.abs(sim.u.cpu().numpy() - ro["u"]).max() < 1e-6
    assert np.array_equal(sim.activation_time.cpu().numpy(), ro["AT"])
    assert sum(sim.iterations) == int(ro["iters"].sum())


def test_served_shapes_match_oracle_and_alone_vs_mix():
    from torchcor_amd.serve import Request, ServeEngine
    _say("[gpu] served pool, each slices-per-block")

    def req(N, T, tag):
        nodes, elems, regions, fib = meshes.cube_workload(N)
        return Request(nodes, elems, regions, fib, COND, "ttp",
                       [(meshes.corner_nodes(N, 4), 0.0, 1.0, 100.0)], T, tag=tag)

    big = req(33, 4.0, "big")          # 562 slices: 71, 36 or 18 blocks
    small = [req(9, 3.0, 0), req(11, 5.0, 1)]
    ro = _oracle(33, 4.0)
    outs = {}
    for spb in autotune.CANDIDATES["serve_slices_per_block"]:
        runs = []
        for reqs in ([big], small[:1] + [big] + small[1:]):
            eng = ServeEngine(cap_nodes=60000, cap_entries=16 * 60000, max_sys=8,
                              dtype=torch.float64, device=DEV, slices_per_block=spb)
            res = eng.run_until_done([eng.prepare(r) for r in reqs], max_chunk=23)
            runs.append(next(o for r, o in zip(reqs, res) if r.tag == "big"))
            eng.close()
        alone, mixed = runs
        assert torch.equal(alone["u"], mixed["u"]) and alone["iterations"] == mixed["iterations"]
        assert np.abs(alone["u"].cpu().numpy() - ro["u"]).max() < 1e-6, spb
        assert np.array_equal(alone["AT"].cpu().numpy(), ro["AT"])
        assert alone["iterations"] == int(ro["iters"].sum())
        outs[spb] = alone


@pytest.mark.parametrize("max_blocks", autotune.CANDIDATES["batch_max_blocks"])
def test_batch_block_caps_match_oracle(max_blocks):
    from torchcor_amd.batch import SharedMeshBatch
    from torchcor_amd.ionic import ModifiedMS2v
    N, T = 21, 3.0
    _say(f"[gpu] batch, block cap {max_blocks}")
    nodes, elems, regions, fib = meshes.cube_workload(N)
    b = SharedMeshBatch(T=T, dt=0.01, device=DEV, dtype=torch.float64)
    b.load_arrays(nodes, elems, regions, fib)
    for c in COND:
        b.add_conductivity(*c)
    for k in range(40):                       # 64 lanes
        b.add_instance(ModifiedMS2v(0.01, device=DEV, dtype=torch.float64), 1.0, 0.0,
                       [(meshes.corner_nodes(N, 4), 0.0, 1.0, 100.0)])
    tot = b.solve(1e-5, 1e-5, 100, max_blocks=max_blocks)
    assert b.max_blocks == max_blocks
    ro = _oracle(N, T, "ms2v", round_f32=False)
    for k in (0, 39):
        assert np.abs(b.u[k].cpu().numpy() - ro["u"]).max() < 1e-6
        assert tot[k] == int(ro["iters"].sum())

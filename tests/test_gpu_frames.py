"""Vm snapshots streamed to host memory (run with -m gpu).

With frames="host" every snapshot goes through a small device ring
(4 slots) and is copied by DMA to pinned host memory while the run goes on;
with frames="device" the snapshots stay in device memory, as the reference
keeps its list of snapshots (torchcor/simulator/monodomain.py:162,186-187).
The two must give bit-identical frames, iteration counts and AT/RT: the
placement changes only where the copies land.  The runs take many more
frames than ring slots (each slot is rewritten about a hundred times) and
start with steps that need more PCG iterations than the runtime speculates,
so gated copies and their redo go through the ring too.
"""
import pytest

torch = pytest.importorskip("torch")

from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _say(msg):
    print(msg, flush=True)


def _cube_sim(model, dtype, T):
    from torchcor_amd.ionic import ModifiedMS2v, TenTusscherPanfilov
    from torchcor_amd.simulator import Monodomain
    ion = (TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype) if model == "ttp"
           else ModifiedMS2v(0.01, device=DEV, dtype=dtype))
    nodes, elems, regions, fib = meshes.cube_workload(21)
    sim = Monodomain(ion, T=T, dt=0.01, device=DEV, dtype=dtype)
    sim.load_arrays(nodes, elems, regions, fib)
    sim.add_conductivity([0], 0.1, 0.1)
    sim.add_conductivity([1], 0.1, 0.1)
    x0 = torch.nonzero(torch.as_tensor(nodes[:, 0]) == 0).flatten()
    sim.add_stimulus(x0.numpy(), 0.0, 1.0, 100)
    return sim


def _same_run(a, b):
    assert a.iterations == b.iterations
    assert a.frames.shape == b.frames.shape
    assert torch.equal(a.frames.cpu(), b.frames.cpu())
    assert torch.equal(a.activation_time.cpu(), b.activation_time.cpu())
    assert torch.equal(a.repolarization_time.cpu(), b.repolarization_time.cpu())


@pytest.mark.parametrize("model,dtype", [("ms2v", torch.float32), ("ttp", torch.float64)])
def test_host_ring_frames_match_device_frames(model, dtype, tmp_path):
    _say(f"[gpu] frames host ring vs device, cube 21, {model} {dtype}")
    runs = {}
    for where in ("device", "host"):
        s = _cube_sim(model, dtype, T=20.0)
        s.solve(1e-5, 1e-5, 100, snapshot_interval=0.05, verbose=False,
                result_path=tmp_path / where, frames=where)
        runs[where] = s
    a, b = runs["device"], runs["host"]
    assert a.frames.is_cuda and not b.frames.is_cuda and b.frames.is_pinned()
    assert b.frames.shape == (401, 21 ** 3)
    assert max(a.iterations) >= 1           # steps beyond the first speculation (K = 1)
    _same_run(a, b)
    assert (b.frames[-1] - b.frames[0]).abs().max().item() > 1.0
    saved = torch.load(tmp_path / "host" / "Vm.pt", weights_only=True)
    assert torch.equal(saved, b.frames)


def test_host_ring_with_rcm_order(tmp_path):
    _say("[gpu] frames host ring vs device, shell with rcm order")
    from torchcor_amd.ionic import ModifiedMS2v
    from torchcor_amd.simulator import Monodomain
    small = dict(outer=(10.0, 7.5, 15.0), wall=2.5)
    nodes, elems, regions, fib = meshes.ellipsoid_shell(5000, 5, **small)
    stims = meshes.shell_stimuli(nodes, 6, radius_mm=1.5, **small)
    runs = []
    for where in ("device", "host"):
        s = Monodomain(ModifiedMS2v(0.01, device=DEV, dtype=torch.float32), T=5.0, dt=0.01,
                       device=DEV, dtype=torch.float32, reorder="rcm")
        s.load_arrays(nodes, elems, regions, fib)
        for c in meshes.SHELL_CONDUCTIVITIES:
            s.add_conductivity(*c)
        for ids, start in stims:
            s.add_stimulus(ids, start, 1.0, 100)
        s.solve(1e-5, 1e-5, 100, snapshot_interval=0.1, verbose=False, frames=where)
        runs.append(s)
    assert runs[0].node_order is not None
    _same_run(*runs)


def test_frames_placement_checked():
    _say("[gpu] frames placement argument")
    s = _cube_sim("ms2v", torch.float32, T=1.0)
    with pytest.raises(ValueError):
        s.solve(1e-5, 1e-5, 100, verbose=False, frames="disk")

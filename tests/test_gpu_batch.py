"""Batched runs on a shared mesh (SharedMeshBatch) against single runs.

Tolerances:
  * float64, each instance against the float64 oracle run alone with the
    instance's rotated fibres and scaled conductivities (plain float64
    assembly): max |dVm| < 1e-6 mV at the end, LAT equal, per-instance sums
    of n_iter equal.
  * float64, an instance alone (B = 1) and in a mix of B instances: max
    |dVm| < 1e-8 mV, LAT/RT equal, n_iter sums equal.
  * float32 (float64 PCG reductions, float64 iterate and right-hand side)
    against the float64 oracle: max |dVm| <= 0.5 mV, |dLAT| <= dt.
  * a mesh with scattered node numbers (jittered Delaunay, randomly
    renumbered) runs in the level order: float64 against the oracle in the
    mesh's numbering as above, and against the batch in the mesh's order
    max |dVm| < 1e-8 mV with the same iteration sums.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import ionic as oion  # noqa: E402
from oracle import monodomain as omono  # noqa: E402
from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
COND = [([0], 0.5272, 0.2076), ([1], 0.9074, 0.3332)]


def _say(msg):
    print(msg, flush=True)


def _gpu_model(name, dtype):
    from torchcor_amd.ionic import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov
    if name == "ttp":
        return TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype)
    if name == "crn":
        return CourtemancheRamirezNattel(0.01, device=DEV, dtype=dtype)
    return ModifiedMS2v(0.01, device=DEV, dtype=dtype)


def _oracle_model(name):
    if name == "ttp":
        return oion.TenTusscherPanfilov("ENDO", 0.01)
    if name == "crn":
        return oion.CourtemancheRamirezNattel(0.01)
    return oion.ModifiedMS2v(0.01)


def _mesh(N, fibres3d):
    nodes, elems, regions, fib = meshes.cube_workload(N)
    if fibres3d:
        rng = np.random.default_rng(11)
        fib = rng.normal(size=fib.shape)
        fib /= np.linalg.norm(fib, axis=1, keepdims=True)
    return nodes, elems, regions, fib


def _specs(N):
    """(model, sigma scale, fibre rotation in degrees, stimuli)."""
    c3, far = meshes.corner_nodes(N, 3), meshes.face_nodes(N, 0, N - 1)
    return [("ms2v", 1.0, 0.0, [(c3, 0.0, 1.0, 100.0)]),
            ("ttp", 1.3, -15.0, [(c3, 0.0, 1.0, 100.0)]),
            ("ms2v", 0.7, 25.0, [(far, 2.0, 1.0, 80.0)]),
            ("ttp", 0.9, 30.0, [(far, 0.0, 1.0, 100.0), (c3, 3.0, 0.5, 60.0, 4.0, 2)])]


def _batch(mesh, specs, dtype, T):
    from torchcor_amd.batch import SharedMeshBatch
    b = SharedMeshBatch(T=T, dt=0.01, device=DEV, dtype=dtype)
    b.load_arrays(*mesh)
    for c in COND:
        b.add_conductivity(*c)
    for name, s, rot, st in specs:
        b.add_instance(_gpu_model(name, dtype), s, rot, st)
    tot = b.solve(1e-5, 1e-5, 100)
    return b, tot


def _oracle(mesh, spec, T):
    from torchcor_amd.batch import rotate_z
    name, s, rot, st = spec
    nodes, elems, regions, fib = mesh
    om = omono.Monodomain(_oracle_model(name), T=T, dt=0.01)
    om.round_f32 = False
    om.load_arrays(nodes, elems, regions, rotate_z(fib, rot))
    for ids, il, it in COND:
        om.add_conductivity(ids, il * s, it * s)
    for x in st:
        om.add_stimulus(*x)
    return om.solve(1e-5, 1e-5, 100)


@pytest.mark.parametrize("fibres3d", [False, True])
def test_batch_float64_matches_oracle_runs(fibres3d):
    N, T = 9, 8.0
    _say(f"[gpu] batch f64 vs oracle, {N}^3, 3-D fibres {fibres3d}")
    mesh = _mesh(N, fibres3d)
    specs = _specs(N)
    b, tot = _batch(mesh, specs, torch.float64, T)
    assert len(b.components) == (5 if fibres3d else 3)
    for k, spec in enumerate(specs):
        ro = _oracle(mesh, spec, T)
        err = np.abs(b.u[k].cpu().numpy() - ro["u"]).max()
        assert err < 1e-6, (k, err)
        assert np.array_equal(b.activation_time[k].cpu().numpy(), ro["AT"]), k
        assert tot[k] == int(ro["iters"].sum()), (k, tot[k], int(ro["iters"].sum()))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_instance_alone_equals_in_mix(dtype):
    N, T = 11, 10.0
    _say(f"[gpu] batch alone vs mix {dtype}")
    mesh = _mesh(N, False)
    specs = _specs(N) * 5          # 20 instances, lanes = 32
    mix, tot = _batch(mesh, specs, dtype, T)
    tol = 1e-8 if dtype == torch.float64 else 1e-3
    for k in (0, 1, 6, 19):
        alone, t1 = _batch(mesh, [specs[k]], dtype, T)
        err = (alone.u[0] - mix.u[k]).abs().max().item()
        assert err < tol, (k, err)
        assert torch.equal(alone.activation_time[0], mix.activation_time[k])
        if dtype == torch.float64:
            assert t1[0] == tot[k]


def test_batch_float32_vs_float64_oracle():
    N, T = 11, 15.0
    _say("[gpu] batch f32 vs f64 oracle")
    mesh = _mesh(N, False)
    specs = _specs(N)
    b, _ = _batch(mesh, specs, torch.float32, T)
    for k, spec in enumerate(specs):
        ro = _oracle(mesh, spec, T)
        err = np.abs(b.u[k].cpu().numpy().astype(np.float64) - ro["u"]).max()
        assert err <= 0.5, (k, err)
        at = b.activation_time[k].cpu().numpy().astype(np.float64)
        both = (at > 0) & (ro["AT"] > 0)
        assert np.array_equal(at > 0, ro["AT"] > 0)
        assert np.abs(at[both] - ro["AT"][both]).max() <= 0.01 + 1e-6


def _scattered_mesh():
    nodes, elems = meshes.jittered_delaunay(10, seed=21)         # randomly renumbered
    regions = (nodes[elems].mean(axis=1)[:, 0] > 5.0).astype(np.int32)
    fib = np.tile(np.array([[1.0, 0.0, 0.0]]), (elems.shape[0], 1))
    corner = np.nonzero((nodes < 2.5).all(axis=1))[0]
    far = np.nonzero(nodes[:, 0] > 8.5)[0]
    specs = [("ms2v", 1.0, 0.0, [(corner, 0.0, 1.0, 100.0)]),
             ("ttp", 1.2, 20.0, [(corner, 0.0, 1.0, 100.0)]),
             ("ms2v", 0.8, -30.0, [(far, 1.0, 1.0, 80.0)])]
    return (nodes, elems, regions, fib), specs


def test_batch_scattered_mesh_in_locality_order():
    from torchcor_amd.batch import SharedMeshBatch
    _say("[gpu] batch scattered mesh: level order vs oracle and mesh order (f64)")
    mesh, specs = _scattered_mesh()
    T = 6.0
    runs = {}
    for mode in ("auto", None):
        b = SharedMeshBatch(T=T, dt=0.01, device=DEV, dtype=torch.float64, reorder=mode)
        b.load_arrays(*mesh)
        for c in COND:
            b.add_conductivity(*c)
        for name, sc, rot, st in specs:
            b.add_instance(_gpu_model(name, torch.float64), sc, rot, st)
        tot = b.solve(1e-5, 1e-5, 100)
        assert (b.node_order is not None) == (mode == "auto")
        runs[mode] = (b, tot)
    (a, ta), (m, tm) = runs["auto"], runs[None]
    assert ta == tm
    assert (a.u - m.u).abs().max().item() < 1e-8
    assert torch.equal(a.activation_time, m.activation_time)
    for k, spec in enumerate(specs):
        ro = _oracle(mesh, spec, T)
        err = np.abs(a.u[k].cpu().numpy() - ro["u"]).max()
        assert err < 1e-6, (k, err)
        assert np.array_equal(a.activation_time[k].cpu().numpy(), ro["AT"]), k
        assert ta[k] == int(ro["iters"].sum()), k
    assert (a.activation_time > 0).sum().item() > 20

"""Batches of 64 instances: the 64-lane kernels (one wavefront per row for
all instances: whole-row pattern load plus lane broadcast), which batches of
33..64 instances select and which fewer instances never reach.

Tolerances:
  * float64, every instance against the float64 oracle run alone with its
    rotated fibres, scaled conductivities and stimuli (plain float64
    assembly): max |dVm| < 1e-6 mV, LAT equal, per-instance sums of n_iter
    equal.
  * an instance alone (B = 1, one lane) and at its lane in the 64-lane mix:
    float64 max |dVm| < 1e-8 mV with equal LAT and n_iter sums; float32
    <= 1e-3 mV.  Not bitwise: the per-instance partial sums of the PCG are
    grouped by the rows one wavefront holds, which depends on the lane count.
  * the 64 x 41^3 workload (32 MS2v + 32 TTP ENDO, float32): two runs
    bitwise equal; instance 0 alone vs in the mix <= 1e-3 mV; an MS2v and a
    TTP instance against the float64 oracle <= 0.5 mV over 200 steps.
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
TIMINGS = [(0.0, 1.0, 100.0), (0.5, 1.0, 80.0), (1.5, 0.5, 120.0), (0.0, 2.0, 60.0)]


def _say(msg):
    print(msg, flush=True)


def _gpu_model(name, dtype):
    from torchcor_amd.ionic import ModifiedMS2v, TenTusscherPanfilov
    if name == "ttp":
        return TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype)
    return ModifiedMS2v(0.01, device=DEV, dtype=dtype)


def _oracle_model(name):
    return oion.TenTusscherPanfilov("ENDO", 0.01) if name == "ttp" else oion.ModifiedMS2v(0.01)


def _mesh(N, fibres3d):
    nodes, elems, regions, fib = meshes.cube_workload(N)
    if fibres3d:
        rng = np.random.default_rng(12)
        fib = rng.normal(size=fib.shape)
        fib /= np.linalg.norm(fib, axis=1, keepdims=True)
    return nodes, elems, regions, fib


def _specs(N, B, seed, patch=20):
    """B instances: models in a shuffled order (so lanes and instance
    numbers differ), sigma scale U[0.5, 1.5], rotation U[-30, 30] degrees, a
    random surface patch with one of four timings; every eighth instance
    also has a periodic stimulus (period 4 ms, 2 beats) on the far face."""
    rng = np.random.default_rng(seed)
    nodes = meshes.cube_workload(N)[0]
    surf = meshes.cube_surface_nodes(N)
    far = meshes.face_nodes(N, 0, N - 1)
    names = np.array(["ms2v", "ttp"] * (B // 2) + ["ms2v"] * (B % 2))
    rng.shuffle(names)
    specs = []
    for k in range(B):
        start, dur, amp = TIMINGS[k % len(TIMINGS)]
        st = [(meshes.surface_patch(nodes, surf, patch, rng), start, dur, amp)]
        if k % 8 == 5:
            st.append((far, 2.0, 0.5, 60.0, 4.0, 2))
        specs.append((str(names[k]), rng.uniform(0.5, 1.5), rng.uniform(-30, 30), st))
    return specs


def _batch(mesh, specs, dtype, T, nsteps=None):
    from torchcor_amd.batch import SharedMeshBatch
    b = SharedMeshBatch(T=T, dt=0.01, device=DEV, dtype=dtype)
    b.load_arrays(*mesh)
    for c in COND:
        b.add_conductivity(*c)
    for name, s, rot, st in specs:
        b.add_instance(_gpu_model(name, dtype), s, rot, st)
    tot = b.solve(1e-5, 1e-5, 100, nsteps=nsteps)
    return b, tot


def _oracle(mesh, spec, T, nsteps=None):
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
    return om.solve(1e-5, 1e-5, 100, nsteps=nsteps)


def _lanes(b):
    return np.asarray(b.lane_of)


@pytest.mark.parametrize("fibres3d", [False, True])
def test_batch64_float64_matches_oracle_runs(fibres3d):
    N, T, B = 9, 5.0, 64
    _say(f"[gpu] batch64 f64 vs oracle, {N}^3, 3-D fibres {fibres3d}")
    mesh = _mesh(N, fibres3d)
    specs = _specs(N, B, seed=31 if fibres3d else 30)
    b, tot = _batch(mesh, specs, torch.float64, T)
    assert b.lanes == 64
    assert len(b.components) == (5 if fibres3d else 3)
    lane = _lanes(b)
    # every instance for in-plane fibres; for 3-D fibres the instances at
    # lanes 0, 31, 32, 63 and every fourth lane
    if fibres3d:
        check = sorted({int(np.flatnonzero(lane == L)[0]) for L in range(0, 64, 4)} |
                       {int(np.flatnonzero(lane == L)[0]) for L in (31, 32, 63)})
    else:
        check = range(B)
    for k in check:
        ro = _oracle(mesh, specs[k], T)
        err = np.abs(b.u[k].cpu().numpy() - ro["u"]).max()
        assert err < 1e-6, (k, int(lane[k]), err)
        assert np.array_equal(b.activation_time[k].cpu().numpy(), ro["AT"]), k
        assert tot[k] == int(ro["iters"].sum()), (k, tot[k], int(ro["iters"].sum()))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch64_instance_alone_equals_in_mix(dtype):
    N, T, B = 11, 8.0, 64
    _say(f"[gpu] batch64 alone vs mix {dtype}")
    mesh = _mesh(N, False)
    specs = _specs(N, B, seed=40)
    mix, tot = _batch(mesh, specs, dtype, T)
    assert mix.lanes == 64
    lane = _lanes(mix)
    tol = 1e-8 if dtype == torch.float64 else 1e-3
    for L in (0, 31, 32, 63):
        k = int(np.flatnonzero(lane == L)[0])
        alone, t1 = _batch(mesh, [specs[k]], dtype, T)
        assert alone.lanes == 1
        err = (alone.u[0] - mix.u[k]).abs().max().item()
        assert err < tol, (k, L, err)
        assert torch.equal(alone.activation_time[0], mix.activation_time[k]), (k, L)
        if dtype == torch.float64:
            assert t1[0] == tot[k], (k, L, t1[0], tot[k])


def _cube41_batch(dtype, steps):
    """The 64 x 41^3 workload of the batched benchmark: 32 MS2v then 32 TTP
    ENDO, sigma scale U[0.5, 1.5], rotation U[-30, 30] degrees, one seeded
    200-node surface patch each (100 uA/cm^2, 0..1 ms)."""
    N, B = 41, 64
    mesh = meshes.cube_workload(N)
    rng = np.random.default_rng(2024)
    surf = meshes.cube_surface_nodes(N)
    specs = []
    for k in range(B):
        patch = meshes.surface_patch(mesh[0], surf, 200, rng)
        specs.append(("ms2v" if k < B // 2 else "ttp", rng.uniform(0.5, 1.5),
                      rng.uniform(-30, 30), [(patch, 0.0, 1.0, 100.0)]))
    b, tot = _batch(mesh, specs, dtype, 100.0, nsteps=steps)
    return mesh, specs, b, tot


def test_batch64_cube41_workload_properties():
    steps = 200
    _say(f"[gpu] batch64 41^3 workload, {steps} steps, f32, twice")
    mesh, specs, b1, tot1 = _cube41_batch(torch.float32, steps)
    _, _, b2, tot2 = _cube41_batch(torch.float32, steps)
    assert b1.lanes == 64
    assert torch.equal(b1.u, b2.u) and tot1 == tot2
    assert torch.equal(b1.activation_time, b2.activation_time)
    assert bool(torch.isfinite(b1.u).all())
    assert bool((b1.activation_time > 0).any())
    _say("[gpu] batch64 41^3: instance 0 alone")
    alone, _ = _batch(mesh, [specs[0]], torch.float32, 100.0, nsteps=steps)
    assert (alone.u[0] - b1.u[0]).abs().max().item() <= 1e-3
    for k in (0, 32):
        _say(f"[gpu] batch64 41^3: instance {k} vs float64 oracle")
        ro = _oracle(mesh, specs[k], 100.0, nsteps=steps)
        err = np.abs(b1.u[k].cpu().numpy().astype(np.float64) - ro["u"]).max()
        assert err <= 0.5, (k, err)
        at = b1.activation_time[k].cpu().numpy().astype(np.float64)
        both = (at > 0) & (ro["AT"] > 0)
        # a node may cross 0 mV within float32 roundoff of the last step
        assert ((at > 0) != (ro["AT"] > 0)).sum() <= max(2, both.sum() // 1000), k
        if both.any():
            assert np.abs(at[both] - ro["AT"][both]).max() <= 0.01 + 1e-6

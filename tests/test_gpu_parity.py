"""HIP path against the CPU oracle (run with -m gpu on an MI355X).

Tolerances:
  * CSR pattern (row_ptr, col): bit-exact.
  * M, K values, float64 assembly without the float32 rounding: 1e-12 relative
    to the row scale; with the reference's float32 rounding of each element
    contribution: 2e-7 relative (one float32 ulp per contribution may differ
    because the element matrices come from different float64 routes).
  * PCG in float64: same iteration count, x within 1e-10 relative.
  * Ionic step in float64: 1e-9 relative to the oracle; float32: 2e-3.
  * Monodomain (float64 vs float64 oracle): max |dVm| < 1e-6 mV, LAT/RT equal,
    per-step n_iter equal.  float32 vs float64 oracle: max |dVm| <= 0.5 mV,
    |dLAT|, |dRT| <= dt.
"""
import functools

import numpy as np
import pytest
import scipy.sparse as sp

torch = pytest.importorskip("torch")

from oracle import fem as ofem  # noqa: E402
from oracle import ionic as oion  # noqa: E402
from oracle import monodomain as omono  # noqa: E402
from oracle import pcg as opcg  # noqa: E402
from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _say(msg):
    print(msg, flush=True)


def _sys(nodes, elems):
    from torchcor_amd.core.system import P1System
    return P1System(torch.from_numpy(nodes), torch.from_numpy(elems), DEV)


def _rand_fibres(E, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(E, 3))
    return f / np.linalg.norm(f, axis=1, keepdims=True)


MESHES = {
    "kuhn9": lambda: meshes.kuhn_cube(9),
    "delaunay": lambda: meshes.jittered_delaunay(9, seed=3),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_pattern_and_values(name):
    _say(f"[gpu] pattern/values {name}")
    nodes, elems = MESHES[name]()
    n, E = len(nodes), len(elems)
    sl = np.full(E, 0.5)
    st = np.full(E, 0.2)
    fib = _rand_fibres(E, 7)
    S = _sys(nodes, elems)
    Me, Ke = ofem.tet_local(nodes[elems], ofem.sigma_tensor(sl, st, fib))
    rp, col, kv = ofem.assemble_csr(elems, Ke, n, round_f32=False)
    _, _, mv = ofem.assemble_csr(elems, Me, n, round_f32=False)
    assert np.array_equal(S.row_ptr.cpu().numpy(), rp)
    assert np.array_equal(S.col.cpu().numpy(), col)
    M, K = S.assemble(sigma_l=torch.tensor(sl), sigma_t=torch.tensor(st),
                      fibres=torch.tensor(fib), round_f32=False)
    scale = np.abs(kv).max()
    assert np.abs(K.cpu().numpy() - kv).max() <= 1e-12 * scale
    assert np.abs(M.cpu().numpy() - mv).max() <= 1e-12 * np.abs(mv).max()
    # reference-compatible float32 rounding of contributions
    _, _, kv32 = ofem.assemble_csr(elems, Ke, n)
    M2, K2 = S.assemble(sigma_l=torch.tensor(sl), sigma_t=torch.tensor(st), fibres=torch.tensor(fib))
    assert np.abs(K2.cpu().numpy() - kv32).max() <= 2e-7 * scale
    # full-tensor conductivity input gives the same values
    sig = torch.tensor(ofem.sigma_tensor(sl, st, fib))
    M3, K3 = S.assemble(sigma=sig, round_f32=False)
    assert torch.allclose(K3, K, rtol=1e-13, atol=1e-15 * scale)
    # determinism
    M4, K4 = S.assemble(sigma_l=torch.tensor(sl), sigma_t=torch.tensor(st), fibres=torch.tensor(fib))
    assert torch.equal(K4, K2) and torch.equal(M4, M2)


def test_surface_triangles():
    _say("[gpu] surface triangles")
    nodes, tris = meshes.surface_of(*meshes.jittered_delaunay(7, seed=5, permute=False))
    E = len(tris)
    fib = _rand_fibres(E, 8)
    sl, st = np.full(E, 0.3), np.full(E, 0.06)
    S = _sys(nodes, tris)
    Me, Ke = ofem.tri_local(nodes[tris], ofem.sigma_tensor(sl, st, fib))
    rp, col, kv = ofem.assemble_csr(tris, Ke, len(nodes), round_f32=False)
    assert np.array_equal(S.row_ptr.cpu().numpy(), rp) and np.array_equal(S.col.cpu().numpy(), col)
    M, K = S.assemble(sigma_l=torch.tensor(sl), sigma_t=torch.tensor(st), fibres=torch.tensor(fib),
                      round_f32=False)
    assert np.abs(K.cpu().numpy() - kv).max() <= 1e-12 * np.abs(kv).max()


def _operators(N=11, dtype=torch.float64, dt=0.01):
    nodes, elems = meshes.kuhn_cube(N)
    E = len(elems)
    S = _sys(nodes, elems)
    sl, st, fib = np.full(E, 0.3), np.full(E, 0.1), _rand_fibres(E, 9)
    M, K = S.assemble(sigma_l=torch.tensor(sl), sigma_t=torch.tensor(st), fibres=torch.tensor(fib))
    S.set_operators(M, K, 140, 0.01, 0.5, dt, dtype)
    rp, col = S.row_ptr.cpu().numpy(), S.col.cpu().numpy()
    Mo, Ko = M.cpu().numpy(), K.cpu().numpy() / 140
    Ao = 0.01 * Mo + 0.5 * dt * Ko
    return S, rp, col, Mo, Ko, Ao


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_spmv_and_rhs(dtype):
    _say(f"[gpu] spmv/rhs {dtype}")
    S, rp, col, Mo, Ko, Ao = _operators(dtype=dtype)
    n = S.n
    rng = np.random.default_rng(1)
    x = rng.normal(size=n)
    w = rng.normal(size=n)
    xt = torch.tensor(x, dtype=dtype, device=DEV)
    wt = torch.tensor(w, dtype=dtype, device=DEV)
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    for which, vals in (("A", Ao), ("M", Mo), ("K", Ko)):
        ref = sp.csr_matrix((vals, col, rp), shape=(n, n)) @ x
        got = S.spmv(which, xt).cpu().numpy()
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max() + 1e-30
    b = torch.empty_like(xt)
    S.rhs(wt, xt, 0.005, b)
    ref = sp.csr_matrix((Mo, col, rp), shape=(n, n)) @ w - 0.005 * (sp.csr_matrix((Ko, col, rp), shape=(n, n)) @ x)
    assert np.abs(b.cpu().numpy() - ref).max() <= tol * np.abs(ref).max()


def test_pcg_float64_matches_oracle():
    _say("[gpu] pcg f64")
    S, rp, col, Mo, Ko, Ao = _operators(dtype=torch.float64)
    n = S.n
    pc = opcg.jacobi(rp, col, Ao)
    assert np.allclose(S.Dinv.cpu().numpy(), pc, rtol=1e-15)
    O = opcg.PCG(rp, col, Ao, pc)
    rng = np.random.default_rng(2)
    u0 = -80 + rng.normal(size=n)
    O.initialize(u0, linear_guess=True)
    x = torch.tensor(u0, device=DEV)
    xp = x.clone()
    xn = torch.empty_like(x)
    for k, (tol, mi) in enumerate([(1e-5, 100), (1e-8, 100), (1e-12, 200), (1e-30, 3), (1e-5, 0)]):
        b = 0.01 * (sp.csr_matrix((Mo, col, rp), shape=(n, n)) @ (u0 + 5 * np.sin(np.arange(n) * (k + 1))))
        xo, it_o = O.solve(b, tol, tol, mi)
        it = S.pcg_solve(torch.tensor(b, device=DEV), x, xp, xn, tol, tol, mi, True)
        xp, x, xn = x, xn, xp
        assert it == it_o, (k, it, it_o)
        assert np.abs(x.cpu().numpy() - xo).max() <= 1e-10 * np.abs(xo).max()


def test_pcg_float32_close_to_oracle():
    _say("[gpu] pcg f32")
    S, rp, col, Mo, Ko, Ao = _operators(dtype=torch.float32)
    n = S.n
    O = opcg.PCG(rp, col, Ao, opcg.jacobi(rp, col, Ao))
    u0 = np.full(n, -80.0)
    O.initialize(u0, linear_guess=False)
    b = 0.01 * (sp.csr_matrix((Mo, col, rp), shape=(n, n)) @ (u0 + 10 * np.cos(np.arange(n) * 0.01)))
    xo, _ = O.solve(b, 1e-6, 1e-6, 100)
    x = torch.tensor(u0, dtype=torch.float32, device=DEV)
    S.pcg_solve(torch.tensor(b, dtype=torch.float32, device=DEV), x, None, None, 1e-6, 1e-6, 100, False)
    assert np.abs(x.cpu().numpy() - xo).max() < 1e-3


def _gpu_model(name, dtype):
    from torchcor_amd.ionic import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov
    if name == "ttp":
        return TenTusscherPanfilov("ENDO", 0.01, device=DEV, dtype=dtype)
    if name == "ttp_epi":
        return TenTusscherPanfilov("EPI", 0.01, device=DEV, dtype=dtype)
    if name == "crn":
        return CourtemancheRamirezNattel(0.01, device=DEV, dtype=dtype)
    return ModifiedMS2v(0.01, device=DEV, dtype=dtype)


def _oracle_model(name):
    if name == "ttp":
        return oion.TenTusscherPanfilov("ENDO", 0.01)
    if name == "ttp_epi":
        return oion.TenTusscherPanfilov("EPI", 0.01)
    if name == "crn":
        return oion.CourtemancheRamirezNattel(0.01)
    return oion.ModifiedMS2v(0.01)


@pytest.mark.parametrize("name", ["ttp", "ttp_epi", "crn"])
def test_tables_float64(name):
    _say(f"[gpu] tables {name}")
    g = _gpu_model(name, torch.float64)
    g.initialize(4)
    o = _oracle_model(name)
    o.initialize(4)
    pairs = [("V_tab", "V_tab")]
    pairs += [("CaSS_tab", "CaSS_tab"), ("VEk_tab", "VEk_tab")] if name.startswith("ttp") else \
        [("Cai_tab", "Cai_tab"), ("fn_tab", "fn_tab")]
    for ga, oa in pairs:
        a = getattr(g, ga).cpu().numpy()
        b = getattr(o, oa)
        if ga == "V_tab" and name.startswith("ttp"):
            assert a.shape[1] == 28 and np.all(a[:, 27] == 0)   # padding column
            a = a[:, :27]
        assert a.shape == b.shape, ga
        fin = np.isfinite(b) & np.isfinite(a)
        assert (np.isfinite(a) == np.isfinite(b)).mean() > 0.999
        err = np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1e-300)
        assert np.quantile(err, 0.999) < 1e-12, ga
        # physiological rows must agree everywhere
        assert np.abs(a[fin] - b[fin]).max() <= 1e-9 * np.abs(b[fin]).max() + 1e-300, ga


def _random_states(name, n, rng):
    """Physiological-range states: the oracle's own trajectory sampled."""
    traj = _trajectory(name)
    pick = rng.integers(0, len(traj), n)
    Vs = np.array([traj[i][0] for i in pick]) + rng.normal(0, 0.5, n)
    st = {k: np.array([traj[i][1][k] for i in pick]) for k in traj[0][1]}
    return Vs, st


@functools.lru_cache(maxsize=None)
def _trajectory(name):
    o = _oracle_model(name)
    V = o.initialize(1)
    traj = []
    t = 0.0
    for s in range(30000):
        V = V + 0.01 * o.differentiate(V)
        t += 0.01
        if t <= 2.0:
            V = V + 0.01 * 40.0
        if s % 100 == 0:
            traj.append((V[0], {k: v[0] for k, v in o.s.items()} if hasattr(o, "s") else {"H": o.H[0]}))
    return traj


@pytest.mark.parametrize("name", ["ms2v", "ttp", "crn"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ionic_step(name, dtype):
    _say(f"[gpu] ionic step {name} {dtype}")
    rng = np.random.default_rng(11)
    n = 3000
    Vs, st = _random_states(name, n, rng)
    o = _oracle_model(name)
    o.initialize(n)
    g = _gpu_model(name, dtype)
    g.initialize(n)
    for k, v in st.items():
        if name == "ms2v":
            o.H = v.copy()
        else:
            o.s[k] = v.copy()
        setattr(g, k, torch.tensor(v, dtype=dtype))
    dVo = o.differentiate(Vs.copy())
    dVg = g.differentiate(torch.tensor(Vs, dtype=dtype, device=DEV)).cpu().numpy()
    scale = np.abs(dVo).max()
    tol = 1e-9 if dtype == torch.float64 else 2e-3
    assert np.abs(dVg - dVo).max() <= tol * scale
    for k in st:
        ov = o.H if name == "ms2v" else o.s[k]
        gv = getattr(g, k).cpu().numpy()
        assert np.abs(gv - ov).max() <= tol * max(np.abs(ov).max(), 1e-12) + (1e-12 if dtype == torch.float64 else 1e-6), k


def test_ionic_fused_outputs():
    """w, stimulus and LAT/RT epilogue of the fused kernel."""
    _say("[gpu] ionic fused epilogue")
    from torchcor_amd.ionic import ModifiedMS2v
    n = 1000
    g = ModifiedMS2v(0.01, device=DEV, dtype=torch.float64)
    g.initialize(n)
    V = torch.linspace(-90, 30, n, dtype=torch.float64, device=DEV)
    mask = torch.zeros(n, dtype=torch.int32, device=DEV)
    mask[:300] = 1
    mask[200:500] |= 4
    w = torch.empty_like(V)
    AT = torch.full_like(V, -1.0)
    RT = torch.full_like(V, -1.0)
    AT[:50] = 3.0
    H0 = g.H.clone()
    io = g._io(V, w=w, Cm=0.01, dt_rhs=0.01, stim_mask=mask, stim_bits=0b101,
               stim_amp=[100.0, 7.0, 50.0], AT=AT, RT=RT, lat_t=4.25)
    g.step(io)
    g.H = H0
    dV = g.differentiate(V)
    Ist = np.zeros(n)
    Ist[:300] += 100.0
    Ist[200:500] += 50.0
    wref = V.cpu().numpy() * 0.01 + 0.01 * (dV.cpu().numpy() / 100) + 0.01 * (Ist / 100)
    assert np.allclose(w.cpu().numpy(), wref, rtol=1e-14, atol=1e-14)
    Vn = V.cpu().numpy()
    ATr = np.where(Vn > 0, 4.25, -1.0)
    ATr[:50] = 3.0
    RTr = np.where((ATr > 0) & (Vn < -70), 4.25, -1.0)
    assert np.array_equal(AT.cpu().numpy(), ATr) and np.array_equal(RT.cpu().numpy(), RTr)


@pytest.mark.parametrize("name", ["ttp_endo", "ttp_epi", "crn"])
def test_single_cell_vs_reference_golden(name):
    _say(f"[gpu] single cell golden {name}")
    from pathlib import Path
    gold = np.load(Path(__file__).parent / "golden" / f"ionic_{name}_f64.npz")
    g = _gpu_model({"ttp_endo": "ttp", "ttp_epi": "ttp_epi", "crn": "crn"}[name], torch.float64)
    V = g.initialize(1)
    tr = [V.item()]
    t = 0.0
    nsteps = 6000
    for s in range(1, nsteps + 1):
        V = V + 0.01 * g.differentiate(V)
        t += 0.01
        if 0.0 <= t <= 2.0:
            V = V + 0.01 * 40.0
        if s % 10 == 0:
            tr.append(V.item())
    assert np.abs(np.array(tr) - gold["V"][: len(tr)]).max() < 1e-6


# ------------------------------------------------------------ monodomain
def _run_pair(N, model, dtype, T, stim_nodes, snapshot=1.0, cond=(0.1, 0.1), mesh=None):
    from torchcor_amd.simulator import Monodomain
    nodes, elems = mesh if mesh is not None else meshes.kuhn_cube(N)
    E = len(elems)
    regions = np.zeros(E, np.int32)
    fib = np.tile([1.0, 0.0, 0.0], (E, 1))
    om = omono.Monodomain(_oracle_model(model), T=T, dt=0.01)
    om.load_arrays(nodes, elems, regions, fib)
    om.add_conductivity([0], *cond)
    om.add_stimulus(stim_nodes, 0.0, 1.0, 100)
    ro = om.solve(1e-5, 1e-5, 100, snapshot_interval=snapshot)
    gm = Monodomain(_gpu_model(model, dtype), T=T, dt=0.01, device=DEV, dtype=dtype)
    gm.load_arrays(nodes, elems, regions, fib)
    gm.add_conductivity([0], *cond)
    gm.add_stimulus(stim_nodes, 0.0, 1.0, 100)
    it = gm.solve(1e-5, 1e-5, 100, snapshot_interval=snapshot, verbose=False)
    Vg = gm.frames.cpu().numpy()
    return ro, gm, Vg, it


def test_monodomain_float64_config_small():
    """21^3 MS2v float64 (the reference's CPU-runnable case), first 20 ms."""
    _say("[gpu] monodomain f64 21^3 MS2v")
    N = 21
    ro, gm, Vg, it = _run_pair(N, "ms2v", torch.float64, 20.0, meshes.face_nodes(N, 0, 0))
    assert np.abs(Vg - ro["Vm"]).max() < 1e-6
    assert np.array_equal(np.array(gm.iterations), ro["iters"])
    assert np.array_equal(gm.activation_time.cpu().numpy(), ro["AT"])
    assert np.array_equal(gm.repolarization_time.cpu().numpy(), ro["RT"])
    assert (ro["AT"] > 0).mean() > 0.9


@pytest.mark.parametrize("model,T", [("ttp", 30.0), ("crn", 30.0), ("ms2v", 30.0)])
def test_monodomain_float32_vs_float64_oracle(model, T):
    _say(f"[gpu] monodomain f32 {model}")
    N = 13
    ro, gm, Vg, it = _run_pair(N, model, torch.float32, T, meshes.face_nodes(N, 0, 0))
    assert np.abs(Vg - ro["Vm"]).max() <= 0.5
    ATg = gm.activation_time.cpu().numpy()
    RTg = gm.repolarization_time.cpu().numpy()
    assert np.array_equal(ATg > 0, ro["AT"] > 0)
    assert np.abs(ATg - ro["AT"]).max() <= 0.01 + 1e-6
    assert np.abs(RTg - ro["RT"]).max() <= 0.01 + 1e-6


def test_monodomain_unstructured_permuted():
    _say("[gpu] monodomain delaunay f64")
    mesh = meshes.jittered_delaunay(10, seed=4)
    nodes = mesh[0]
    stim = np.flatnonzero(nodes[:, 0] < 1.0)
    ro, gm, Vg, it = _run_pair(None, "ms2v", torch.float64, 10.0, stim, mesh=mesh)
    assert np.abs(Vg - ro["Vm"]).max() < 1e-6
    assert np.array_equal(np.array(gm.iterations), ro["iters"])


@pytest.mark.parametrize("model,dtype", [("ms2v", torch.float64), ("ttp", torch.float32)])
def test_monodomain_surface_triangles(model, dtype):
    """Monodomain on a triangle surface (the reference's Matrices3DSurface
    branch, monodomain.py:86-89): the closed surface of the 21^3 cube,
    anisotropic conductivity with fibres along x (normal to two of the six
    faces), stimulus on the x = 0 face."""
    _say(f"[gpu] monodomain surface triangles {model} {dtype}")
    nodes, tris = meshes.surface_of(*meshes.kuhn_cube(21))
    stim = np.flatnonzero(nodes[:, 0] == 0.0)
    ro, gm, Vg, it = _run_pair(None, model, dtype, 15.0, stim, cond=(0.3, 0.1),
                               mesh=(nodes, tris))
    assert gm.system.k == 3
    if dtype == torch.float64:
        assert np.abs(Vg - ro["Vm"]).max() < 1e-6
        assert np.array_equal(np.array(gm.iterations), ro["iters"])
        assert np.array_equal(gm.activation_time.cpu().numpy(), ro["AT"])
    else:
        assert np.abs(Vg - ro["Vm"]).max() <= 0.5
        ATg = gm.activation_time.cpu().numpy()
        assert np.array_equal(ATg > 0, ro["AT"] > 0)
        assert np.abs(ATg - ro["AT"]).max() <= 0.01 + 1e-6
    assert (ro["AT"] > 0).mean() > 0.5


def test_monodomain_files_api(tmp_path):
    """load_mesh/add_stimulus from files, outputs ATs.pt/RTs.pt/Vm.pt."""
    _say("[gpu] monodomain files api")
    import torchcor as tc
    from torchcor.ionic import TenTusscherPanfilov
    from torchcor.simulator import Monodomain
    N = 9
    nodes, elems, regions, fib = meshes.cube_workload(N)
    mesh_dir = meshes.write_mesh(tmp_path / "mesh", nodes, elems, regions, fib)
    vtx = meshes.write_vtx(tmp_path / "mesh" / "stim.vtx", meshes.corner_nodes(N, 3))
    ion = TenTusscherPanfilov(cell_type="ENDO", dt=0.01, dtype=tc.float32)
    sim = Monodomain(ion, T=5.0, dt=0.01, dtype=tc.float32)
    sim.load_mesh(path=mesh_dir)
    sim.add_conductivity([0], il=0.5272, it=0.2076)
    sim.add_conductivity([1], il=0.9074, it=0.3332)
    sim.add_stimulus(vtx, start=0.0, duration=1.0, intensity=100)
    n_iter = sim.solve(a_tol=1e-5, r_tol=1e-5, max_iter=100, calculate_AT_RT=True,
                       linear_guess=True, snapshot_interval=1, verbose=True,
                       result_path=tmp_path / "out")
    assert isinstance(n_iter, int)
    AT = torch.load(tmp_path / "out" / "ATs.pt", weights_only=True)
    Vm = torch.load(tmp_path / "out" / "Vm.pt", weights_only=True)
    assert AT.shape == (N ** 3,) and Vm.shape == (6, N ** 3)
    assert (AT[torch.from_numpy(meshes.corner_nodes(N, 3))] > 0).all()


def test_large_cube_properties():
    """101^3 cube: pattern size, mass = volume, K 1 = 0, a few TTP steps, determinism."""
    _say("[gpu] large cube 101^3")
    from torchcor_amd.simulator import Monodomain
    from torchcor_amd.ionic import TenTusscherPanfilov
    N = 101
    nodes, elems, regions, fib = meshes.cube_workload(N)
    outs = []
    for rep in range(2):
        sim = Monodomain(TenTusscherPanfilov("ENDO", 0.01, device=DEV), T=100, dt=0.01, device=DEV)
        sim.load_arrays(nodes, elems, regions, fib)
        sim.add_conductivity([0], 0.5272, 0.2076)
        sim.add_conductivity([1], 0.9074, 0.3332)
        sim.add_stimulus(meshes.corner_nodes(N, 10), 0.0, 1.0, 100)
        sim.solve(1e-5, 1e-5, 100, nsteps=200, verbose=False, snapshot_interval=100)
        outs.append(sim.u.clone())
        if rep == 0:
            S = sim.system
            assert S.nnz == 15210901
            ones = torch.ones(S.n, dtype=torch.float32, device=DEV)
            assert abs(S.spmv("M", ones).double().sum().item() - 1000.0) < 1e-2
            assert S.spmv("K", ones).abs().max().item() < 1e-4
            assert (sim.activation_time > 0).sum().item() >= 1000
    assert torch.equal(outs[0], outs[1])


def _cube_sim(N, model, dtype, T=100.0):
    from torchcor_amd.simulator import Monodomain
    nodes, elems, regions, fib = meshes.cube_workload(N)
    sim = Monodomain(_gpu_model(model, dtype), T=T, dt=0.01, device=DEV, dtype=dtype)
    sim.load_arrays(nodes, elems, regions, fib)
    sim.add_conductivity([0], 0.5272, 0.2076)
    sim.add_conductivity([1], 0.9074, 0.3332)
    sim.add_stimulus(meshes.corner_nodes(N, 3), 0.0, 1.0, 100)
    sim.add_stimulus(meshes.corner_nodes(N, 2), 3.0, 0.5, 60, period=4.0, count=3)
    return sim


@pytest.mark.parametrize("model", ["ttp", "crn"])
def test_runtime_pipeline_depths_identical(model):
    """The native pipelined loop gives bit-identical results at any depth
    (speculative iterations with stalls are exact)."""
    _say(f"[gpu] runtime depths {model}")
    N, steps = 11, 600
    runs = []
    for depth in (1, 6):
        sim = _cube_sim(N, model, torch.float32)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1.0, depth=depth)
        sim.advance(steps)
        sim.finish()
        runs.append((sim.u.clone(), list(sim.iterations), sim.activation_time.clone(),
                     sim.frames.clone()))
    assert torch.equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]
    assert torch.equal(runs[0][2], runs[1][2]) and torch.equal(runs[0][3], runs[1][3])
    assert len(set(runs[0][1])) > 1        # iteration counts vary, so stalls were exercised


@pytest.mark.parametrize("model", ["ttp", "ms2v"])
def test_runtime_matches_python_step_loop(model):
    """The native loop (fused RHS + PCG start, A x0 from M and K) against the
    per-step Python path (separate RHS and PCG start with A values), float64:
    same stimulus timing, LAT/RT, rotation and iteration counts."""
    _say(f"[gpu] runtime vs step() {model}")
    N, steps = 11, 500
    sim = _cube_sim(N, model, torch.float64)
    sim.start(1e-5, 1e-5, 100, snapshot_interval=1.0, depth=4)
    sim.advance(steps)
    sim.finish()
    ref_u, ref_it, ref_at = sim.u.clone(), list(sim.iterations), sim.activation_time.clone()
    sim = _cube_sim(N, model, torch.float64)
    sim.assemble()
    u = sim.ionic_model.initialize(sim.n_nodes)
    sim._init_solver(u, True)
    AT = torch.full_like(u, -1.0)
    RT = torch.full_like(u, -1.0)
    t, prev, its = 0, None, []
    for n in range(steps):
        t += 0.01
        lat = (AT, RT, prev) if prev is not None else None
        u, it, _, _ = sim.step(sim._x, t, 1e-5, 1e-5, 100, False, lat)
        prev = t
        its.append(it)
    assert its == ref_it
    assert (u - ref_u).abs().max().item() < 1e-9
    import torchcor_amd._lib as L
    L.call("tc_lat_update", L.ptr(u), L.ptr(AT), L.ptr(RT), u.numel(), float(prev), 1,
           L.stream_handle(DEV))
    assert torch.equal(AT, ref_at)


@pytest.mark.parametrize("max_iter", [1, 2])
def test_runtime_small_max_iter_matches_step_loop(max_iter):
    """max_iter below what the steps need: steps end unconverged at
    n_iter = max_iter (solver.py:57-76) and must not stall the pipeline;
    the native loop equals the per-step Python path, float64."""
    _say(f"[gpu] runtime vs step() with max_iter {max_iter}")
    N, steps = 11, 300
    sim = _cube_sim(N, "ttp", torch.float64)
    sim.start(1e-7, 1e-7, max_iter, snapshot_interval=1.0, depth=4)
    sim.advance(steps)
    sim.finish()
    ref_u, ref_it = sim.u.clone(), list(sim.iterations)
    sim = _cube_sim(N, "ttp", torch.float64)
    sim.assemble()
    u = sim.ionic_model.initialize(sim.n_nodes)
    sim._init_solver(u, True)
    AT = torch.full_like(u, -1.0)
    RT = torch.full_like(u, -1.0)
    t, prev, its = 0, None, []
    for n in range(steps):
        t += 0.01
        lat = (AT, RT, prev) if prev is not None else None
        u, it, _, _ = sim.step(sim._x, t, 1e-7, 1e-7, max_iter, False, lat)
        prev = t
        its.append(it)
    assert its == ref_it
    assert max(its) == max_iter and its.count(max_iter) > steps // 2
    assert (u - ref_u).abs().max().item() < 1e-9


@pytest.mark.parametrize("name", ["ttp_endo", "crn"])
def test_single_cell_float32_vs_reference_float32(name):
    """float32 kernels against the reference's own float32 single-cell trace
    (both against its float64 trace for scale)."""
    _say(f"[gpu] single cell f32 {name}")
    from pathlib import Path
    g32 = np.load(Path(__file__).parent / "golden" / f"ionic_{name}_f32.npz")["V"]
    g64 = np.load(Path(__file__).parent / "golden" / f"ionic_{name}_f64.npz")["V"]
    g = _gpu_model("ttp" if name == "ttp_endo" else "crn", torch.float32)
    V = g.initialize(1)
    tr = [V.item()]
    t = 0.0
    nsteps = 6000
    for s in range(1, nsteps + 1):
        V = V + 0.01 * g.differentiate(V)
        t += 0.01
        if 0.0 <= t <= 2.0:
            V = V + 0.01 * 40.0
        if s % 10 == 0:
            tr.append(V.item())
    tr = np.array(tr)
    e32 = np.abs(tr - g32[: len(tr)]).max()
    e64 = np.abs(tr - g64[: len(tr)]).max()
    ref = np.abs(g32[: len(tr)] - g64[: len(tr)]).max()
    _say(f"[gpu] {name}: |gpu32-ref32| {e32:.3e}  |gpu32-ref64| {e64:.3e}  |ref32-ref64| {ref:.3e}")
    assert e32 < 0.05 and e64 < 0.05


def test_runtime_int16_columns_identical(monkeypatch):
    """The hot passes on int16 column offsets give bit-identical results to
    the int32 columns, and a permuted mesh whose offsets do not fit falls
    back to int32 columns."""
    from torchcor_amd.core.system import P1System
    _say("[gpu] runtime int16 vs int32 columns")
    runs = []
    for flag in (True, False):
        monkeypatch.setattr(P1System, "int16_columns", flag)
        sim = _cube_sim(11, "ttp", torch.float32)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1.0, depth=4)
        assert (sim.system.sell.dcol is not None) == flag
        sim.advance(300)
        sim.finish()
        runs.append((sim.u.clone(), list(sim.iterations), sim.frames.clone()))
    assert torch.equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]
    assert torch.equal(runs[0][2], runs[1][2])
    monkeypatch.setattr(P1System, "int16_columns", True)
    nodes, elems = meshes.kuhn_cube(41)
    perm = np.random.default_rng(0).permutation(len(nodes))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    S = _sys(nodes[perm], inv[elems].astype(elems.dtype))
    E = len(elems)
    M, K = S.assemble(sigma_l=torch.full((E,), 0.3, dtype=torch.float64),
                      sigma_t=torch.full((E,), 0.1, dtype=torch.float64),
                      fibres=torch.tensor(_rand_fibres(E, 2)))
    S.set_operators(M, K, 140, 0.01, 0.5, 0.01, torch.float32)
    assert S.bandwidth > 32767 and S.sell.dcol is None


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_shared_slices_identical(monkeypatch, dtype):
    """Slices whose rows share one stencil (the structured cube) are laid out
    on it and read one column offset per entry position: each row keeps its
    entries in column order with zeros where it lacks an offset, so runs,
    products and the external CSR are bit-identical to the per-row layout;
    an unstructured mesh gets no shared slices."""
    from torchcor_amd.core.system import P1System
    _say(f"[gpu] shared vs per-row slices {dtype}")
    runs = []
    for flag in (True, False):
        monkeypatch.setattr(P1System, "share_slices", flag)
        sim = _cube_sim(21, "ttp", dtype)
        sim.start(1e-5, 1e-5, 100, snapshot_interval=1.0, depth=4)
        S = sim.system
        ns = (S.n + 63) // 64
        if flag:
            assert S.sell.slice_tag is not None and S.n_shared >= 0.95 * ns
            assert torch.equal(S.slice_tag & ~1, S.slice_ptr)
        else:
            assert S.sell.slice_tag is None and S.n_shared == 0
        sim.advance(400)
        sim.finish()
        rng = np.random.default_rng(3)
        x = torch.tensor(rng.normal(size=S.n), dtype=dtype, device=DEV)
        prods = [S.spmv(w, x) for w in ("M", "K", "A")]
        A = sim.A
        runs.append((sim.u.clone(), list(sim.iterations), sim.frames.clone(), prods,
                     A.col_indices().clone(), A.values().clone()))
    a, b = runs
    assert torch.equal(a[0], b[0]) and a[1] == b[1] and torch.equal(a[2], b[2])
    for p, q in zip(a[3], b[3]):
        assert torch.equal(p, q)
    assert torch.equal(a[4], b[4]) and torch.equal(a[5], b[5])
    monkeypatch.setattr(P1System, "share_slices", True)
    nodes, elems = meshes.jittered_delaunay(9, seed=3)
    S = _sys(nodes, elems)
    E = len(elems)
    M, K = S.assemble(sigma_l=torch.full((E,), 0.3, dtype=torch.float64),
                      sigma_t=torch.full((E,), 0.1, dtype=torch.float64),
                      fibres=torch.tensor(_rand_fibres(E, 2)))
    S.set_operators(M, K, 140, 0.01, 0.5, 0.01, torch.float32)
    assert S.n_shared <= 0.05 * ((S.n + 63) // 64)

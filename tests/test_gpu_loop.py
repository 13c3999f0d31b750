"""Whole monodomain runs against the float64 oracle: the small configuration
over its full length, runs long enough to repolarise (so RT holds times, not
-1), and the 1M-node cube at full size.

Tolerances:
  * float64 vs the float64 oracle: max |dVm| < 1e-6 mV over every frame,
    per-step n_iter equal, AT and RT equal.
  * float32 (float64 PCG reductions, float64 iterate and right-hand side)
    vs the float64 oracle: max |dVm| <= 0.5 mV over every frame, |dAT| <=
    dt, |dRT| <= dt, the same nodes activated and repolarised (measured:
    MS2v 0.0013 mV, TTP 0.021 mV; RT off by one step at some nodes).  The
    times are stored in float32, so dt is widened by the float32 spacing of
    the latest time (3e-5 ms at 400 ms).
  * 101^3 TTP float32 vs the float64 oracle over the first 20 steps, every
    step's Vm compared: max |dVm| <= 0.5 mV; total n_iter within 2.
"""
import functools

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import ionic as oion  # noqa: E402
from oracle import monodomain as omono  # noqa: E402
from torchcor_amd import meshes  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
DT = 0.01


def _say(msg):
    print(msg, flush=True)


def _gpu_model(name, dtype, **params):
    from torchcor_amd.ionic import ModifiedMS2v, TenTusscherPanfilov
    m = TenTusscherPanfilov("ENDO", DT, device=DEV, dtype=dtype) if name == "ttp" else \
        ModifiedMS2v(DT, device=DEV, dtype=dtype)
    for k, v in params.items():
        setattr(m, k, v)
    return m


def _oracle_model(name, **params):
    m = oion.TenTusscherPanfilov("ENDO", DT) if name == "ttp" else oion.ModifiedMS2v(DT)
    for k, v in params.items():
        setattr(m, k, v)
    return m


def _cube(N):
    nodes, elems = meshes.kuhn_cube(N)
    E = len(elems)
    return nodes, elems, np.zeros(E, np.int32), np.tile([1.0, 0.0, 0.0], (E, 1))


@functools.lru_cache(maxsize=None)
def _oracle_run(N, name, T, params=(), snapshot=1.0, sigma=0.1):
    nodes, elems, regions, fib = _cube(N)
    om = omono.Monodomain(_oracle_model(name, **dict(params)), T=T, dt=DT)
    om.load_arrays(nodes, elems, regions, fib)
    om.add_conductivity([0], sigma, sigma)
    om.add_stimulus(meshes.face_nodes(N, 0, 0), 0.0, 1.0, 100)
    return om.solve(1e-5, 1e-5, 100, snapshot_interval=snapshot)


def _gpu_run(N, name, dtype, T, params=(), snapshot=1.0, sigma=0.1):
    from torchcor_amd.simulator import Monodomain
    nodes, elems, regions, fib = _cube(N)
    gm = Monodomain(_gpu_model(name, dtype, **dict(params)), T=T, dt=DT, device=DEV, dtype=dtype)
    gm.load_arrays(nodes, elems, regions, fib)
    gm.add_conductivity([0], sigma, sigma)
    gm.add_stimulus(meshes.face_nodes(N, 0, 0), 0.0, 1.0, 100)
    gm.solve(1e-5, 1e-5, 100, snapshot_interval=snapshot, verbose=False)
    return gm


def _compare(gm, ro, dtype):
    Vg = gm.frames.cpu().numpy().astype(np.float64)
    AT = gm.activation_time.cpu().numpy().astype(np.float64)
    RT = gm.repolarization_time.cpu().numpy().astype(np.float64)
    assert Vg.shape == ro["Vm"].shape
    err = np.abs(Vg - ro["Vm"]).max()
    if dtype == torch.float64:
        assert err < 1e-6, err
        assert np.array_equal(np.array(gm.iterations), ro["iters"])
        assert np.array_equal(AT, ro["AT"])
        assert np.array_equal(RT, ro["RT"])
    else:
        assert err <= 0.5, err
        for got, ref in ((AT, ro["AT"]), (RT, ro["RT"])):
            assert np.array_equal(got > 0, ref > 0)
            on = ref > 0
            d = np.abs(got[on] - ref[on]).max() if on.any() else 0.0
            _say(f"[gpu] float32: max |dVm| {err:.4f} mV, max |dT| {d:.6f} ms")
            # one step, plus the float32 rounding of the stored time stamp
            assert d <= DT + float(np.spacing(np.float32(ref.max()))) + 1e-9
    return err


def test_config_small_full_length_float64():
    """21^3 MS2v float64 over the configuration's full T = 50 ms (5000 steps)."""
    N, T = 21, 50.0
    _say(f"[gpu] loop {N}^3 MS2v f64, T = {T} ms")
    ro = _oracle_run(N, "ms2v", T)
    gm = _gpu_run(N, "ms2v", torch.float64, T)
    assert len(gm.iterations) == 5000
    _compare(gm, ro, torch.float64)
    assert (ro["AT"] > 0).all()


# tau_close = 6 ms shortens the MS2v action potential to about 45 ms, so a
# 100 ms run repolarises every node
MS2V_SHORT = (("tau_close", 6.0),)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_repolarisation_ms2v(dtype):
    N, T = 11, 100.0
    _say(f"[gpu] loop repolarisation {N}^3 MS2v tau_close 6, T = {T} ms, {dtype}")
    ro = _oracle_run(N, "ms2v", T, MS2V_SHORT)
    assert (ro["RT"] > 0).all() and ro["RT"].min() > ro["AT"].max() - 30.0
    gm = _gpu_run(N, "ms2v", dtype, T, MS2V_SHORT)
    _compare(gm, ro, dtype)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_repolarisation_ttp(dtype):
    # 7^3 nodes (h = 1.67 mm) at sigma = 0.3: the wave crosses the cube in
    # 16 ms with a sharp upstroke (a coarser or less conductive cube is near
    # conduction block, where the slow upstroke makes AT roundoff-sensitive)
    N, T = 7, 400.0
    _say(f"[gpu] loop repolarisation {N}^3 TTP ENDO, T = {T} ms, {dtype}")
    ro = _oracle_run(N, "ttp", T, (), 10.0, 0.3)
    assert (ro["RT"] > 0).all()
    gm = _gpu_run(N, "ttp", dtype, T, (), 10.0, 0.3)
    _compare(gm, ro, dtype)


def test_cube101_float32_vs_oracle_first_steps():
    """The 1M-node TTP cube (101^3, two conductivity regions, rotating
    fibres, 1000-node corner stimulus) in float32 against the float64
    oracle, every one of the first 20 steps."""
    from torchcor_amd.simulator import Monodomain
    from torchcor_amd.ionic import TenTusscherPanfilov
    N, steps = 101, 20
    # T covers the compared steps only (frames every step); the stimulus
    # keeps the 100 ms run's period, so its windows are those of that run
    T = steps * DT
    assert int(T / DT) == steps
    nodes, elems, regions, fib = meshes.cube_workload(N)
    stim = meshes.corner_nodes(N, 10)
    _say(f"[gpu] loop {N}^3 TTP f32, {steps} steps")
    gm = Monodomain(TenTusscherPanfilov("ENDO", DT, device=DEV, dtype=torch.float32), T=T,
                    dt=DT, device=DEV, dtype=torch.float32)
    gm.load_arrays(nodes, elems, regions, fib)
    gm.add_conductivity([0], 0.5272, 0.2076)
    gm.add_conductivity([1], 0.9074, 0.3332)
    gm.add_stimulus(stim, 0.0, 1.0, 100, period=100.0)
    gm.solve(1e-5, 1e-5, 100, snapshot_interval=DT, verbose=False)
    Vg = gm.frames.cpu().numpy().astype(np.float64)
    _say(f"[gpu] loop {N}^3 TTP: float64 oracle")
    om = omono.Monodomain(oion.TenTusscherPanfilov("ENDO", DT), T=T, dt=DT)
    om.load_arrays(nodes, elems, regions, fib)
    om.add_conductivity([0], 0.5272, 0.2076)
    om.add_conductivity([1], 0.9074, 0.3332)
    om.add_stimulus(stim, 0.0, 1.0, 100, period=100.0)
    ro = om.solve(1e-5, 1e-5, 100, snapshot_interval=DT)
    assert Vg.shape == ro["Vm"].shape == (steps + 1, N ** 3)
    err = np.abs(Vg - ro["Vm"]).max(axis=1)
    _say(f"[gpu] 101^3 max |dVm| per step (mV): {np.array2string(err, precision=4)}")
    _say(f"[gpu] 101^3 n_iter per step: GPU {gm.iterations}, oracle {ro['iters'].tolist()}")
    assert err.max() <= 0.5
    assert abs(sum(gm.iterations) - int(ro["iters"].sum())) <= 2
    # the stimulated corner depolarises on both sides
    assert (Vg[-1, stim] - Vg[0, stim]).mean() > 5.0

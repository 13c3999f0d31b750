"""Monodomain driver with the reference's API (torchcor/simulator/monodomain.py:18-226).

Per time step the device runs three stages, each one HIP kernel or a short
fixed sequence of them:

1. the fused ionic kernel: states advance, w = Cm u + dt(dV/100 + Istim/100),
   and the LAT/RT test of the previous step's solution;
2. the right-hand side b = M w - (1-theta) dt K u (one SELL-64 pass);
3. the Jacobi-PCG on A = Cm M + theta dt K, warm-started from the linear
   extrapolation 2x - x_prev, two fused passes per iteration.

The solution vector is kept in the working dtype and rotated between three
buffers (x_prev <- x <- x_next) instead of copied.
"""
import time
from pathlib import Path

import numpy as np
import torch

from .. import _lib as L
from ..core.conductivity import Conductivity
from ..core.mesh import MeshReader
from ..core.stimulation import Stimuli
from ..core.system import P1System
from ..ionic.base import default_device


class Monodomain:
    def __init__(self, ionic_model, T, dt, device=None, dtype=None):
        self.device = torch.device(device) if device is not None else default_device()
        self.dtype = dtype if dtype is not None else torch.float32
        self.T = T
        self.dt = dt
        self.nt = int(T / dt)
        self.ionic_model = ionic_model
        self.Chi = 140
        self.Cm = 0.01
        self.theta = 0.5
        self.n_nodes = None
        self.nodes = self.elems = self.fibres = self.regions = None
        self.stimuli = None
        self.conductivity = None
        self.system = None
        self.mesh_path = None
        self.result_path = None
        self._x = None
        self.kernel_timing = False
        self._events = {"ionic": [], "rhs": []}

    # ------------------------------------------------------------ set-up
    def load_mesh(self, path="Data/ventricle/Case_1", unit_conversion=1000):
        """Read .pts/.elem/.lon from a folder (monodomain.py:60-73)."""
        self.mesh_path = Path(path)
        nodes, elems, regions, fibres = MeshReader(path).read(unit_conversion=unit_conversion)
        self.load_arrays(nodes, elems, regions, fibres)

    def load_arrays(self, nodes, elems, regions, fibres):
        """Same as ``load_mesh`` from in-memory arrays (nodes already in mm)."""
        dev = self.device
        self.n_nodes = int(nodes.shape[0])
        self.nodes = torch.as_tensor(np.asarray(nodes)).to(dtype=self.dtype, device=dev)
        self.elems = torch.as_tensor(np.asarray(elems)).to(dtype=torch.int32, device=dev)
        self.regions = torch.as_tensor(np.asarray(regions)).to(dtype=torch.int32, device=dev)
        self.fibres = torch.as_tensor(np.asarray(fibres)).to(dtype=self.dtype, device=dev)
        self.stimuli = Stimuli(self.n_nodes, dev, self.dtype)
        self.conductivity = Conductivity(self.regions, dtype=self.dtype)
        self.system = None

    def add_stimulus(self, vtx_filepath, start, duration, intensity, period=None, count=1):
        if period is None:
            period = self.T
        self.stimuli.add(vtx_filepath, start, duration, intensity, period, count)

    def add_conductivity(self, region_ids, il, it, el=None, et=None):
        self.conductivity.add(region_ids, il, it, el, et)

    def assemble(self):
        """Pattern, M, K, A and the Jacobi scaling on the GPU (monodomain.py:83-106)."""
        sys_ = P1System(self.nodes, self.elems, self.device)
        # The reference forms elements from coordinates and fibres held in the
        # working dtype; the same rounded inputs are used here.
        M, K = sys_.assemble(sigma_l=self.conductivity.sigma_l, sigma_t=self.conductivity.sigma_t,
                             fibres=self.fibres)
        sys_.release_assembly_maps()
        sys_.set_operators(M, K, self.Chi, self.Cm, self.theta, self.dt, self.dtype)
        self.system = sys_

    @property
    def M(self):
        return self.system.csr(self.system.Mv)

    @property
    def K(self):
        return self.system.csr(self.system.Kv)

    @property
    def A(self):
        return self.system.csr(self.system.Av)

    # ------------------------------------------------------------ stepping
    def _init_solver(self, u, linear_guess=True):
        """ConjugateGradient.initialize (solver.py:23-33): x = u, x_prev = u."""
        self._x = u.clone()
        self._x_prev = u.clone()
        self._x_next = torch.empty_like(u)
        self._w = torch.empty_like(u)
        self._b = torch.empty_like(u)
        self._linear_guess = linear_guess

    def _ionic(self, u, t, lat=None):
        stim = self.stimuli
        bits = stim.active(t) if stim is not None else 0
        AT = RT = None
        lat_t = 0.0
        if lat is not None:
            AT, RT, lat_t = lat
        io = self.ionic_model._io(u, w=self._w, Cm=self.Cm, dt_rhs=self.dt,
                                  stim_mask=stim.mask if bits else None, stim_bits=bits,
                                  stim_amp=stim.amplitudes() if bits else (), AT=AT, RT=RT,
                                  lat_t=lat_t)
        ev = self._event_pair("ionic")
        self.ionic_model.step(io)
        if ev:
            ev[1].record()

    def _electric(self, u, a_tol, r_tol, max_iter):
        sys_ = self.system
        ev = self._event_pair("rhs")
        sys_.rhs(self._w, u, (1 - self.theta) * self.dt, self._b)
        if ev:
            ev[1].record()
        n_iter = sys_.pcg_solve(self._b, self._x, self._x_prev, self._x_next, a_tol, r_tol,
                                max_iter, self._linear_guess)
        if self._linear_guess:
            self._x_prev, self._x, self._x_next = self._x, self._x_next, self._x_prev
        return self._x, n_iter

    def _event_pair(self, name):
        if not self.kernel_timing:
            return None
        pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        pair[0].record()
        self._events[name].append(pair)
        return pair

    def set_kernel_timing(self, on):
        """Time the ionic and RHS kernels and the PCG passes with HIP events
        on the stream they run on (for measurement)."""
        self.kernel_timing = bool(on)
        self._events = {"ionic": [], "rhs": []}
        if self.system is not None:
            self.system.pcg_timing(bool(on))

    def kernel_times(self):
        """{kernel: (total_ms, launches)} since ``set_kernel_timing(True)``."""
        torch.cuda.synchronize(self.device)
        out = {k: (sum(a.elapsed_time(b) for a, b in v), len(v)) for k, v in self._events.items()}
        out.update({f"pcg_{k}": v for k, v in self.system.pcg_timing().items()})
        return out

    def step(self, u, t, a_tol, r_tol, max_iter, verbose=False, _lat=None):
        """One time step (monodomain.py:109-139): returns (u, n_iter, ionic_s, electric_s)."""
        if self._x is None:
            self._init_solver(u)
        if u.data_ptr() != self._x.data_ptr():
            self._x.copy_(u)
        ionic_time = electric_time = 0.0
        if verbose:
            torch.cuda.synchronize(self.device)
            t0 = time.time()
        self._ionic(self._x, t, _lat)
        if verbose:
            torch.cuda.synchronize(self.device)
            ionic_time = time.time() - t0
            t0 = time.time()
        u, n_iter = self._electric(self._x, a_tol, r_tol, max_iter)
        if verbose:
            torch.cuda.synchronize(self.device)
            electric_time = time.time() - t0
        return u, n_iter, ionic_time, electric_time

    def solve(self, a_tol, r_tol, max_iter, calculate_AT_RT=True, linear_guess=True,
              snapshot_interval=1, verbose=True, result_path=None, nsteps=None):
        """Time loop (monodomain.py:141-226).  Writes ATs.pt, RTs.pt and Vm.pt
        like the reference and returns the total PCG iteration count.
        ``nsteps`` (extension) stops after that many steps."""
        if result_path is not None:
            self.result_path = Path(result_path)
            self.result_path.mkdir(parents=True, exist_ok=True)
        self.start(a_tol, r_tol, max_iter, calculate_AT_RT, linear_guess, snapshot_interval, verbose)
        nt = self.nt if nsteps is None else min(int(nsteps), self.nt)
        self.advance(nt)
        return self.finish()

    # The loop of ``solve`` split in three so that callers (benchmarks,
    # servers) can run it in pieces: start -> advance(k)... -> finish.
    def start(self, a_tol, r_tol, max_iter, calculate_AT_RT=True, linear_guess=True,
              snapshot_interval=1, verbose=False):
        self.assemble()
        u = self.ionic_model.initialize(self.n_nodes)
        self._init_solver(u, linear_guess)
        self._run = dict(a_tol=a_tol, r_tol=r_tol, max_iter=max_iter, calc=calculate_AT_RT,
                         tpf=int(snapshot_interval / self.dt), snapshot_interval=snapshot_interval,
                         verbose=verbose, t=0, n=0, prev_t=None, n_total_iter=0, tot_ion=0.0,
                         tot_ele=0.0, t0=time.time())
        self.activation_time = self.repolarization_time = None
        if calculate_AT_RT:
            self.activation_time = torch.ones_like(u) * -1
            self.repolarization_time = torch.ones_like(u) * -1
        self.frames = [u.clone()]
        self.iterations = []

    def advance(self, k):
        R = self._run
        AT, RT = self.activation_time, self.repolarization_time
        for _ in range(int(k)):
            R["n"] += 1
            R["t"] += self.dt
            t = R["t"]
            lat = (AT, RT, R["prev_t"]) if (R["calc"] and R["prev_t"] is not None) else None
            u, n_iter, it_s, et_s = self.step(self._x, t, R["a_tol"], R["r_tol"], R["max_iter"],
                                              R["verbose"], lat)
            R["prev_t"] = t
            R["n_total_iter"] += n_iter
            self.iterations.append(n_iter)
            R["tot_ion"] += it_s
            R["tot_ele"] += et_s
            if R["tpf"] > 0 and R["n"] % R["tpf"] == 0:
                self.frames.append(u.clone())
                if R["verbose"] and R["snapshot_interval"] != self.T:
                    print(f"t: {round(t, 1)}/{self.T}", "Time elapsed:",
                          round(time.time() - R["t0"], 2), "Total CG iter:", R["n_total_iter"],
                          flush=True)

    def finish(self):
        R = self._run
        AT, RT = self.activation_time, self.repolarization_time
        if R["calc"] and R["prev_t"] is not None:
            L.call("tc_lat_update", L.ptr(self._x), L.ptr(AT), L.ptr(RT), self.n_nodes,
                   float(R["prev_t"]), L.dtype_code(self.dtype), L.stream_handle(self.device))
        torch.cuda.synchronize(self.device)
        self.solving_time = time.time() - R["t0"]
        self.u = self._x
        if self.result_path is not None:
            if R["calc"]:
                torch.save(AT.cpu(), self.result_path / "ATs.pt")
                torch.save(RT.cpu(), self.result_path / "RTs.pt")
            if R["snapshot_interval"] < self.T:
                torch.save(torch.stack(self.frames, dim=0).cpu(), self.result_path / "Vm.pt")
        if R["verbose"]:
            mem = torch.cuda.max_memory_allocated(self.device) / 1e9
            print(self.ionic_model.name, self.n_nodes, round(self.solving_time, 2),
                  round(R["tot_ion"], 2), round(R["tot_ele"], 2), R["n_total_iter"], "n/a",
                  f"{round(mem, 2)}", flush=True)
            if R["calc"]:
                print("ATs: ", AT.min().item(), AT.max().item(), flush=True)
                print("RTs: ", RT.min().item(), RT.max().item(), flush=True)
        return R["n_total_iter"]

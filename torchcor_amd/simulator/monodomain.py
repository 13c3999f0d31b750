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

Node order.  When the mesh's own numbering scatters neighbours (bandwidth of
the pattern above n/8 and at least twice what the new order gives), the
operators and every per-node array live in a locality order on the device:
the Cuthill-McKee level structure with each level along a Z-order curve of
the node coordinates (``level_order``), and in "auto" mode, when the staged
split's column sets fit a stage, the rows of each of its windows then sorted
by length (``slice_sort``: less SELL padding, the same column set a window).  The mesh arrays, stimulus node ids and every
output (u, frames, AT, RT, M, K, A) stay in the caller's numbering.
``reorder`` selects this: "auto" (default), "levels" (always, the level
order alone), "rcm" (always, plain reverse Cuthill-McKee) or None (never).
"""
import time
from pathlib import Path

import numpy as np
import torch

from .. import _lib as L
from ..core.conductivity import Conductivity
from ..core.mesh import MeshReader
from ..core.reorder import order_bandwidth, pick_order, slice_sort, window_columns
from ..core.stimulation import Stimuli
from ..core.system import P1System
from ..ionic.base import default_device

# Vm snapshots up to this size stay in device memory; larger runs stream them
# to pinned host memory through a ring of RING_SLOTS device frames.
FRAMES_DEVICE_BYTES = 16 << 30
RING_SLOTS = 4
# with verbose output, every VERBOSE_TIMING_PERIOD-th step is timed per kernel
VERBOSE_TIMING_PERIOD = 16


class Monodomain:
    #: the time-step runtime keeps the ionic states in blocks of 64 nodes
    #: (all of a node block's states in one contiguous span, tc_sim_desc
    #: state_block); the model's (S, n) states are copied in at start and
    #: back at finish.  Same results either way.
    block_states = True

    def __init__(self, ionic_model, T, dt, device=None, dtype=None, reorder="auto"):
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
        self._sim = None
        if reorder not in ("auto", "levels", "rcm", None):
            raise ValueError("reorder must be 'auto', 'levels', 'rcm' or None")
        self.reorder = reorder
        self.node_order = None   # internal position i holds mesh node node_order[i]
        self.node_map = None     # mesh node j sits at internal position node_map[j]
        self.row_window = 0      # rows sorted by length in windows of this many (0: not)
        self._state_blocks = None

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

    def _choose_order(self, sys_):
        """Node order for the pattern of ``sys_`` if the mode asks for one
        (``pick_order``; "auto" with staged gathers also groups rows by
        length inside the staged split's block windows)."""
        self.node_order = self.node_map = None
        self.row_window = 0
        n = sys_.n
        res = pick_order(sys_.row_ptr, sys_.col, n, self.nodes, self.reorder)
        if res is None:
            return None
        order, self.bandwidth_natural, self.bandwidth_reordered = res
        if self.reorder == "auto" and P1System.stage_gathers:
            # rows grouped by length inside the staged split's windows, when
            # that split's column sets fit a stage (else the passes gather per
            # entry, which the grouping would slow)
            window = P1System.sort_window(n)
            if window_columns(sys_.row_ptr, sys_.col, order, window) <= L.STAGE_MAX_COLS:
                order = slice_sort(order, sys_.row_ptr, window)
                self.bandwidth_reordered = order_bandwidth(sys_.row_ptr, sys_.col, order)
                self.row_window = window
        dev = self.device
        self.node_order = torch.as_tensor(order, dtype=torch.long, device=dev)
        self.node_map = torch.empty_like(self.node_order)
        self.node_map[self.node_order] = torch.arange(n, dtype=torch.long, device=dev)
        return self.node_order

    def _to_mesh_order(self, v):
        """A per-node array (last axis) from the internal order to the mesh's."""
        return v if self.node_map is None else v[..., self.node_map.to(v.device)]

    def _to_internal(self, v):
        return v if self.node_order is None else v[..., self.node_order]

    def assemble(self):
        """Pattern, M, K, A and the Jacobi scaling on the GPU (monodomain.py:83-106)."""
        self._mask_int = None
        sys_ = P1System(self.nodes, self.elems, self.device)
        if self._choose_order(sys_) is not None:
            elems = self.node_map[self.elems.to(torch.long)].to(torch.int32)
            sys_ = P1System(self.nodes[self.node_order], elems, self.device)
            sys_.row_window = self.row_window
        # The reference forms elements from coordinates and fibres held in the
        # working dtype; the same rounded inputs are used here.
        M, K = sys_.assemble(sigma_l=self.conductivity.sigma_l, sigma_t=self.conductivity.sigma_t,
                             fibres=self.fibres)
        sys_.release_assembly_maps()
        sys_.set_operators(M, K, self.Chi, self.Cm, self.theta, self.dt, self.dtype)
        self.system = sys_

    def _csr(self, vals):
        if self.node_order is None:
            return self.system.csr(vals)
        # relabel the internal pattern to mesh numbering: the same pattern and,
        # entry by entry, the same values as assembling in mesh order
        sys_, n = self.system, self.system.n
        rows = torch.repeat_interleave(torch.arange(n, device=sys_.device),
                                       sys_.row_ptr.diff().to(torch.int64))
        r = self.node_order[rows]
        c = self.node_order[sys_.col.to(torch.int64)]
        perm = torch.argsort(r * n + c)
        crow = torch.zeros(n + 1, dtype=torch.int64, device=sys_.device)
        crow[1:] = torch.cumsum(torch.bincount(r, minlength=n), 0)
        return torch.sparse_csr_tensor(crow, c[perm], vals[perm], size=(n, n))

    @property
    def M(self):
        return self._csr(self.system.Mv)

    @property
    def K(self):
        return self._csr(self.system.Kv)

    @property
    def A(self):
        return self._csr(self.system.Av)

    # ------------------------------------------------------------ stepping
    def _init_solver(self, u, linear_guess=True):
        """ConjugateGradient.initialize (solver.py:23-33): the iterate x = u in
        float64 and x_prev = u in u's dtype.  x_prev is held in float64
        storage with values rounded to the working dtype, as the reference's
        float32 x_prev rounds each copy of x (solver.py:26, 35-38)."""
        f64 = torch.float64
        self._x = u.to(f64).clone()
        self._x_prev = u.to(f64).clone()
        self._x_next = torch.empty_like(self._x)
        self._w = torch.empty_like(u)
        self._b = torch.empty_like(u)
        self._linear_guess = linear_guess

    def _stim_mask(self):
        """The stimulus bit mask in the internal node order (cached per assembly)."""
        if getattr(self, "_mask_int", None) is None:
            m = self.stimuli.mask
            self._mask_int = m if self.node_order is None else m[self.node_order].contiguous()
        return self._mask_int

    def _ionic(self, u, t, lat=None):
        stim = self.stimuli
        bits = stim.active(t) if stim is not None else 0
        AT = RT = None
        lat_t = 0.0
        if lat is not None:
            AT, RT, lat_t = lat
        io = self.ionic_model._io(u, w=self._w, Cm=self.Cm, dt_rhs=self.dt,
                                  stim_mask=self._stim_mask() if bits else None, stim_bits=bits,
                                  stim_amp=stim.amplitudes() if bits else (), AT=AT, RT=RT,
                                  lat_t=lat_t)
        self.ionic_model.step(io)

    def _electric(self, u, a_tol, r_tol, max_iter):
        """b = M w - (1-theta) dt K u in the working dtype, then the float64
        PCG (monodomain.py:127-133, solver.py:40-76): b converted to float64,
        the float64 iterate warm-started by the linear guess, the solution
        returned in the working dtype."""
        sys_ = self.system
        sys_.rhs(self._w, u, (1 - self.theta) * self.dt, self._b)
        f64 = torch.float64
        b = self._b if self._b.dtype == f64 else self._b.to(f64)
        n_iter = sys_.pcg_solve(b, self._x, self._x_prev, self._x_next, a_tol, r_tol, max_iter,
                                self._linear_guess, plan=sys_.pcg_plan_f64())
        if self._linear_guess:
            self._x_prev, self._x, self._x_next = self._x, self._x_next, self._x_prev
            if self.dtype != f64:
                self._x_prev.copy_(self._x_prev.to(self.dtype))
        return self._x.to(self.dtype), n_iter

    def step(self, u, t, a_tol, r_tol, max_iter, verbose=False, _lat=None):
        """One time step (monodomain.py:109-139): returns (u, n_iter, ionic_s, electric_s).
        ``u`` is in mesh numbering; the returned u too.  As in the reference,
        ``u`` feeds the ionic update and K u, while the PCG keeps its own
        float64 iterate between steps (set from the first step's u)."""
        if self.system is None:
            self.assemble()
        if self.node_order is not None and _lat is not None:
            raise NotImplementedError("LAT arrays in step() need the mesh's own node order")
        ui = self._to_internal(u).to(self.dtype).contiguous()
        if self._x is None:
            self._init_solver(ui)
        ionic_time = electric_time = 0.0
        if verbose:
            torch.cuda.synchronize(self.device)
            t0 = time.time()
        self._ionic(ui, t, _lat)
        if verbose:
            torch.cuda.synchronize(self.device)
            ionic_time = time.time() - t0
            t0 = time.time()
        u, n_iter = self._electric(ui, a_tol, r_tol, max_iter)
        if verbose:
            torch.cuda.synchronize(self.device)
            electric_time = time.time() - t0
        return self._to_mesh_order(u), n_iter, ionic_time, electric_time

    def solve(self, a_tol, r_tol, max_iter, calculate_AT_RT=True, linear_guess=True,
              snapshot_interval=1, verbose=True, result_path=None, nsteps=None, frames="auto"):
        """Time loop (monodomain.py:141-226).  Writes ATs.pt, RTs.pt and Vm.pt
        like the reference and returns the total PCG iteration count.
        Extensions: ``nsteps`` stops after that many steps; ``frames`` places
        the Vm snapshots (see ``start``)."""
        if result_path is not None:
            self.result_path = Path(result_path)
            self.result_path.mkdir(parents=True, exist_ok=True)
        self.start(a_tol, r_tol, max_iter, calculate_AT_RT, linear_guess, snapshot_interval, verbose,
                   frames=frames)
        nt = self.nt if nsteps is None else min(int(nsteps), self.nt)
        self.advance(nt)
        return self.finish()

    # The loop of ``solve`` split in three so that callers (benchmarks,
    # servers) can run it in pieces: start -> advance(k)... -> finish.  The
    # loop itself runs in the native runtime (tc_sim_*): steps are pipelined
    # on the device and no host round trip is needed per step.
    def start(self, a_tol, r_tol, max_iter, calculate_AT_RT=True, linear_guess=True,
              snapshot_interval=1, verbose=False, depth=4, x_f64=True, precise_rhs=True,
              frames="auto", separate_check=False, k_history=0, graph=False):
        """Set up a run.  ``x_f64`` keeps the PCG iterate in float64 in float32
        runs, as the reference's ConjugateGradient does (solver.py:24-26);
        x_prev stays in the working dtype.  ``precise_rhs`` also keeps the
        right-hand side in float64 (w stored in float64, b never rounded);
        with it off a float32 run rounds w and b as the reference does.
        ``frames`` places the Vm snapshots: "device" keeps them in device
        memory, as the reference's list of snapshots does
        (monodomain.py:162,186-187); "host" streams each one through a
        device ring into pinned host memory while the run goes on; "auto"
        is "device" up to FRAMES_DEVICE_BYTES of frames, else "host".
        ``separate_check`` runs each step's PCG stopping test in a pass of
        its own instead of in the step's last update pass (same results).
        ``k_history``: each step launches as many PCG iterations as the most
        any of the last ``k_history`` steps needed (0: the runtime's
        default) before its status is read; any value gives the same
        results.  ``graph`` replays each step's RHS and PCG launches as a
        captured HIP graph (one per buffer parity and iteration count; same
        results)."""
        import ctypes as C
        self.assemble()
        ion = self.ionic_model
        u = ion.initialize(self.n_nodes)
        x_dtype = torch.float64 if (x_f64 or self.dtype == torch.float64) else self.dtype
        self._X = [u.to(x_dtype).clone(), torch.empty_like(u, dtype=x_dtype)]
        self._XP = [u.clone(), torch.empty_like(u)]
        w_f64 = bool(precise_rhs and x_dtype == torch.float64)
        self._w = torch.empty_like(u, dtype=torch.float64 if w_f64 else self.dtype)
        tpf = int(snapshot_interval / self.dt)
        nframes = self.nt // tpf if tpf > 0 else 0
        nbytes = (nframes + 1) * self.n_nodes * u.element_size()
        if frames == "auto":
            frames = "host" if nbytes > FRAMES_DEVICE_BYTES else "device"
        if frames == "device":
            self._ring = None
            self._frames = torch.empty((nframes + 1, self.n_nodes), dtype=self.dtype,
                                       device=self.device)
        elif frames == "host":
            self._ring = torch.empty((RING_SLOTS, self.n_nodes), dtype=self.dtype,
                                     device=self.device)
            self._frames = torch.empty((nframes + 1, self.n_nodes), dtype=self.dtype,
                                       pin_memory=True)
        else:
            raise ValueError(f"frames must be 'auto', 'device' or 'host', not {frames!r}")
        self._frames[0].copy_(u)
        self.activation_time = self.repolarization_time = None
        if calculate_AT_RT:
            self.activation_time = torch.ones_like(u) * -1
            self.repolarization_time = torch.ones_like(u) * -1
        d = L.SimDesc()
        d.model = ion.MODEL
        self._cparams = ion.c_params()
        d.params = C.cast(C.pointer(self._cparams), C.c_void_p)
        self._state_blocks = None
        if self.block_states:
            # (S, n) rows -> (ceil(n/64), S, 64) blocks
            S, n = ion.states.shape
            nb = -(-n // 64)
            rows = torch.zeros((S, nb * 64), dtype=ion.states.dtype, device=ion.states.device)
            rows[:, :n] = ion.states
            self._state_blocks = rows.view(S, nb, 64).permute(1, 0, 2).contiguous()
            d.states = L.ptr(self._state_blocks)
            d.state_stride = 0
            d.state_block = 1
        else:
            d.states = L.ptr(ion.states)
            d.state_stride = ion.states.stride(0)
        for k, tab in enumerate(ion.c_tables()):
            d.tab[k] = L.ptr(tab)
        d.pcg = self.system._pcg
        d.Mv, d.Kv = L.ptr(self.system.sM), L.ptr(self.system.sK)
        d.dtype = L.dtype_code(self.dtype)
        d.n = self.n_nodes
        d.dt, d.dt_ionic, d.Cm, d.theta = self.dt, ion.dt, self.Cm, self.theta
        d.x_f64 = 1 if x_dtype == torch.float64 else 0
        d.w_f64 = 1 if w_f64 else 0
        for k in range(2):
            d.x[k] = L.ptr(self._X[k])
            d.xp[k] = L.ptr(self._XP[k])
        d.w = L.ptr(self._w)
        d.AT = L.ptr(self.activation_time)
        d.RT = L.ptr(self.repolarization_time)
        st = self.stimuli.stimulus_list
        d.n_stim = len(st)
        d.stim_mask = L.ptr(self._stim_mask()) if st else None
        for k, sc in enumerate(st):
            d.stim[k] = L.Stimulus(float(sc.start), float(sc.duration), float(sc.period),
                                   float(sc.count), float(sc.intensity))
        d.a_tol, d.r_tol, d.max_iter = float(a_tol), float(r_tol), int(max_iter)
        d.linear_guess = 1 if linear_guess else 0
        d.depth = int(depth)
        d.separate_check = 1 if separate_check else 0
        d.k_history = int(k_history)
        d.graph = 1 if graph else 0
        self._free_sim()
        h = C.c_void_p(0)
        L.call("tc_sim_create", C.byref(h), C.byref(d))
        self._sim = h
        self._run = dict(calc=calculate_AT_RT, tpf=tpf, nframes=nframes,
                         snapshot_interval=snapshot_interval, verbose=verbose, n_total_iter=0,
                         t0=time.time())
        self.iterations = []
        if verbose:
            # per-phase totals for the closing log line, from HIP-event
            # timings of every 16th step (timing every step slows the loop)
            self.set_kernel_timing(True, VERBOSE_TIMING_PERIOD)

    def _free_sim(self):
        if getattr(self, "_sim", None):
            L.lib.tc_sim_destroy(self._sim)
            self._sim = None

    def __del__(self):
        try:
            self._free_sim()
        except Exception:
            pass

    def sim_status(self):
        import ctypes as C
        t, steps, xi, fr = C.c_double(), C.c_int64(), C.c_int(), C.c_int64()
        L.call("tc_sim_status", self._sim, C.byref(t), C.byref(steps), C.byref(xi), C.byref(fr))
        return t.value, steps.value, xi.value, fr.value

    def advance(self, k):
        """Run k more time steps in the native runtime."""
        import ctypes as C
        R = self._run
        k = int(k)
        while k > 0:
            chunk = k
            if R["verbose"] and R["tpf"] > 0:
                _, steps, _, _ = self.sim_status()
                chunk = min(k, R["tpf"] - steps % R["tpf"])
            iters = (C.c_int32 * chunk)()
            frames = self._frames[1:] if R["nframes"] > 0 else None
            if frames is not None and self._ring is not None:
                L.call("tc_sim_advance_ring", self._sim, chunk, iters, frames.data_ptr(),
                       L.ptr(self._ring), RING_SLOTS, R["tpf"], R["nframes"],
                       L.stream_handle(self.device))
            else:
                L.call("tc_sim_advance", self._sim, chunk, iters,
                       L.ptr(frames) if frames is not None else None, R["tpf"], R["nframes"],
                       L.stream_handle(self.device))
            self.iterations.extend(iters)
            R["n_total_iter"] += sum(iters)
            k -= chunk
            if R["verbose"] and R["tpf"] > 0 and R["snapshot_interval"] != self.T:
                t, steps, _, _ = self.sim_status()
                if steps % R["tpf"] == 0:
                    print(f"t: {round(t, 1)}/{self.T}", "Time elapsed:",
                          round(time.time() - R["t0"], 2), "Total CG iter:", R["n_total_iter"],
                          flush=True)

    def set_kernel_timing(self, on, period=1):
        """Time each kernel of the runtime with HIP events around its launch,
        on every ``period``-th step."""
        L.call("tc_sim_set_timing", self._sim, int(period) if on else 0)

    def graph_stats(self):
        """(graphs recorded, graph replays) of the current run."""
        import ctypes as C
        cap, rep = C.c_int64(0), C.c_int64(0)
        L.call("tc_sim_graph_stats", self._sim, C.byref(cap), C.byref(rep))
        return cap.value, rep.value

    def kernel_times(self):
        """{kernel: (total_ms, launches)} since ``set_kernel_timing(True)``."""
        import ctypes as C
        ms = (C.c_double * 5)()
        cnt = (C.c_int64 * 5)()
        L.call("tc_sim_get_timing", self._sim, ms, cnt)
        names = ("ionic", "rhs_init", "pcg_spmv", "pcg_update", "pcg_check")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(names)}

    def phase_totals(self, steps):
        """(ionic seconds, electric seconds, GPU-busy % of the wall time) for
        ``steps`` steps, from the sampled kernel timings (zeros when no step
        was timed)."""
        kt = self.kernel_times()
        sampled = kt["ionic"][1]
        if sampled <= 0:
            return 0.0, 0.0, 0.0
        ion = kt["ionic"][0] / sampled * steps / 1e3
        ele = sum(kt[k][0] for k in ("rhs_init", "pcg_spmv", "pcg_update", "pcg_check")) \
            / sampled * steps / 1e3
        wall = getattr(self, "solving_time", 0.0)
        util = 100.0 * (ion + ele) / wall if wall > 0 else 0.0
        return ion, ele, min(util, 100.0)

    # ------------------------------------------------------------ output
    def pt_to_vtk(self):
        """Vm.pt frames, ATs.pt and RTs.pt of ``result_path`` as legacy VTK
        files (monodomain.py:229-251): Vm_vtk/frame_{i}.vtk, ATs.vtk, RTs.vtk."""
        from ..core.vtk import VTK3D, VTK3DSurface
        t0 = time.time()
        viz = (VTK3DSurface if self.elems.shape[1] == 3 else VTK3D)(self.nodes, self.elems)
        frames = torch.load(self.result_path / "Vm.pt", weights_only=True)
        for i in range(frames.shape[0]):
            viz.save_frame(frames[i], self.result_path / f"Vm_vtk/frame_{i}.vtk")
        for name in ("ATs", "RTs"):
            viz.save_frame(torch.load(self.result_path / f"{name}.pt", weights_only=True),
                           self.result_path / f"{name}.vtk")
        print(f"Saved vtk files in {round(time.time() - t0, 2)}", flush=True)

    # ------------------------------------------------------ phi_e recovery
    def phie_recovery(self, a_tol=1e-5, r_tol=1e-5, max_iter=100, Vm=None, save=True):
        """Extracellular potential of every Vm snapshot (monodomain.py:254-286).

        For each frame V: b = -K_i V, then K_ie phi = b by the Jacobi-PCG
        (float64, warm-started from the previous frames' solutions by the
        linear guess, as one ConjugateGradient object carried over the
        frames), and phi minus its mean.  K_i and K_ie are the stiffness
        matrices of sigma_i and sigma_i + sigma_e (conductivities given with
        el/et), assembled like the monodomain operators.  As in the
        reference, a float64 run subtracts the mean from the solver's own
        iterate (the returned solution is that iterate), a float32 run from
        a float32 copy.  ``Vm`` (frames, n) defaults to ``result_path /
        'Vm.pt'``; the result is saved as ``Phi_e.pt`` there when ``save``
        and returned, (frames, n) in the working dtype, mesh numbering.
        """
        import ctypes as C
        dev, dt_ = self.device, self.dtype
        if Vm is None:
            if self.result_path is None:
                raise ValueError("phie_recovery needs Vm frames or a result_path holding Vm.pt")
            Vm = torch.load(self.result_path / "Vm.pt", weights_only=True)
        Vm = torch.as_tensor(Vm).to(device=dev, dtype=dt_)
        if Vm.dim() != 2 or Vm.shape[1] != self.n_nodes:
            raise ValueError(f"Vm must be (frames, {self.n_nodes})")
        si, se, _ = self.conductivity.calculate_sigma(self.fibres)
        if si is None or se is None:
            raise ValueError("phie_recovery needs intra- and extracellular conductivities "
                             "(add_conductivity with el and et)")
        if self.system is None:
            self.assemble()
        order = self.node_order
        if order is not None:
            elems = self.node_map[self.elems.to(torch.long)].to(torch.int32)
            P = P1System(self.nodes[order], elems, dev)
            P.row_window = self.row_window
            Vm = Vm[:, order]
        else:
            P = P1System(self.nodes, self.elems, dev)
        _, K_ie = P.assemble(sigma=si + se)
        _, K_i = P.assemble(sigma=si)
        P.release_assembly_maps()
        f64 = torch.float64
        # the reference's CG converts K_ie (held in the working dtype) to
        # float64 and keeps its Jacobi scaling in the working dtype
        K_ie = K_ie.to(dt_).to(f64)
        P.set_operators(K_ie, torch.zeros_like(K_ie), 1.0, 1.0, 0.0, 0.0, f64)
        P.Dinv.copy_(P.Dinv.to(dt_).to(f64))
        sK_i = P.to_sell(K_i.to(dt_))
        n = P.n
        x = torch.zeros(n, dtype=f64, device=dev)
        x_prev, x_next = torch.zeros_like(x), torch.empty_like(x)
        b = torch.empty(n, dtype=dt_, device=dev)
        out = torch.empty_like(Vm)
        self.phie_iterations = []
        for k in range(Vm.shape[0]):
            V = Vm[k].contiguous()
            L.call("tc_sell_spmv", C.byref(P.sell), L.ptr(sK_i), L.ptr(V), L.ptr(b),
                   L.dtype_code(dt_), L.stream_handle(dev))
            rhs = (-b).to(f64)
            it = P.pcg_solve(rhs, x, x_prev, x_next, a_tol, r_tol, max_iter, True)
            x_prev, x, x_next = x, x_next, x_prev
            self.phie_iterations.append(it)
            if dt_ == f64:
                x -= x.mean()
                out[k].copy_(x)
            else:
                phi = x.to(dt_)
                out[k].copy_(phi - phi.mean())
        if order is not None:
            out = out[:, self.node_map]
        if save and self.result_path is not None:
            torch.save(out.cpu(), self.result_path / "Phi_e.pt")
        return out

    def finish(self):
        R = self._run
        L.call("tc_sim_finish", self._sim, L.stream_handle(self.device))
        if self._state_blocks is not None:
            # the runtime's state blocks back into the model's (S, n) rows
            st = self.ionic_model.states
            S, n = st.shape
            st.copy_(self._state_blocks.permute(1, 0, 2).reshape(S, -1)[:, :n])
        torch.cuda.synchronize(self.device)
        t, steps, xi, nfr = self.sim_status()
        self.solving_time = time.time() - R["t0"]
        self.u = self._to_mesh_order(self._X[xi].to(self.dtype))
        self.frames = self._to_mesh_order(self._frames[: 1 + nfr])
        if self.node_order is not None and self.activation_time is not None:
            self.activation_time = self._to_mesh_order(self.activation_time)
            self.repolarization_time = self._to_mesh_order(self.repolarization_time)
        AT, RT = self.activation_time, self.repolarization_time
        if self.result_path is not None:
            if R["calc"]:
                torch.save(AT.cpu(), self.result_path / "ATs.pt")
                torch.save(RT.cpu(), self.result_path / "RTs.pt")
            if R["snapshot_interval"] < self.T:
                torch.save(self.frames.cpu(), self.result_path / "Vm.pt")
        if R["verbose"]:
            # the reference's closing line (monodomain.py:211-220): model, nodes,
            # wall time, ionic and electric totals, CG iterations, GPU
            # utilisation (%), memory (GB).  The phase totals and the
            # utilisation (kernel-busy time over wall time) are the sampled
            # per-step kernel durations scaled to all steps.
            ion_s, ele_s, util = self.phase_totals(steps)
            mem = torch.cuda.max_memory_allocated(self.device) / 1e9
            print(self.ionic_model.name, self.n_nodes, round(self.solving_time, 2),
                  round(ion_s, 2), round(ele_s, 2), R["n_total_iter"], f"{round(util, 2)}",
                  f"{round(mem, 2)}", flush=True)
            if R["calc"]:
                print("ATs: ", AT.min().item(), AT.max().item(), flush=True)
                print("RTs: ", RT.min().item(), RT.max().item(), flush=True)
        return R["n_total_iter"]

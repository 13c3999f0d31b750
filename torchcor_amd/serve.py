"""Served stream: monodomain runs on different meshes, continuously batched.

Requests (a mesh, conductivities, an ionic model, stimuli and a duration T)
arrive over time.  ``ServeEngine`` keeps a device pool sized by an HBM
budget; a request whose operator and state fit is admitted as soon as it
has arrived and room is free, and from then on advances with every other
admitted run in the same kernels (one time step of every live run per
engine step, each with its own PCG scalars and stopping test, clock and
stimuli).  A finished run is retired, its result copied out and its room
reused.  ``serve`` drives the loop in real time and reports the latency per
simulated second of every request and the throughput.

Each run follows the single-run path exactly: the same GPU assembly (with
the reference's float32 rounding of element contributions), the same SELL
operator values and Jacobi scaling, the same time loop semantics; its
result does not depend on which other runs share the pool.  A mesh whose
node numbering scatters neighbours runs in a locality order inside the pool
(as ``Monodomain(reorder=...)``, without the row grouping of the staged
single-run passes, which the pool's per-entry gathers do not use); its
stimulus node ids and its results stay in the mesh's numbering.
"""
import collections
import concurrent.futures
import ctypes as C
import time

import numpy as np
import torch

from . import _lib as L
from .core.conductivity import Conductivity
from .core.reorder import pick_order
from .core.system import P1System
from .ionic.base import default_device
from .ionic.models import CourtemancheRamirezNattel, ModifiedMS2v, TenTusscherPanfilov

MODEL_CLASSES = {L.MODEL_MS2V: lambda dt, dev, dt_: ModifiedMS2v(dt, device=dev, dtype=dt_),
                 L.MODEL_TTP: lambda dt, dev, dt_: TenTusscherPanfilov("ENDO", dt, device=dev,
                                                                        dtype=dt_),
                 L.MODEL_CRN: lambda dt, dev, dt_: CourtemancheRamirezNattel(dt, device=dev,
                                                                            dtype=dt_)}
MODEL_NAMES = {"ms2v": L.MODEL_MS2V, "ttp": L.MODEL_TTP, "crn": L.MODEL_CRN}


class Request:
    """One simulation: mesh arrays (mm), conductivity calls as for
    ``Monodomain.add_conductivity``, model name ('ms2v', 'ttp' (ENDO), 'crn'),
    stimuli (node_ids, start, duration, intensity[, period[, count]]), T (ms).
    """

    def __init__(self, nodes, elems, regions, fibres, conductivities, model, stimuli, T,
                 arrival=0.0, tag=None):
        self.nodes, self.elems, self.regions, self.fibres = nodes, elems, regions, fibres
        self.conductivities = list(conductivities)
        self.model = MODEL_NAMES[model] if isinstance(model, str) else int(model)
        self.stimuli = list(stimuli)
        self.T = float(T)
        self.arrival = float(arrival)
        self.tag = tag


class Prepared:
    pass


class ServeEngine:
    Chi = 140
    Cm = 0.01
    theta = 0.5

    def __init__(self, cap_nodes, cap_entries, max_sys=256, dt=0.01, dtype=torch.float32,
                 device=None, a_tol=1e-5, r_tol=1e-5, max_iter=100, x_f64=True,
                 precise_rhs=True, lat=True, depth=4, int16_columns=True, slices_per_block=None,
                 k_history=0, graph=False, reorder="auto"):
        if reorder not in ("auto", "levels", "rcm", None):
            raise ValueError("reorder must be 'auto', 'levels', 'rcm' or None")
        self.reorder = reorder
        self.device = torch.device(device) if device is not None else default_device()
        dev = self.device
        self.dtype = dtype
        self.dt = dt
        cap_slices = (int(cap_nodes) + 63) // 64
        rows = cap_slices * 64
        x_dtype = torch.float64 if (x_f64 or dtype == torch.float64) else dtype
        w_f64 = bool(precise_rhs and x_dtype == torch.float64)
        z = lambda m, t: torch.zeros(m, dtype=t, device=dev)  # noqa: E731
        self.pool = dict(
            slice_off=z(cap_slices, torch.int32), slice_w=z(cap_slices, torch.int32),
            slice_sys=z(cap_slices, torch.int32), col=z(cap_entries, torch.int32),
            dcol=z(cap_entries, torch.int16) if int16_columns else None,
            pat=z(max(int(cap_entries) // 64, 1), torch.int32) if int16_columns else None,
            Mv=z(cap_entries, dtype), Kv=z(cap_entries, dtype), Av=z(cap_entries, dtype),
            Dinv=z(rows, dtype), x0=z(rows, x_dtype), x1=z(rows, x_dtype), xp0=z(rows, dtype),
            xp1=z(rows, dtype), w=z(rows, torch.float64 if w_f64 else dtype), r=z(rows, dtype),
            p0=z(rows, dtype), p1=z(rows, dtype), Ap=z(rows, dtype),
            AT=z(rows, dtype) if lat else None, RT=z(rows, dtype) if lat else None,
            stim_mask=z(rows, torch.int32))
        self.models = {}
        for m, make in MODEL_CLASSES.items():
            obj = make(dt, dev, dtype)
            obj.construct_tables()
            # states in blocks of one slice: (slice, state, row in slice)
            st = torch.zeros((cap_slices, len(obj.STATES), 64), dtype=dtype, device=dev)
            self.models[m] = (obj, obj.c_params(), obj.c_tables(), st)
        d = L.ServeDesc()
        P = self.pool
        d.dtype = L.dtype_code(dtype)
        d.x_f64 = 1 if x_dtype == torch.float64 else 0
        d.w_f64 = 1 if w_f64 else 0
        d.cap_slices, d.cap_entries, d.max_sys = cap_slices, int(cap_entries), int(max_sys)
        for k in ("slice_off", "slice_w", "slice_sys", "col", "dcol", "pat", "Mv", "Kv", "Av", "Dinv",
                  "w", "r", "Ap", "AT", "RT", "stim_mask"):
            setattr(d, k, L.ptr(P[k]))
        d.x[0], d.x[1] = L.ptr(P["x0"]), L.ptr(P["x1"])
        d.xp[0], d.xp[1] = L.ptr(P["xp0"]), L.ptr(P["xp1"])
        d.p[0], d.p[1] = L.ptr(P["p0"]), L.ptr(P["p1"])
        for m, (obj, cp, tabs, st) in self.models.items():
            d.states[m] = L.ptr(st)
            d.params[m] = C.cast(C.pointer(cp), C.c_void_p)
            for k, tab in enumerate(tabs):
                d.tab[m][k] = L.ptr(tab)
        d.dt, d.dt_ionic, d.Cm, d.theta = dt, dt, self.Cm, self.theta
        d.a_tol, d.r_tol, d.max_iter = float(a_tol), float(r_tol), int(max_iter)
        d.linear_guess, d.depth = 1, int(depth)
        d.k_history = int(k_history)
        d.graph = 1 if graph else 0
        nb = C.c_size_t(0)
        L.call("tc_serve_workspace_size", cap_slices, int(max_sys), C.byref(nb))
        self._ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        d.ws, d.ws_bytes = L.ptr(self._ws), nb.value
        h = C.c_void_p(0)
        L.call("tc_serve_create", C.byref(h), C.byref(d), L.stream_handle(dev))
        self._h = h
        self._desc = d
        self.lat = lat
        # slices per block of each system's passes: fixed, or by its size
        # from the tuning table
        self.slices_per_block = slices_per_block
        self.live = {}
        self.step_iters = []      # largest n_iter among the live runs, per engine step

    # ----------------------------------------------------------- requests
    def prepare(self, req):
        """Assemble a request's operator on the GPU (as Monodomain.assemble)."""
        dev, dt_ = self.device, self.dtype
        nodes = torch.as_tensor(np.asarray(req.nodes)).to(dtype=dt_, device=dev)
        elems = torch.as_tensor(np.asarray(req.elems)).to(dtype=torch.int32, device=dev)
        regions = torch.as_tensor(np.asarray(req.regions)).to(dtype=torch.int32, device=dev)
        fibres = torch.as_tensor(np.asarray(req.fibres)).to(dtype=dt_, device=dev)
        cond = Conductivity(regions, dtype=dt_)
        for c in req.conductivities:
            cond.add(*c)
        S = P1System(nodes, elems, dev)
        node_map = None
        res = pick_order(S.row_ptr, S.col, S.n, nodes, self.reorder)
        if res is not None:
            # a locality order: internal position i holds mesh node order[i]
            order = torch.as_tensor(res[0], dtype=torch.long, device=dev)
            node_map = torch.empty_like(order)
            node_map[order] = torch.arange(order.shape[0], dtype=torch.long, device=dev)
            elems = node_map[elems.to(torch.long)].to(torch.int32)
            S = P1System(nodes[order], elems, dev)
        M, K = S.assemble(sigma_l=cond.sigma_l, sigma_t=cond.sigma_t, fibres=fibres)
        S.release_assembly_maps()
        S.set_operators(M, K, self.Chi, self.Cm, self.theta, self.dt, dt_)
        n = S.n
        obj = MODEL_CLASSES[req.model](self.dt, dev, dt_)
        obj._alloc_states(n)
        pr = Prepared()
        pr.req, pr.system, pr.n = req, S, n
        pr.node_map = node_map          # mesh node j at internal row node_map[j] (None: same)
        pr.states = obj.states
        pr.V0 = torch.full((n,), float(obj.V_init), dtype=dt_, device=dev)
        pr.mask = torch.zeros(n, dtype=torch.int32, device=dev)
        pr.stim = []
        if len(req.stimuli) > L.POOL_STIM:
            raise ValueError(f"at most {L.POOL_STIM} stimuli per served run")
        for k, s in enumerate(req.stimuli):
            s = tuple(s)
            ids, start, duration, amp = s[:4]
            period = s[4] if len(s) > 4 and s[4] is not None else req.T
            count = s[5] if len(s) > 5 else 1
            ids = torch.as_tensor(np.asarray(ids), dtype=torch.long, device=dev)
            if node_map is not None:
                ids = node_map[ids]
            pr.mask[ids] |= (1 << k)
            pr.stim.append(L.Stimulus(float(start), float(duration), float(period), float(count),
                                      float(amp)))
        pr.nt = int(req.T / self.dt)
        pr.n_slices = (n + 63) // 64
        pr.n_entries = S.padded
        return pr

    def fits(self, pr):
        f = C.c_int(0)
        L.call("tc_serve_fits", self._h, pr.n_slices, pr.n_entries, C.byref(f))
        return bool(f.value)

    def add(self, pr, slices_per_block=None):
        """Admit a prepared run.  ``slices_per_block`` (8..64) overrides the
        engine's choice for this system (engine value, else the tuning
        table by the system's size)."""
        S = pr.system
        y = L.ServeSystem()
        y.n, y.n_entries = pr.n, pr.n_entries
        y.slice_ptr, y.col = L.ptr(S.slice_ptr), L.ptr(S.scol)
        y.dcol = L.ptr(S.dcol) if S.sell.dcol is not None else None
        if S.sell.dcol is not None and S.slice_tag is not None:
            y.pat, y.slice_tag = L.ptr(S.pat), L.ptr(S.slice_tag)
        y.Mv, y.Kv, y.Av, y.Dinv = L.ptr(S.sM), L.ptr(S.sK), L.ptr(S.sA), L.ptr(S.Dinv)
        y.model = pr.req.model
        y.n_states = pr.states.shape[0]
        y.states = L.ptr(pr.states)
        y.V0 = L.ptr(pr.V0)
        y.stim_mask = L.ptr(pr.mask)
        y.n_stim = len(pr.stim)
        for k, s in enumerate(pr.stim):
            y.stim[k] = s
        y.lat = 1 if self.lat else 0
        from . import autotune
        spb = slices_per_block or self.slices_per_block
        y.slices_per_block = int(spb) if spb else autotune.lookup("serve_slices_per_block", pr.n)
        slot = C.c_int(-1)
        L.call("tc_serve_add", self._h, C.byref(y), C.byref(slot), L.stream_handle(self.device))
        self.live[slot.value] = pr
        pr.steps_left = pr.nt
        return slot.value

    def remove(self, slot):
        pr = self.live.pop(slot)
        dev, dt_ = self.device, self.dtype
        V = torch.empty(pr.n, dtype=dt_, device=dev)
        AT = torch.empty(pr.n, dtype=dt_, device=dev) if self.lat else None
        RT = torch.empty(pr.n, dtype=dt_, device=dev) if self.lat else None
        tot, steps = C.c_int64(0), C.c_int64(0)
        L.call("tc_serve_remove", self._h, slot, L.ptr(V), L.ptr(AT), L.ptr(RT), C.byref(tot),
               C.byref(steps), L.stream_handle(dev))
        if pr.node_map is not None:      # back to the mesh's numbering
            V = V[pr.node_map]
            AT = AT[pr.node_map] if AT is not None else None
            RT = RT[pr.node_map] if RT is not None else None
        return dict(u=V, AT=AT, RT=RT, iterations=int(tot.value), steps=int(steps.value),
                    req=pr.req)

    def advance(self, k):
        k = int(k)
        if k <= 0:
            return []
        iters = (C.c_int32 * k)()
        L.call("tc_serve_advance", self._h, k, iters, L.stream_handle(self.device))
        for pr in self.live.values():
            pr.steps_left -= k
        out = list(iters)
        self.step_iters.extend(out)
        return out

    def stats(self):
        """Engine steps so far and how many needed more PCG iterations than
        the pipeline had launched (each such step drains the pipeline)."""
        steps, stalls = C.c_int64(0), C.c_int64(0)
        L.call("tc_serve_stats", self._h, C.byref(steps), C.byref(stalls))
        return dict(steps=int(steps.value), stalls=int(stalls.value))

    def graph_stats(self):
        """(graphs recorded, graph replays) so far (graph mode)."""
        cap, rep = C.c_int64(0), C.c_int64(0)
        L.call("tc_serve_graph_stats", self._h, C.byref(cap), C.byref(rep))
        return cap.value, rep.value

    # -------------------------------------------------------------- loop
    def run_until_done(self, prepared, max_chunk=200):
        """Admit all prepared runs (room permitting) and run them to the end;
        returns their results in order (no arrival clock)."""
        out = [None] * len(prepared)
        queue = collections.deque(enumerate(prepared))
        slots = {}
        while queue or self.live:
            while queue and self.fits(queue[0][1]):
                i, pr = queue.popleft()
                slots[self.add(pr)] = i
            if not self.live:
                raise RuntimeError("a request does not fit the empty pool")
            k = min(max_chunk, min(pr.steps_left for pr in self.live.values()))
            self.advance(k)
            for s in [s for s, pr in self.live.items() if pr.steps_left <= 0]:
                out[slots.pop(s)] = self.remove(s)
        return out

    def serve(self, requests, max_chunk=1000, log=None, admit_chunk=25):
        """Real-time serving loop: requests become visible at their arrival
        time (seconds after the start); returns per-request records with
        arrival, admission and completion times and the latency per
        simulated second, plus summary statistics.

        Visible requests are assembled by one worker thread on its own
        stream while the pool keeps stepping (the library call that steps
        the pool releases the interpreter lock); they are admitted in
        arrival order as their assembly finishes and the pool has room.
        While a request waits for admission the pool steps in chunks of at
        most ``admit_chunk`` steps; otherwise a chunk ends at the next
        retirement, at ``max_chunk`` steps, or near the next arrival (from
        the measured time per step), whichever comes first."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = collections.deque(reqs)
        waiting = collections.deque()       # (request, future of its Prepared)
        slots = {}
        records = []
        spent = {"prepare": 0.0}
        side = torch.cuda.Stream(device=self.device)

        def work(req):
            torch.cuda.set_device(self.device)
            ta = time.perf_counter()
            with torch.cuda.stream(side):
                pr = self.prepare(req)
            side.synchronize()
            spent["prepare"] += time.perf_counter() - ta
            return pr

        worker = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        s_per_step = None                   # measured, smoothed
        n_calls, t_adv = 0, 0.0             # advance calls and the time spent in them
        st0 = self.stats()
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        try:
            while pending or waiting or self.live:
                now = clock()
                while pending and pending[0].arrival <= now:
                    req = pending.popleft()
                    waiting.append((req, worker.submit(work, req)))
                while waiting:
                    req, fut = waiting[0]
                    if self.live and not fut.done():
                        break
                    pr = fut.result()          # waits only when nothing is stepping
                    if not self.fits(pr):
                        if not self.live:
                            raise RuntimeError("a request does not fit the empty pool")
                        break
                    waiting.popleft()
                    slot = self.add(pr)
                    slots[slot] = dict(req=req, admitted=clock())
                if not self.live:
                    if pending:
                        time.sleep(max(0.0, pending[0].arrival - clock()))
                    continue
                k = min(max_chunk, min(pr.steps_left for pr in self.live.values()))
                if waiting:
                    k = min(k, admit_chunk)
                elif pending and s_per_step:
                    k = min(k, max(admit_chunk, int((pending[0].arrival - clock()) / s_per_step) + 1))
                ta = time.perf_counter()
                self.advance(k)
                el = time.perf_counter() - ta
                n_calls += 1
                t_adv += el
                dt_step = el / k
                s_per_step = dt_step if s_per_step is None else 0.7 * s_per_step + 0.3 * dt_step
                for s in [s for s, pr in self.live.items() if pr.steps_left <= 0]:
                    res = self.remove(s)
                    info = slots.pop(s)
                    done = clock()
                    req = info["req"]
                    rec = dict(tag=req.tag, n=int(np.asarray(req.nodes).shape[0]), T=req.T,
                               model=req.model, arrival=req.arrival, admitted=info["admitted"],
                               done=done, latency=done - req.arrival,
                               latency_per_sim_s=(done - req.arrival) / (req.T / 1000.0),
                               iterations=res["iterations"], result=res)
                    records.append(rec)
                    if log:
                        log(rec)
        finally:
            worker.shutdown(wait=True)
        wall = clock()
        lat = np.array([r["latency_per_sim_s"] for r in records])
        st1 = self.stats()
        summary = dict(requests=len(records), wall_s=wall, prepare_s=spent["prepare"],
                       steps=st1["steps"] - st0["steps"], stalls=st1["stalls"] - st0["stalls"],
                       sims_per_hour=len(records) * 3600.0 / wall,
                       advance_calls=n_calls, advance_s=t_adv,
                       latency_per_sim_s_p50=float(np.percentile(lat, 50)),
                       latency_per_sim_s_p99=float(np.percentile(lat, 99)))
        return records, summary

    def close(self):
        if getattr(self, "_h", None):
            L.lib.tc_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

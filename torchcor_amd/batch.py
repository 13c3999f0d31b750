"""Many monodomain runs on one shared mesh, advanced together on one GPU.

The reference runs one ``Monodomain`` at a time (torchcor/simulator/
monodomain.py:141-226).  ``SharedMeshBatch`` runs up to 64 of them on the
same mesh in one time loop; instances differ in conductivity scale, fibre
rotation about the z axis, stimulus sites and timing, and ionic model.

Operator sharing.  With sigma = s [sigma_t I + (sigma_l - sigma_t) f f^T] and
the fibre f rotated by delta about z, the rotated tensor splits exactly into
parts that transform as a scalar, a vector and a rank-2 tensor under that
rotation:
    sigma_b = s_b (S0 + cos d S1c + sin d S1s + cos 2d S2c + sin 2d S2s),
so every instance's stiffness matrix is a combination of NK shared component
matrices (NK = 3 when no fibre leaves the xy plane, 1 when no instance is
rotated).  The device keeps M and the components once and forms
A_b = Cm M + theta dt K_b per instance inside the PCG passes: the batch is a
multi-right-hand-side PCG with per-instance scalars and stopping tests.
Assembly is float64 without the float32 rounding of element contributions
that the single-run path reproduces (that rounding does not commute with the
split), so a batched instance equals a single run assembled in float64.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from .core.conductivity import Conductivity
from .core.mesh import MeshReader
from .core.stimulation import MAX_STIMULI
from .core.reorder import pick_order
from .core.system import P1System
from .ionic.base import default_device

MAX_INSTANCES = 64


def sigma_components(sigma_l, sigma_t, fibres):
    """(S0, S1c, S1s, S2c, S2s), each (E,3,3) float64, of the conductivity
    tensor sigma_t I + (sigma_l - sigma_t) f f^T under rotations about z."""
    sl = torch.as_tensor(sigma_l, dtype=torch.float64)
    st = torch.as_tensor(sigma_t, dtype=torch.float64)
    f = torch.as_tensor(fibres, dtype=torch.float64)
    d = (sl - st).view(-1, 1, 1)
    E = f.shape[0]
    fx, fy, fz = f[:, 0], f[:, 1], f[:, 2]
    Fxx, Fyy, Fzz, Fxy, Fxz, Fyz = fx * fx, fy * fy, fz * fz, fx * fy, fx * fz, fy * fz
    z = lambda: torch.zeros((E, 3, 3), dtype=torch.float64, device=f.device)  # noqa: E731
    S0 = z()
    S0[:, 0, 0] = S0[:, 1, 1] = 0.5 * (Fxx + Fyy)
    S0[:, 2, 2] = Fzz
    S0 = st.view(-1, 1, 1) * torch.eye(3, dtype=torch.float64, device=f.device) + d * S0

    def sym(a, b, v, out):
        out[:, a, b] = out[:, b, a] = v

    S1c, S1s, S2c, S2s = z(), z(), z(), z()
    sym(0, 2, Fxz, S1c)
    sym(1, 2, Fyz, S1c)
    sym(0, 2, -Fyz, S1s)
    sym(1, 2, Fxz, S1s)
    D, Exy = 0.5 * (Fxx - Fyy), Fxy
    S2c[:, 0, 0], S2c[:, 1, 1] = D, -D
    sym(0, 1, Exy, S2c)
    S2s[:, 0, 0], S2s[:, 1, 1] = -Exy, Exy
    sym(0, 1, D, S2s)
    return S0, d * S1c, d * S1s, d * S2c, d * S2s


def rotate_z(fibres, deg):
    """Fibres rotated by ``deg`` degrees about the z axis."""
    f = np.asarray(fibres, dtype=np.float64)
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    return f @ R.T


class _Instance:
    def __init__(self, model, scale, rot, stimuli):
        self.model = model
        self.scale = float(scale)
        self.rot = float(rot)
        self.stimuli = stimuli


class SharedMeshBatch:
    """Up to 64 monodomain runs on one mesh (same T, dt, dtype).

    Set-up mirrors ``Monodomain``: ``load_mesh``/``load_arrays``,
    ``add_conductivity`` (the base conductivities, shared), then
    ``add_instance(ionic_model, sigma_scale, fibre_rotation_deg, stimuli)``
    once per run, where ``stimuli`` lists (node_ids, start, duration,
    intensity[, period[, count]]) like ``Monodomain.add_stimulus``.
    ``solve`` returns the per-instance sums of the PCG iteration counts;
    results are (B, n) tensors in instance order: ``u``,
    ``activation_time``, ``repolarization_time``.  A mesh whose node
    numbering scatters neighbours runs in a locality order (``reorder`` as
    for ``Monodomain``: "auto", "levels", "rcm" or None); stimulus node ids
    and results stay in the mesh's numbering.
    """
    Chi = 140
    Cm = 0.01
    theta = 0.5

    def __init__(self, T, dt, device=None, dtype=None, reorder="auto"):
        if reorder not in ("auto", "levels", "rcm", None):
            raise ValueError("reorder must be 'auto', 'levels', 'rcm' or None")
        self.reorder = reorder
        self.node_order = self.node_map = None
        self.device = torch.device(device) if device is not None else default_device()
        self.dtype = dtype if dtype is not None else torch.float32
        self.T = T
        self.dt = dt
        self.nt = int(T / dt)
        self.instances = []
        self.system = None
        self._h = None

    # ------------------------------------------------------------ set-up
    def load_mesh(self, path, unit_conversion=1000):
        nodes, elems, regions, fibres = MeshReader(path).read(unit_conversion=unit_conversion)
        self.load_arrays(nodes, elems, regions, fibres)

    def load_arrays(self, nodes, elems, regions, fibres):
        dev = self.device
        self.n_nodes = int(np.asarray(nodes).shape[0])
        self.nodes = torch.as_tensor(np.asarray(nodes)).to(dtype=self.dtype, device=dev)
        self.elems = torch.as_tensor(np.asarray(elems)).to(dtype=torch.int32, device=dev)
        self.regions = torch.as_tensor(np.asarray(regions)).to(dtype=torch.int32, device=dev)
        self.fibres = torch.as_tensor(np.asarray(fibres)).to(dtype=self.dtype, device=dev)
        self.conductivity = Conductivity(self.regions, dtype=self.dtype)
        self.system = None

    def add_conductivity(self, region_ids, il, it, el=None, et=None):
        self.conductivity.add(region_ids, il, it, el, et)

    def add_instance(self, ionic_model, sigma_scale=1.0, fibre_rotation_deg=0.0, stimuli=()):
        """Add one run; returns its index.  ``ionic_model`` is a model object
        (its class and parameters pick the run's ionic model; instances with
        equal class and parameters share one state block)."""
        if len(self.instances) >= MAX_INSTANCES:
            raise ValueError(f"at most {MAX_INSTANCES} instances per batch")
        st = []
        for s in stimuli:
            s = tuple(s)
            nodes, start, duration, intensity = s[:4]
            period = s[4] if len(s) > 4 and s[4] is not None else self.T
            count = s[5] if len(s) > 5 else 1
            st.append((np.asarray(nodes, dtype=np.int64), float(start), float(duration),
                       float(intensity), float(period), float(count)))
        self.instances.append(_Instance(ionic_model, sigma_scale, fibre_rotation_deg, st))
        return len(self.instances) - 1

    # ---------------------------------------------------------- assembly
    def assemble(self):
        """Shared pattern, M and the stiffness components on the GPU."""
        if not self.instances:
            raise ValueError("no instances")
        dev = self.device
        S = P1System(self.nodes, self.elems, dev)
        self.node_order = self.node_map = None
        res = pick_order(S.row_ptr, S.col, S.n, self.nodes, self.reorder)
        if res is not None:
            # internal position i holds mesh node node_order[i]
            self.node_order = torch.as_tensor(res[0], dtype=torch.long, device=dev)
            self.node_map = torch.empty_like(self.node_order)
            self.node_map[self.node_order] = torch.arange(S.n, dtype=torch.long, device=dev)
            elems = self.node_map[self.elems.to(torch.long)].to(torch.int32)
            S = P1System(self.nodes[self.node_order], elems, dev)
        sl, st = self.conductivity.sigma_l, self.conductivity.sigma_t
        comps = sigma_components(sl, st, self.fibres)
        rots = [i.rot for i in self.instances]
        if all(r == 0.0 for r in rots):
            keep = [None]           # the full tensor, unrotated
        elif bool((self.fibres[:, 2] == 0).all()):
            keep = [0, 3, 4]
        else:
            keep = [0, 1, 2, 3, 4]
        Ks = []
        M = None
        for k in keep:
            sig = comps[0] + comps[3] + comps[1] if k is None else comps[k]
            Mk, Kk = S.assemble(sigma=sig, round_f32=False)
            M = Mk if M is None else M
            Ks.append(Kk)
        S.release_assembly_maps()
        self.system = S
        self.components = keep
        self.Mv = M.to(self.dtype)
        self.Kv = [K.to(self.dtype) for K in Ks]
        self.nnz = S.nnz

    def _coef(self, lanes, order):
        """coef[k, lane] = scale * g_k(rotation) / Chi for the kept components."""
        g = {None: lambda r: 1.0, 0: lambda r: 1.0,
             1: lambda r: math.cos(math.radians(r)), 2: lambda r: math.sin(math.radians(r)),
             3: lambda r: math.cos(math.radians(2 * r)), 4: lambda r: math.sin(math.radians(2 * r))}
        c = np.zeros((len(self.components), lanes))
        for lane, idx in enumerate(order):
            inst = self.instances[idx]
            for k, comp in enumerate(self.components):
                c[k, lane] = inst.scale * g[comp](inst.rot) / self.Chi
        return torch.tensor(c, dtype=torch.float64, device=self.device)

    # ------------------------------------------------------------- run
    def start(self, a_tol, r_tol, max_iter, calculate_AT_RT=True, linear_guess=True, depth=4,
              x_f64=True, precise_rhs=True, max_blocks=None, graph=False):
        """Set up the run.  ``max_blocks`` caps the blocks of the PCG passes
        (default: the tuning table's value for n x lanes).  ``graph``
        replays each step's RHS and PCG launches as a captured HIP graph
        (same results)."""
        self.assemble()
        dev, dt_ = self.device, self.dtype
        n, B = self.n_nodes, len(self.instances)
        lanes = 1 << (B - 1).bit_length()
        # model groups: equal class and parameters -> one contiguous lane range
        groups = {}
        for idx, inst in enumerate(self.instances):
            m = inst.model
            key = (type(m), tuple(sorted(m._param_values().items())), float(m.dt))
            groups.setdefault(key, []).append(idx)
        if len(groups) > 8:
            raise ValueError("at most 8 distinct ionic models per batch")
        order = [i for members in groups.values() for i in members]
        self.lane_of = np.empty(B, dtype=np.int64)
        self.lane_of[order] = np.arange(B)
        x_dtype = torch.float64 if (x_f64 or dt_ == torch.float64) else dt_
        w_f64 = bool(precise_rhs and x_dtype == torch.float64)
        mk = lambda dtype: torch.zeros((n, lanes), dtype=dtype, device=dev)  # noqa: E731
        X0 = mk(x_dtype)
        self._groups = []
        gdesc = []
        b0 = 0
        for members in groups.values():
            model = self.instances[members[0]].model
            V0 = model.initialize(n * len(members))
            X0[:, b0:b0 + len(members)] = V0[0].to(x_dtype)
            self._groups.append((model, model.c_params(), model.c_tables()))
            gdesc.append((model, b0, len(members)))
            b0 += len(members)
        self._X = [X0, mk(x_dtype)]
        self._XP = [X0.to(dt_, copy=True), mk(dt_)]
        self._w = mk(torch.float64 if w_f64 else dt_)
        self.Dinv = mk(dt_)
        self.activation_time = self.repolarization_time = None
        if calculate_AT_RT:
            self.activation_time = torch.full((n, lanes), -1.0, dtype=dt_, device=dev)
            self.repolarization_time = torch.full((n, lanes), -1.0, dtype=dt_, device=dev)
        # stimuli: one slot per distinct (start, duration, intensity, period, count)
        slots = []
        mask = torch.zeros((n, lanes), dtype=torch.int32, device=dev)
        for idx, inst in enumerate(self.instances):
            for nodes, *timing in inst.stimuli:
                key = tuple(timing)
                if key not in slots:
                    if len(slots) >= MAX_STIMULI:
                        raise ValueError(f"at most {MAX_STIMULI} distinct stimulus timings")
                    slots.append(key)
                k = slots.index(key)
                bit = 1 << k if k < 31 else -(1 << 31)
                nd = torch.as_tensor(nodes, device=dev)
                if self.node_map is not None:
                    nd = self.node_map[nd]
                mask[nd, int(self.lane_of[idx])] |= bit
        self._mask = mask
        self._coefs = self._coef(lanes, order)
        d = L.BatchDesc()
        d.n, d.nnz, d.B, d.lanes, d.NK = n, self.nnz, B, lanes, len(self.components)
        d.dtype = L.dtype_code(dt_)
        d.x_f64 = 1 if x_dtype == torch.float64 else 0
        d.w_f64 = 1 if w_f64 else 0
        d.row_ptr, d.col = L.ptr(self.system.row_ptr), L.ptr(self.system.col)
        d.Mv = L.ptr(self.Mv)
        for k, K in enumerate(self.Kv):
            d.Kv[k] = L.ptr(K)
        d.coef = L.ptr(self._coefs)
        d.dt, d.dt_ionic, d.Cm, d.theta = self.dt, float(self.instances[0].model.dt), self.Cm, self.theta
        for k in range(2):
            d.x[k] = L.ptr(self._X[k])
            d.xp[k] = L.ptr(self._XP[k])
        d.w, d.Dinv = L.ptr(self._w), L.ptr(self.Dinv)
        d.AT = L.ptr(self.activation_time)
        d.RT = L.ptr(self.repolarization_time)
        d.stim_mask = L.ptr(mask) if slots else None
        d.n_stim = len(slots)
        for k, (start, duration, intensity, period, count) in enumerate(slots):
            d.stim[k] = L.Stimulus(start, duration, period, count, intensity)
        d.n_groups = len(gdesc)
        for g, (model, gb0, cnt) in enumerate(gdesc):
            _, cp, tabs = self._groups[g]
            d.groups[g].model = model.MODEL
            d.groups[g].params = C.cast(C.pointer(cp), C.c_void_p)
            d.groups[g].states = L.ptr(model.states)
            d.groups[g].state_stride = model.states.stride(0)
            for k, tab in enumerate(tabs):
                d.groups[g].tab[k] = L.ptr(tab)
            d.groups[g].b0, d.groups[g].count = gb0, cnt
        for inst in self.instances:
            if float(inst.model.dt) != d.dt_ionic:
                raise ValueError("all ionic models of a batch need the same dt")
        d.a_tol, d.r_tol, d.max_iter = float(a_tol), float(r_tol), int(max_iter)
        d.linear_guess = 1 if linear_guess else 0
        d.depth = int(depth)
        from . import autotune
        d.max_blocks = int(max_blocks) if max_blocks else \
            autotune.lookup("batch_max_blocks", n * lanes)
        self.max_blocks = d.max_blocks
        d.graph = 1 if graph else 0
        nb = C.c_size_t(0)
        L.call("tc_batch_workspace_size", n, lanes, d.dtype, C.byref(nb))
        self._ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        d.ws, d.ws_bytes = L.ptr(self._ws), nb.value
        self._free()
        h = C.c_void_p(0)
        L.call("tc_batch_create", C.byref(h), C.byref(d), L.stream_handle(dev))
        self._h = h
        self._desc = d
        self.iterations = []
        self.lanes = lanes

    def graph_stats(self):
        """(graphs recorded, graph replays) of the current run."""
        cap, rep = C.c_int64(0), C.c_int64(0)
        L.call("tc_batch_graph_stats", self._h, C.byref(cap), C.byref(rep))
        return cap.value, rep.value

    def advance(self, k):
        k = int(k)
        if k <= 0:
            return
        iters = (C.c_int32 * k)()
        L.call("tc_batch_advance", self._h, k, iters, L.stream_handle(self.device))
        self.iterations.extend(iters)

    def finish(self):
        B = len(self.instances)
        tot = (C.c_int64 * B)()
        L.call("tc_batch_finish", self._h, tot, L.stream_handle(self.device))
        t, steps, xi = C.c_double(), C.c_int64(), C.c_int()
        L.call("tc_batch_status", self._h, C.byref(t), C.byref(steps), C.byref(xi))
        lanes = torch.as_tensor(self.lane_of, device=self.device)
        # (n, lanes) in the internal order -> (B, n) in the mesh's numbering
        rows = self.node_map if self.node_map is not None else slice(None)
        self.u = self._X[xi.value][rows][:, lanes].T.to(self.dtype).contiguous()
        if self.activation_time is not None:
            self.activation_time = self.activation_time[rows][:, lanes].T.contiguous()
            self.repolarization_time = self.repolarization_time[rows][:, lanes].T.contiguous()
        self.instance_iterations = [int(tot[self.lane_of[b]]) for b in range(B)]
        self.t = t.value
        return self.instance_iterations

    def solve(self, a_tol, r_tol, max_iter, calculate_AT_RT=True, linear_guess=True, nsteps=None,
              **kw):
        self.start(a_tol, r_tol, max_iter, calculate_AT_RT, linear_guess, **kw)
        self.advance(self.nt if nsteps is None else min(int(nsteps), self.nt))
        return self.finish()

    def _free(self):
        if getattr(self, "_h", None):
            L.lib.tc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass

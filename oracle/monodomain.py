"""Monodomain time loop of the reference, restated in float64 numpy.

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np

from . import fem, pcg


def stimulus_active(t, start, duration, period, count):
    """Stimulus window test of stimulation.py:51-63, with ``t`` a Python double."""
    p_t = t % period
    t_t = t / period
    return t_t <= count and start <= p_t <= start + duration


class Monodomain:
    """Float64 restatement of ``Monodomain`` (monodomain.py:18-226).

    Mesh arrays are passed in directly (coordinates already in mm) instead of
    being read from files; ``add_stimulus`` takes the node indices of the
    region.  ``solve`` returns a dict with AT, RT, the Vm frames and the
    per-step PCG iteration counts.
    """
    Chi = 140
    Cm = 0.01
    theta = 0.5
    #: the reference rounds each element contribution to float32
    #: (assemble.py:233-243); False assembles in plain float64
    round_f32 = True

    def __init__(self, ionic_model, T, dt):
        self.ionic = ionic_model
        self.T = T
        self.dt = dt
        self.nt = int(T / dt)
        self.stimuli = []
        self.cond_calls = []

    def load_arrays(self, nodes, elems, regions, fibres):
        self.nodes = np.asarray(nodes, dtype=np.float64)
        self.elems = np.asarray(elems, dtype=np.int64)
        self.regions = np.asarray(regions)
        self.fibres = np.asarray(fibres, dtype=np.float64)
        self.n = self.nodes.shape[0]

    def add_conductivity(self, region_ids, il, it, el=None, et=None):
        self.cond_calls.append((list(region_ids), il, it, el, et))

    def add_stimulus(self, node_ids, start, duration, intensity, period=None, count=1):
        if period is None:
            period = self.T
        mask = np.zeros(self.n)
        mask[np.asarray(node_ids, dtype=np.int64)] = 1.0
        self.stimuli.append((mask * intensity, start, duration, period, count))

    def stim(self, t):
        """Summed stimulus at time t (stimulation.py:51-63); zeros when none is on."""
        on = [m for (m, s, d, p, c) in self.stimuli if stimulus_active(t, s, d, p, c)]
        if not on:
            return np.zeros(self.n)
        return np.sum(np.stack(on), axis=0)

    def assemble(self):
        """M, K, A on one CSR pattern plus the Jacobi preconditioner (monodomain.py:83-106)."""
        sl, st = fem.conductivity_fields(self.regions, self.cond_calls)
        sig = fem.sigma_tensor(sl, st, self.fibres)
        coords = self.nodes[self.elems]
        if self.elems.shape[1] == 3:
            Me, Ke = fem.tri_local(coords, sig)
        else:
            Me, Ke = fem.tet_local(coords, sig)
        rp, col, kv = fem.assemble_csr(self.elems, Ke, self.n, round_f32=self.round_f32)
        rp2, col2, mv = fem.assemble_csr(self.elems, Me, self.n, round_f32=self.round_f32)
        assert np.array_equal(rp, rp2) and np.array_equal(col, col2)
        self.row_ptr, self.col = rp, col
        self.K = kv / self.Chi
        self.M = mv
        self.A = self.M * self.Cm + self.K * self.dt * self.theta
        self.pc = pcg.jacobi(rp, col, self.A)

    def spmv(self, vals, x):
        return fem.spmv(self.row_ptr, self.col, vals, x)

    def phie_recovery(self, Vm, a_tol=1e-5, r_tol=1e-5, max_iter=100):
        """phi_e of every Vm frame (monodomain.py:254-286), float64 run:
        b = -K_i V; K_ie phi = b by one PCG object carried over the frames
        with the linear guess; phi -= mean(phi) applied to the solver's own
        iterate (in a float64 run the reference's returned solution is that
        iterate).  Returns (frames (F, n), n_iter per frame)."""
        fields = fem.bidomain_fields(self.regions, self.cond_calls)
        if fields is None:
            raise ValueError("phie_recovery needs el and et")
        il, it, el, et = fields
        si = fem.sigma_tensor(il, it, self.fibres)
        se = fem.sigma_tensor(el, et, self.fibres)
        coords = self.nodes[self.elems]
        local = fem.tri_local if self.elems.shape[1] == 3 else fem.tet_local
        _, k_ie = local(coords, si + se)
        _, k_i = local(coords, si)
        rp, col, kie = fem.assemble_csr(self.elems, k_ie, self.n, round_f32=self.round_f32)
        _, _, ki = fem.assemble_csr(self.elems, k_i, self.n, round_f32=self.round_f32)
        cg = pcg.PCG(rp, col, kie, pcg.jacobi(rp, col, kie))
        cg.initialize(np.zeros(self.n), linear_guess=True)
        out, iters = [], []
        for V in np.asarray(Vm, dtype=np.float64):
            b = -fem.spmv(rp, col, ki, V)
            x, it_ = cg.solve(b, a_tol, r_tol, max_iter)
            cg.x = x - x.mean()
            out.append(cg.x.copy())
            iters.append(it_)
        return np.stack(out), np.array(iters)

    def solve(self, a_tol, r_tol, max_iter, calculate_AT_RT=True, linear_guess=True,
              snapshot_interval=1, nsteps=None):
        self.assemble()
        u = self.ionic.initialize(self.n)
        frames = [u.copy()]
        cg = pcg.PCG(self.row_ptr, self.col, self.A, self.pc)
        cg.initialize(u, linear_guess=linear_guess, x_prev_dtype=np.float64)
        ts_per_frame = int(snapshot_interval / self.dt)
        AT = -np.ones(self.n)
        RT = -np.ones(self.n)
        iters = []
        t = 0
        nt = self.nt if nsteps is None else nsteps
        for step in range(1, nt + 1):
            t += self.dt
            du = self.ionic.differentiate(u) / 100
            b = u * self.Cm + self.dt * du
            b = b + self.dt * (self.stim(t) / 100)
            b = self.spmv(self.M, b)
            b = b - self.spmv((1 - self.theta) * self.dt * self.K, u)
            u, n_iter = cg.solve(b, a_tol, r_tol, max_iter)
            u = u.copy()
            iters.append(n_iter)
            if calculate_AT_RT:
                AT[(u > 0) & (AT == -1)] = t
                RT[(AT > 0) & (RT == -1) & (u < -70)] = t
            if step % ts_per_frame == 0:
                frames.append(u.copy())
        return dict(AT=AT, RT=RT, Vm=np.stack(frames), iters=np.array(iters), u=u, t=t)

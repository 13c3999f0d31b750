"""The monodomain time step on torch CPU operations — MEASUREMENT
INFRASTRUCTURE ONLY (the CPU baseline of bench.py; checked against the numpy
oracle by tests/test_oracle.py).

The same algorithm as ``oracle.monodomain`` / ``oracle.ionic`` (which cite
the reference lines), written with torch CPU tensor operations in the
reference's own order and storage types (torchcor/simulator/
monodomain.py:109-139, torchcor/core/solver.py:35-76,
torchcor/ionic/TenTusscherPanfilov.py:188-454): M, K and A as torch sparse
CSR in the working dtype, the ionic states and tables in the working dtype,
the Jacobi-PCG in float64 with x_prev in the working dtype.  torch runs its
CPU kernels on ``torch.get_num_threads()`` threads.  The operators and the
lookup tables come from the numpy oracle (assembly is not timed).
"""
import math

import numpy as np
import torch

from . import ionic as oion
from . import monodomain as omono


def _interp(X, table, T, dtype):
    """Linear table lookup with a float32 lower_pos (the numpy oracle's
    ``interpolate``)."""
    mn, mx, res, step, mxi = T
    X = X.clamp(mn, mx)
    idx = ((X - mn) * step).to(torch.int64).clamp(0, mxi - 1)
    lower_pos = idx.to(torch.float32) * np.float32(res) + np.float32(mn)
    w = ((X - lower_pos.to(dtype)) / res).unsqueeze(1)
    return (1 - w) * table[idx] + w * table[idx + 1]


class TenTusscherPanfilov:
    """ten Tusscher-Panfilov on torch CPU tensors of ``dtype``."""
    name = "TenTusscherPanfilov"

    def __init__(self, cell_type, dt, dtype=torch.float32):
        o = oion.TenTusscherPanfilov(cell_type, dt)
        o.construct_tables()
        self.o = o
        self.dt = dt
        self.dtype = dtype
        self.V_tab = torch.from_numpy(o.V_tab).to(dtype)
        self.CaSS_tab = torch.from_numpy(o.CaSS_tab).to(dtype)
        self.VEk_tab = torch.from_numpy(o.VEk_tab).to(dtype)
        self.c = {k: i for i, k in enumerate(o.V_COLS)}

    def initialize(self, n):
        self.s = {k: torch.full((n,), v, dtype=self.dtype) for k, v in self.o.INIT.items()}
        return torch.full((n,), self.o.V_init, dtype=self.dtype)

    def differentiate(self, V):
        o, s, c, dt = self.o, self.s, self.c, self.dt
        RTONF = (o.Rconst * o.T) / o.Fconst
        inverseVcF2 = 1. / ((2. * o.Vc) * o.Fconst)
        inverseVssF2 = 1. / ((2. * o.Vss) * o.Fconst)
        inverseVcF = 1. / (o.Vc * o.Fconst)
        pmf_INaK = o.knak * (o.Ko / (o.Ko + o.KmK))
        sqrt_Ko = math.sqrt(o.Ko / 5.4)
        F_RT = 1. / RTONF
        invVcF_Cm = inverseVcF * o.CAPACITANCE
        Vr = _interp(V, self.V_tab, o.V_T, self.dtype)
        Cr = _interp(s["CaSS"], self.CaSS_tab, o.CaSS_T, self.dtype)
        Cai, CaSS, CaSR, Ki, Nai = s["Cai"], s["CaSS"], s["CaSR"], s["Ki"], s["Nai"]
        Eca = (0.5 * RTONF) * torch.log(o.Cao / Cai)
        Ek = RTONF * torch.log(o.Ko / Ki)
        Eks = RTONF * torch.log((o.Ko + o.pKNa * o.Nao) / (Ki + o.pKNa * Nai))
        Ena = RTONF * torch.log(o.Nao / Nai)
        IpCa = (o.GpCa * Cai) / (o.KpCa + Cai)
        ghk = torch.where(V == 15., torch.full_like(V, 0.5 * F_RT),
                          (V - 15.) / torch.expm1((2. * (V - 15.)) * F_RT))
        a1 = o.GCaL * o.Fconst * F_RT * 4. * ghk
        ICaL_A = a1 * Vr[:, c["a2"]]
        ICaL_B = a1 * o.Cao
        IKr = o.GKr * sqrt_Ko * s["Xr1"] * s["Xr2"] * (V - Ek)
        IKs = o.GKs * s["Xs"] * s["Xs"] * (V - Eks)
        INa = o.GNa * s["M"] * s["M"] * s["M"] * s["H"] * s["J"] * (V - Ena)
        INaK = pmf_INaK * (Nai / (Nai + o.KmNa)) * Vr[:, c["rec_iNaK"]]
        IbCa = o.GbCa * (V - Eca)
        IbNa = o.GbNa * (V - Ena)
        IpK = o.GpK * Vr[:, c["rec_ipK"]] * (V - Ek)
        Ito = o.Gto * s["R"] * s["S"] * (V - Ek)
        Kr = _interp(V - Ek, self.VEk_tab, o.VEk_T, self.dtype)
        ICaL = s["D"] * s["F"] * s["F2"] * s["FCaSS"] * (ICaL_A * CaSS - ICaL_B)
        INaCa = Vr[:, c["INaCa_A"]] * Nai * Nai * Nai - Vr[:, c["INaCa_B"]] * Cai
        IK1 = o.GK1 * Kr[:, 0] * (V - Ek)
        Iion = IKr + IKs + IK1 + Ito + INa + IbNa + ICaL + IbCa + INaK + INaCa + IpCa + IpK
        Ileak = o.Vleak * (CaSR - Cai)
        Iup = o.Vmaxup / (1. + (o.Kup * o.Kup) / (Cai * Cai))
        Ixfer = o.Vxfer * (CaSS - Cai)
        diff_Ki = -((IK1 + Ito + IKr + IKs - 2. * INaK) + IpK) * invVcF_Cm
        diff_Nai = -((INa + IbNa + 3. * INaK) + 3. * INaCa) * invVcF_Cm
        kCaSR = o.maxsr - (o.maxsr - o.minsr) / (1. + (o.EC / CaSR) * (o.EC / CaSR))
        diff_Cai = ((Ixfer - ((IbCa + IpCa - 2. * INaCa) * inverseVcF2) * o.CAPACITANCE)
                    - (Iup - Ileak) * (o.Vsr / o.Vc)) / \
            (1. + ((o.Bufc * o.Kbufc) / (o.Kbufc + Cai)) / (o.Kbufc + Cai))
        diff_R_ = o.k4 * (1. - s["R_"]) - o.k2_ * kCaSR * CaSS * s["R_"]
        k1 = o.k1_ / kCaSR
        O = (k1 * Cr[:, 0] * s["R_"]) / (o.k3 + k1 * Cr[:, 0])
        Irel = o.Vrel * O * (CaSR - CaSS)
        diff_CaSR = (Iup - Irel - Ileak) / \
            (1. + ((o.Bufsr * o.Kbufsr) / (CaSR + o.Kbufsr)) / (CaSR + o.Kbufsr))
        diff_CaSS = ((-Ixfer * (o.Vc / o.Vss) + Irel * (o.Vsr / o.Vss))
                     + (-ICaL * inverseVssF2) * o.CAPACITANCE) / \
            (1. + ((o.Bufss * o.Kbufss) / (CaSS + o.Kbufss)) / (CaSS + o.Kbufss))
        new = {"CaSR": CaSR + diff_CaSR * dt, "CaSS": CaSS + diff_CaSS * dt,
               "Cai": Cai + diff_Cai * dt, "Ki": Ki + diff_Ki * dt, "Nai": Nai + diff_Nai * dt,
               "R_": s["R_"] + diff_R_ * dt, "FCaSS": Cr[:, 1] + Cr[:, 2] * s["FCaSS"]}
        for g in ("D", "F", "F2", "H", "J", "M", "R", "S", "Xr1", "Xr2", "Xs"):
            new[g] = Vr[:, c[g + "_A"]] + Vr[:, c[g + "_B"]] * s[g]
        self.s = new
        return -Iion


class Monodomain:
    """The time loop on torch CPU tensors.  ``om`` holds the operators and
    stimuli: an assembled ``oracle.monodomain.Monodomain``, or any object
    with its fields (row_ptr, col, n, M, K = K/Chi, Cm, dt, theta, stimuli as
    (mask * intensity, start, duration, period, count)); ``ionic`` is a
    torch-CPU model of this module."""

    def __init__(self, om, ionic, dtype=torch.float32):
        self.om, self.ionic, self.dtype = om, ionic, dtype
        rp = torch.from_numpy(om.row_ptr.astype(np.int64))
        col = torch.from_numpy(om.col.astype(np.int64))
        n = om.n
        csr = lambda v, d: torch.sparse_csr_tensor(rp, col, torch.as_tensor(v).to(d),  # noqa: E731
                                                   size=(n, n))
        self.M, self.K = csr(om.M, dtype), csr(om.K, dtype)
        A = csr(om.M * om.Cm + om.K * om.dt * om.theta, dtype)
        self.A64 = torch.sparse_csr_tensor(rp, col, A.values().to(torch.float64), size=(n, n))
        # Jacobi scaling 1/diag(A) in A's dtype (preconditioner.py:9-27)
        rows = torch.repeat_interleave(torch.arange(n), rp.diff())
        on = rows == col
        d = torch.zeros(n, dtype=dtype)
        d[rows[on]] = A.values()[on]
        self.pc = (1.0 / d).to(torch.float64)
        self.stims = [(torch.as_tensor(m).to(dtype), s, d, p, c) for (m, s, d, p, c) in om.stimuli]
        self.n = n

    def initialize(self):
        u = self.ionic.initialize(self.n)
        self.x = u.to(torch.float64).clone()
        self.x_prev = u.clone()
        return u

    def _cg(self, b, a_tol, r_tol, max_iter):
        b = b.to(torch.float64)
        xc = self.x.clone()
        self.x.mul_(2).sub_(self.x_prev)
        self.x_prev.copy_(xc)
        A, pc = self.A64, self.pc
        r = b - A @ self.x
        z = pc * r
        p = z.clone()
        b_norm = torch.linalg.vector_norm(pc * b)
        n_iter = 0
        for _ in range(max_iter):
            Ap = A @ p
            rz = torch.dot(r, z)
            alpha = rz / torch.dot(p, Ap)
            self.x.add_(alpha * p)
            r.sub_(alpha * Ap)
            z = pc * r
            zn = torch.linalg.vector_norm(z).item()
            if zn < a_tol or zn / b_norm.item() < r_tol:
                break
            beta = torch.dot(r, z) / rz
            p.mul_(beta).add_(z)
            n_iter += 1
        return self.x.to(self.dtype), n_iter

    def step(self, u, t, a_tol=1e-5, r_tol=1e-5, max_iter=100):
        om = self.om
        du = self.ionic.differentiate(u) / 100
        b = u * om.Cm + om.dt * du
        on = [m for (m, s, d, p, c) in self.stims if omono.stimulus_active(t, s, d, p, c)]
        if on:
            b = b + om.dt * (torch.stack(on).sum(0) / 100)
        b = self.M @ b
        b = b - (1 - om.theta) * om.dt * (self.K @ u)
        return self._cg(b, a_tol, r_tol, max_iter)


def cpu_model_name():
    """The host CPU's model name (from /proc/cpuinfo), or ''."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def physical_cores():
    """Distinct (physical id, core id) pairs in /proc/cpuinfo, or 0."""
    try:
        pairs, phys = set(), None
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    phys = line.split(":", 1)[1].strip()
                elif line.startswith("core id"):
                    pairs.add((phys, line.split(":", 1)[1].strip()))
        return len(pairs)
    except OSError:
        return 0

"""Ionic cell models of the reference, restated in float64 numpy.

Test infrastructure only (see ``oracle/__init__.py``).

Each class mirrors the reference's duck-typed protocol: ``initialize(n)``
returns the resting potential array and sets the states; ``differentiate(V)``
advances the gating/concentration states by one ``dt`` from the *old* state
and returns dV/dt = -Iion.

Lookup tables follow the reference's float32 semantics even in a float64 run:
the grid is torch's float32 CPU ``arange`` and the interpolation's
``lower_pos`` is float32 (int64 x Python float promotes to float32).
"""
import math
import numpy as np


def arange_f32(start, end, step):
    """Values of ``torch.arange(start, end, step)`` (float32, CPU).

    torch's CPU kernel evaluates the base of each 8-wide float vector in
    double and fills the lanes as ``base + k*float(step)`` with one rounding.
    Length is ``ceil((end-start)/step)`` in double.
    """
    n = int(math.ceil((end - start) / step))
    i = np.arange(n)
    q = (i // 8) * 8
    base = (start + step * q.astype(np.float64)).astype(np.float32)
    k = (i - q).astype(np.float64)
    return (base.astype(np.float64) + k * np.float64(np.float32(step))).astype(np.float32)


def interpolate(X, table, mn, mx, res, step, mx_idx):
    """Linear table lookup (TenTusscherPanfilov.py:188-195).

    ``idx`` truncates; it is clamped to [0, mx_idx-1]; ``lower_pos`` is
    computed in float32 as the reference's int64*float promotion does.
    """
    X = np.clip(X, mn, mx)
    idx = ((X - mn) * step).astype(np.int64)
    lo = np.clip(idx, 0, mx_idx - 1)
    lower_pos = (lo.astype(np.float32) * np.float32(res)).astype(np.float32) + np.float32(mn)
    w = ((X - lower_pos.astype(np.float64)) / res)[:, None]
    return (1 - w) * table[lo] + w * table[lo + 1]


class ModifiedMS2v:
    """Modified Mitchell-Schaeffer, two variables (ModifiedMS2v.py:4-44)."""
    name = "ModifiedMS2v"

    def __init__(self, dt):
        self.tau_in = 0.1
        self.tau_out = 9.0
        self.tau_open = 100.0
        self.tau_close = 120.0
        self.u_gate = 0.13
        self.u_crit = 0.13
        self.vmin = -80.0
        self.vmax = 20.0
        self.dt = dt

    @property
    def DV(self):
        return self.vmax - self.vmin

    def initialize(self, n):
        self.H = np.full(n, 1.0)
        return np.full(n, self.vmin)

    def differentiate(self, U):
        DV = self.vmax - self.vmin
        Uad = (U - self.vmin) / DV
        H = self.H
        J_in = -1.0 * H * Uad * (Uad - self.u_crit) * (1.0 - Uad) / self.tau_in
        J_out = (1.0 - H) * Uad / self.tau_out
        dU = -DV * (J_in + J_out)
        dH = np.where(Uad > self.u_gate, -H / self.tau_close, (1 - H) / self.tau_open)
        self.H = H + self.dt * dH
        return dU


class TenTusscherPanfilov:
    """ten Tusscher-Panfilov 2006 with Rush-Larsen gates (TenTusscherPanfilov.py:6-454)."""
    name = "TenTusscherPanfilov"
    STATES = ("CaSR", "CaSS", "Cai", "D", "F", "F2", "FCaSS", "H", "J", "Ki",
              "M", "Nai", "R", "R_", "S", "Xr1", "Xr2", "Xs")
    INIT = dict(CaSR=1.3, CaSS=0.00007, Cai=0.00007, D=0.0, F=1.0, F2=1.0, FCaSS=1.0,
                H=0.75, J=0.75, Ki=138.3, M=0.0, Nai=7.67, R=0.0, R_=1.0, S=1.0,
                Xr1=0.0, Xr2=1.0, Xs=0.0)
    V_init = -86.2
    # V_tab columns (TenTusscherPanfilov.py:98-125)
    V_COLS = ("D_A", "D_B", "F2_A", "F2_B", "F_A", "F_B", "H_A", "H_B", "INaCa_A", "INaCa_B",
              "J_A", "J_B", "M_A", "M_B", "R_A", "R_B", "S_A", "S_B", "Xr1_A", "Xr1_B",
              "Xr2_A", "Xr2_B", "Xs_A", "Xs_B", "a2", "rec_iNaK", "rec_ipK")

    def __init__(self, cell_type, dt):
        self.cell_type = "EPI" if cell_type is None else cell_type
        self.dt = dt
        p = dict(Bufc=0.2, Bufsr=10., Bufss=0.4, CAPACITANCE=0.185, Cao=2.0, D_CaL_off=0.,
                 Fconst=96485.3415, GK1=5.405, GNa=14.838, GbCa=0.000592, GbNa=0.00029,
                 GpCa=0.1238, GpK=0.0146, Kbufc=0.001, Kbufsr=0.3, Kbufss=0.00025, KmCa=1.38,
                 KmK=1.0, KmNa=40.0, KmNai=87.5, Ko=5.4, KpCa=0.0005, Kup=0.00025, Nao=140.0,
                 Rconst=8314.472, T=310.0, Vc=0.016404, Vleak=0.00036, Vmaxup=0.006375,
                 Vrel=0.102, Vsr=0.001094, Vss=0.00005468, knaca=1000., knak=2.724, ksat=0.1,
                 n=0.35, pKNa=0.03, scl_tau_f=1., vHalfXs=5., xr2_off=0., EC=1.5, Vxfer=0.0038,
                 k1_=0.15, k2_=0.045, k3=0.060, k4=0.005, maxsr=2.5, minsr=1.)
        self.__dict__.update(p)
        self.GCaL = 0.00003980
        self.GKr = 0.153
        # TenTusscherPanfilov.py:88-89 (keyed on the argument, not the default)
        self.GKs = 0.392 if cell_type == "EPI" else 0.098 if cell_type == "MCELL" else 0.392
        self.Gto = 0.294 if cell_type == "EPI" else 0.294 if cell_type == "MCELL" else 0.073
        # table parameters (TenTusscherPanfilov.py:131-150)
        self.CaSS_T = (0.00001, 10.0, 0.00001, 1 / 0.00001,
                       int((10.0 - 0.00001) * (1 / 0.00001)) - 1)
        self.V_T = (-800., 800., 0.05, 1 / 0.05, int((800. + 800.) * (1 / 0.05)) - 1)
        self.VEk_T = self.V_T

    def construct_tables(self):
        dt = self.dt
        RTONF = (self.Rconst * self.T) / self.Fconst
        invKmCa_Cao = 1. / (self.KmCa + self.Cao)
        F_RT = 1. / RTONF
        KmNai3 = self.KmNai * self.KmNai * self.KmNai
        Nao3 = self.Nao * self.Nao * self.Nao
        invKmNai3_Nao3 = 1. / (KmNai3 + Nao3)
        pmf_INaCa = self.knaca * invKmNai3_Nao3 * invKmCa_Cao
        exp, expm1 = np.exp, np.expm1

        C = arange_f32(*self.CaSS_T[:3]).astype(np.float64)
        FCaSS_inf = 0.6 / (1. + (C / 0.05) * (C / 0.05)) + 0.4
        tau_FCaSS = 80. / (1. + (C / 0.05) * (C / 0.05)) + 2.
        # Column 0 (CaSS2) is never filled by the reference and stays zero
        # (TenTusscherPanfilov.py:208-214), so the release term O is zero.
        self.CaSS_tab = np.stack([
            np.zeros_like(C),
            -FCaSS_inf * expm1(-dt / tau_FCaSS),
            exp(-dt / tau_FCaSS)], axis=1)

        V = arange_f32(*self.V_T[:3]).astype(np.float64)
        D_inf = 1. / (1. + exp(((-8. + self.D_CaL_off) - V) / 7.5))
        F2_inf = 0.67 / (1. + exp((V + 35.) / 7.)) + 0.33
        F_inf = 1. / (1. + exp((V + 20.) / 7.))
        H_inf = 1. / ((1. + exp((V + 71.55) / 7.43)) * (1. + exp((V + 71.55) / 7.43)))
        M_inf = 1. / ((1. + exp((-56.86 - V) / 9.03)) * (1. + exp((-56.86 - V) / 9.03)))
        R_inf = 1. / (1. + exp((20. - V) / 6.))
        S_inf = 1. / (1. + exp((V + 20.) / 5.))
        Xr1_inf = 1. / (1. + exp((-26. - V) / 7.))
        Xr2_inf = 1. / (1. + exp((V - (-88. + self.xr2_off)) / 24.))
        Xs_inf = 1. / (1. + exp((-self.vHalfXs - V) / 14.))
        a2 = 0.25 * exp((2. * (V - 15.)) * F_RT)
        aa_D = 1.4 / (1. + exp((-35. - V) / 13.)) + 0.25
        aa_F = 1102.5 * exp((-(V + 27.) * (V + 27.)) / 225.)
        aa_F2 = 562. * exp((-(V + 27.) * (V + 27.)) / 240.)
        aa_H = np.where(V >= -40., 0., 0.057 * exp(-(V + 80.) / 6.8))
        aa_J = np.where(V >= -40., 0., ((-2.5428e4 * exp(0.2444 * V) - 6.948e-6 * exp(-0.04391 * V))
                                        * (V + 37.78)) / (1. + exp(0.311 * (V + 79.23))))
        aa_M = 1. / (1. + exp((-60. - V) / 5.))
        aa_Xr1 = 450. / (1. + exp((-45. - V) / 10.))
        aa_Xr2 = 3. / (1. + exp((-60. - V) / 20.))
        aa_Xs = 1400. / np.sqrt(1. + exp((5. - V) / 6.))
        bb_D = 1.4 / (1. + exp((V + 5.) / 5.))
        bb_F = 200. / (1. + exp((13. - V) / 10.))
        bb_F2 = 31. / (1. + exp((25. - V) / 10.))
        bb_H = np.where(V >= -40., 0.77 / (0.13 * (1. + exp(-(V + 10.66) / 11.1))),
                        2.7 * exp(0.079 * V) + 3.1e5 * exp(0.3485 * V))
        bb_J = np.where(V >= -40., (0.6 * exp(0.057 * V)) / (1. + exp(-0.1 * (V + 32.))),
                        (0.02424 * exp(-0.01052 * V)) / (1. + exp(-0.1378 * (V + 40.14))))
        bb_M = 0.1 / (1. + exp((V + 35.) / 5.)) + 0.10 / (1. + exp((V - 50.) / 200.))
        bb_Xr1 = 6. / (1. + exp((V - (-30.)) / 11.5))
        bb_Xr2 = 1.12 / (1. + exp((V - 60.) / 20.))
        bb_Xs = 1. / (1. + exp((V - 35.) / 15.))
        cc_D = 1. / (1. + exp((50. - V) / 20.))
        cc_F = 180. / (1. + exp((V + 30.) / 10.)) + 20.
        cc_F2 = 80. / (1. + exp((V + 30.) / 10.))
        den = pmf_INaCa / (1. + self.ksat * exp(((self.n - 1.) * V) * F_RT))
        rec_iNaK = 1. / ((1. + 0.1245 * exp((-0.1 * V) * F_RT)) + 0.0353 * exp(-V * F_RT))
        rec_ipK = 1. / (1. + exp((25. - V) / 5.98))
        tau_R = 9.5 * exp(-(V + 40.) * (V + 40.) / 1800.) + 0.8
        if self.cell_type == "ENDO":
            tau_S = 1000. * exp(-(V + 67.) * (V + 67.) / 1000.) + 8.
        else:
            tau_S = 85. * exp(-(V + 45.) * (V + 45.) / 320.) + 5. / (1. + exp((V - 20.) / 5.)) + 3.
        INaCa_A = den * self.Cao * exp((self.n * V) * F_RT)
        INaCa_B = den * exp(((self.n - 1.) * V) * F_RT) * Nao3 * 2.5
        J_inf = H_inf
        tau_D = aa_D * bb_D + cc_D
        tau_F2 = aa_F2 + bb_F2 + cc_F2
        tau_F_factor = aa_F + bb_F + cc_F
        tau_H = 1.0 / (aa_H + bb_H)
        tau_J = 1.0 / (aa_J + bb_J)
        tau_M = aa_M * bb_M
        tau_Xr1 = aa_Xr1 * bb_Xr1
        tau_Xr2 = aa_Xr2 * bb_Xr2
        tau_Xs = aa_Xs * bb_Xs + 80.
        tau_F = np.where(V > 0., tau_F_factor * self.scl_tau_f, tau_F_factor)

        def rl(inf, tau):
            return -inf * expm1(-dt / tau), exp(-dt / tau)

        cols = {}
        cols["D_A"], cols["D_B"] = rl(D_inf, tau_D)
        cols["F2_A"], cols["F2_B"] = rl(F2_inf, tau_F2)
        cols["F_A"], cols["F_B"] = rl(F_inf, tau_F)
        cols["H_A"], cols["H_B"] = rl(H_inf, tau_H)
        cols["J_A"], cols["J_B"] = rl(J_inf, tau_J)
        cols["M_A"], cols["M_B"] = rl(M_inf, tau_M)
        cols["R_A"], cols["R_B"] = rl(R_inf, tau_R)
        cols["S_A"], cols["S_B"] = rl(S_inf, tau_S)
        cols["Xr1_A"], cols["Xr1_B"] = rl(Xr1_inf, tau_Xr1)
        cols["Xr2_A"], cols["Xr2_B"] = rl(Xr2_inf, tau_Xr2)
        cols["Xs_A"], cols["Xs_B"] = rl(Xs_inf, tau_Xs)
        cols["INaCa_A"], cols["INaCa_B"] = INaCa_A, INaCa_B
        cols["a2"], cols["rec_iNaK"], cols["rec_ipK"] = a2, rec_iNaK, rec_ipK
        self.V_tab = np.stack([cols[c] for c in self.V_COLS], axis=1)

        Vk = arange_f32(*self.VEk_T[:3]).astype(np.float64)
        a_K1 = 0.1 / (1. + exp(0.06 * (Vk - 200.)))
        b_K1 = (3. * exp(0.0002 * (Vk + 100.)) + exp(0.1 * (Vk - 10.))) / (1. + exp(-0.5 * Vk))
        self.VEk_tab = (a_K1 / (a_K1 + b_K1))[:, None]

    def initialize(self, n):
        self.construct_tables()
        self.s = {k: np.full(n, v) for k, v in self.INIT.items()}
        return np.full(n, self.V_init)

    def differentiate(self, V):
        s = self.s
        c = self.V_COLS.index
        RTONF = (self.Rconst * self.T) / self.Fconst
        inverseVcF2 = 1. / ((2. * self.Vc) * self.Fconst)
        inverseVssF2 = 1. / ((2. * self.Vss) * self.Fconst)
        inverseVcF = 1. / (self.Vc * self.Fconst)
        pmf_INaK = self.knak * (self.Ko / (self.Ko + self.KmK))
        sqrt_Ko = math.sqrt(self.Ko / 5.4)
        F_RT = 1. / RTONF
        invVcF_Cm = inverseVcF * self.CAPACITANCE
        mn, mx, res, step, mxi = self.V_T
        Vr = interpolate(V, self.V_tab, mn, mx, res, step, mxi)
        mn, mx, res, step, mxi = self.CaSS_T
        Cr = interpolate(s["CaSS"], self.CaSS_tab, mn, mx, res, step, mxi)

        Cai, CaSS, CaSR, Ki, Nai = s["Cai"], s["CaSS"], s["CaSR"], s["Ki"], s["Nai"]
        Eca = (0.5 * RTONF) * np.log(self.Cao / Cai)
        Ek = RTONF * np.log(self.Ko / Ki)
        Eks = RTONF * np.log((self.Ko + self.pKNa * self.Nao) / (Ki + self.pKNa * Nai))
        Ena = RTONF * np.log(self.Nao / Nai)
        IpCa = (self.GpCa * Cai) / (self.KpCa + Cai)
        with np.errstate(divide="ignore", invalid="ignore"):
            ghk = np.where(V == 15., 0.5 * F_RT, (V - 15.) / np.expm1((2. * (V - 15.)) * F_RT))
        a1 = self.GCaL * self.Fconst * F_RT * 4. * ghk
        ICaL_A = a1 * Vr[:, c("a2")]
        ICaL_B = a1 * self.Cao
        IKr = self.GKr * sqrt_Ko * s["Xr1"] * s["Xr2"] * (V - Ek)
        IKs = self.GKs * s["Xs"] * s["Xs"] * (V - Eks)
        INa = self.GNa * s["M"] * s["M"] * s["M"] * s["H"] * s["J"] * (V - Ena)
        INaK = pmf_INaK * (Nai / (Nai + self.KmNa)) * Vr[:, c("rec_iNaK")]
        IbCa = self.GbCa * (V - Eca)
        IbNa = self.GbNa * (V - Ena)
        IpK = self.GpK * Vr[:, c("rec_ipK")] * (V - Ek)
        Ito = self.Gto * s["R"] * s["S"] * (V - Ek)
        mn, mx, res, step, mxi = self.VEk_T
        Kr = interpolate(V - Ek, self.VEk_tab, mn, mx, res, step, mxi)
        ICaL = s["D"] * s["F"] * s["F2"] * s["FCaSS"] * (ICaL_A * CaSS - ICaL_B)
        INaCa = Vr[:, c("INaCa_A")] * Nai * Nai * Nai - Vr[:, c("INaCa_B")] * Cai
        IK1 = self.GK1 * Kr[:, 0] * (V - Ek)
        Iion = IKr + IKs + IK1 + Ito + INa + IbNa + ICaL + IbCa + INaK + INaCa + IpCa + IpK

        Ileak = self.Vleak * (CaSR - Cai)
        Iup = self.Vmaxup / (1. + (self.Kup * self.Kup) / (Cai * Cai))
        Ixfer = self.Vxfer * (CaSS - Cai)
        diff_Ki = -((IK1 + Ito + IKr + IKs - 2. * INaK) + IpK) * invVcF_Cm
        diff_Nai = -((INa + IbNa + 3. * INaK) + 3. * INaCa) * invVcF_Cm
        kCaSR = self.maxsr - (self.maxsr - self.minsr) / (1. + (self.EC / CaSR) * (self.EC / CaSR))
        diff_Cai = ((Ixfer - ((IbCa + IpCa - 2. * INaCa) * inverseVcF2) * self.CAPACITANCE)
                    - (Iup - Ileak) * (self.Vsr / self.Vc)) / \
                   (1. + ((self.Bufc * self.Kbufc) / (self.Kbufc + Cai)) / (self.Kbufc + Cai))
        diff_R_ = self.k4 * (1. - s["R_"]) - self.k2_ * kCaSR * CaSS * s["R_"]
        k1 = self.k1_ / kCaSR
        O = (k1 * Cr[:, 0] * s["R_"]) / (self.k3 + k1 * Cr[:, 0])
        Irel = self.Vrel * O * (CaSR - CaSS)
        diff_CaSR = (Iup - Irel - Ileak) / \
                    (1. + ((self.Bufsr * self.Kbufsr) / (CaSR + self.Kbufsr)) / (CaSR + self.Kbufsr))
        diff_CaSS = ((-Ixfer * (self.Vc / self.Vss) + Irel * (self.Vsr / self.Vss))
                     + (-ICaL * inverseVssF2) * self.CAPACITANCE) / \
                    (1. + ((self.Bufss * self.Kbufss) / (CaSS + self.Kbufss)) / (CaSS + self.Kbufss))
        dt = self.dt
        new = {}
        new["CaSR"] = CaSR + diff_CaSR * dt
        new["CaSS"] = CaSS + diff_CaSS * dt
        new["Cai"] = Cai + diff_Cai * dt
        new["Ki"] = Ki + diff_Ki * dt
        new["Nai"] = Nai + diff_Nai * dt
        new["R_"] = s["R_"] + diff_R_ * dt
        new["FCaSS"] = Cr[:, 1] + Cr[:, 2] * s["FCaSS"]
        for g in ("D", "F", "F2", "H", "J", "M", "R", "S", "Xr1", "Xr2", "Xs"):
            new[g] = Vr[:, c(g + "_A")] + Vr[:, c(g + "_B")] * s[g]
        self.s = new
        return -Iion


class CourtemancheRamirezNattel:
    """Courtemanche-Ramirez-Nattel 1998 atrial model (CourtemancheRamirezNattel.py:5-444)."""
    name = "CourtemancheRamirezNattel"
    STATES = ("Ca_rel", "Ca_up", "Cai", "Ki", "d", "f", "f_Ca", "h", "j", "m", "oa", "oi",
              "u", "ua", "ui", "v", "w", "xr", "xs")
    INIT = dict(Ca_rel=1.49, Ca_up=1.49, Cai=1.02e-1, Ki=139.0, d=1.37e-4, f=0.999, f_Ca=0.775,
                h=0.965, j=0.978, m=2.91e-3, oa=3.04e-2, oi=0.999, u=0., ua=4.96e-3, ui=0.999,
                v=1., w=0.999, xr=3.29e-5, xs=1.87e-2)
    V_init = -81.2
    CAI_COLS = ("carow_1", "carow_2", "carow_3", "conCa", "f_Ca_A")
    V_COLS = ("GKur", "INaK", "IbNa", "d_A", "d_B", "f_A", "f_B", "h_A", "h_B", "j_A", "j_B",
              "m_A", "m_B", "oa_A", "oa_B", "oi_A", "oi_B", "ua_A", "ua_B", "ui_A", "ui_B",
              "vrow_29", "vrow_31", "vrow_32", "vrow_36", "vrow_7", "w_A", "w_B", "xr_A", "xr_B",
              "xs_A", "xs_B")
    FN_COLS = ("u_A", "v_A", "v_B")

    def __init__(self, dt):
        self.dt = dt
        p = dict(C_B1a=3.79138232501097e-05, C_B1b=0.0811764705882353, C_B1c=0.00705882352941176,
                 C_B1d=0.00537112496043221, C_B1e=11.5, C_Fn1=9.648e-13, C_Fn2=2.5910306809e-13,
                 C_dCa_rel=8., C_dCaup=0.0869565217391304, F=96.4867, K_Q10=3., K_up=0.00092,
                 KmCa=1.38, KmCmdn=0.00238, KmCsqn=0.8, KmKo=1.5, KmNa=87.5, KmNa3=669921.875,
                 KmNai=10., KmTrpn=0.0005, Nai=11.2, R=8.3143, T=310., Volcell=20100.,
                 Voli=13668., Volrel=96.48, Volup=1109.52, gamma=0.35, k_rel=30., k_sat=0.1,
                 maxCmdn=0.05, maxCsqn=10., maxTrpn=0.07, tau_f_Ca=2., tau_tr=180., tau_u=8.,
                 ACh=0.000001, Cao=1.8, Cm=100., GACh=0., GCaL=0.1238, GK1=0.09, GKr=0.0294,
                 GKs=0.129, GNa=7.8, GbCa=0.00113, GbNa=0.000674, Gto=0.1652, Ko=5.4, Nao=140.,
                 factorGKur=1., factorGrel=1., factorGtr=1., factorGup=1., factorhGate=0.,
                 factormGate=0., factoroaGate=0., factorxrGate=1., maxCaup=15., maxINaCa=1600.,
                 maxINaK=0.60, maxIpCa=0.275, maxIup=0.005)
        self.__dict__.update(p)
        self.f_Ca_B = math.exp(-dt / self.tau_f_Ca)
        self.u_B = math.exp(-dt / self.tau_u)
        # (mn, mx, res, step, mx_idx): CourtemancheRamirezNattel.py:145-167
        self.Cai_T = (3e-4, 30.0, 3e-4, 1 / 3e-4, int((30.0 - 3e-4) * (1 / 3e-4)) - 1)
        self.V_T = (-200., 200., 0.1, 1 / 0.1, int((200. + 200.) * (1 / 0.1)) - 1)
        self.fn_T = (-2e-11, 10.0e-11, 2e-15, 1 / 2e-15, int((10.0e-11 + 2e-11) * (1 / 2e-15)) - 1)

    def construct_tables(self):
        dt = self.dt
        exp, expm1 = np.exp, np.expm1
        R, T, F = self.R, self.T, self.F
        E_Na = ((R * T) / F) * math.log(self.Nao / self.Nai)
        f_Ca_C = math.expm1(-dt / self.tau_f_Ca)
        sigma = (math.exp(self.Nao / 67.3) - 1.) / 7.
        u_C = math.expm1(-dt / self.tau_u)

        Ca = arange_f32(*self.Cai_T[:3]).astype(np.float64)
        conCa = Ca / 1000
        cols = {}
        cols["conCa"] = conCa
        cols["carow_1"] = (self.factorGup * self.maxIup) / (1. + self.K_up / conCa)
        cols["carow_2"] = ((((self.maxTrpn * self.KmTrpn) / (conCa + self.KmTrpn)) / (conCa + self.KmTrpn)
                            + ((self.maxCmdn * self.KmCmdn) / (conCa + self.KmCmdn)) / (conCa + self.KmCmdn)
                            + 1.) / self.C_B1c) / 1000.
        cols["carow_3"] = (self.maxIpCa * conCa) / (0.0005 + conCa) - \
            ((((self.GbCa * R) * T) / 2.) / F) * np.log(self.Cao / conCa)
        f_Ca_inf = 1. / (1. + conCa / 0.00035)
        cols["f_Ca_A"] = -f_Ca_inf * f_Ca_C
        self.Cai_tab = np.stack([cols[k] for k in self.CAI_COLS], axis=1)

        V = arange_f32(*self.V_T[:3]).astype(np.float64)
        c = {}
        c["GKur"] = 0.005 + 0.05 / (1. + exp((15. - V) / 13.))
        c["IbNa"] = self.GbNa * (V - E_Na)
        with np.errstate(divide="ignore", invalid="ignore"):
            a_h = np.where(V >= -40., 0.0, 0.135 * exp(((V + 80.) - self.factorhGate) / -6.8))
            a_j = np.where(V < -40., ((-127140. * exp(0.2444 * V) - 3.474e-5 * exp(-0.04391 * V))
                                      * (V + 37.78)) / (1. + exp(0.311 * (V + 79.23))), 0.)
            a_m = np.where(V == -47.13, 3.2, (0.32 * (V + 47.13)) / (1. - exp(-0.1 * (V + 47.13))))
            aa_oa = 0.65 / (exp((V + 10.) / -8.5) + exp((30. - V) / 59.0))
            aa_oi = 1. / (18.53 + exp((V + 113.7) / 10.95))
            aa_ua = 0.65 / (exp((V + 10.) / -8.5) + exp((V - 30.) / -59.0))
            aa_ui = 1. / (21. + exp((V - 185.) / -28.))
            aa_xr = self.factorxrGate * ((0.0003 * (V + 14.1)) / (1. - exp((V + 14.1) / -5.)))
            aa_xs = (4.e-5 * (V - 19.9)) / (1. - exp((19.9 - V) / 17.))
            b_h = np.where(V >= -40., (1. / 0.13) / (1. + exp(-(V + 10.66) / 11.1)),
                           3.56 * exp(0.079 * V) + 3.1e5 * exp(0.35 * V))
            b_j = np.where(V >= -40., (0.3 * exp(-2.535e-7 * V)) / (1. + exp(-0.1 * (V + 32.))),
                           (0.1212 * exp(-0.01052 * V)) / (1. + exp(-0.1378 * (V + 40.14))))
            b_m = 0.08 * exp(-(V - self.factormGate) / 11.)
            bb_oa = 0.65 / (2.5 + exp((V + 82.) / 17.))
            bb_oi = 1. / (35.56 + exp(-(V + 1.26) / 7.44))
            bb_ua = 0.65 / (2.5 + exp((V + 82.) / 17.))
            bb_ui = exp((V - 158.) / 16.)
            bb_xr = (1.0 / self.factorxrGate) * ((7.3898e-5 * (V - 3.3328)) / (exp((V - 3.3328) / 5.1237) - 1.))
            bb_xs = (3.5e-5 * (V - 19.9)) / (exp((V - 19.9) / 9.) - 1.)
            d_inf = 1. / (1. + exp(-(V + 10.) / 8.))
            f_NaK = 1. / ((1. + 0.1245 * exp((((-0.1 * F) * V) / R) / T))
                          + (0.0365 * sigma) * exp(((-F * V) / R) / T))
            f_inf = 1. / (1. + exp((V + 28.) / 6.9))
            oa_inf = 1. / (1. + exp(-((V + 20.47) - self.factoroaGate) / 17.54))
            oi_inf = 1. / (1. + exp((V + 43.1) / 5.3))
            tau_d = np.where(V == -10., ((1. / 6.24) / 0.035) / 2.,
                             ((1. - exp(-(V + 10.) / 6.24)) / 0.035) / (V + 10.) / (1. + exp(-(V + 10.) / 6.24)))
            tau_f = 9. / (0.0197 * exp(((-0.0337 * 0.0337) * (V + 10.)) * (V + 10.)) + 0.02)
            tau_w = ((6. * (1. - exp((7.9 - V) / 5.))) / (1. + 0.3 * exp((7.9 - V) / 5.))) / (V - 7.9)
            ua_inf = 1. / (1. + exp((V + 30.3) / -9.6))
            ui_inf = 1. / (1. + exp((V - 99.45) / 27.48))
            c["vrow_29"] = self.GCaL * (V - 65.)
            kden = (self.KmNa3 + self.Nao * self.Nao * self.Nao) * (self.KmCa + self.Cao)
            sat = 1. + self.k_sat * exp(((((self.gamma - 1.) * F) * V) / R) / T)
            c["vrow_31"] = ((((self.maxINaCa * exp((((self.gamma * F) * V) / R) / T)) * self.Nai * self.Nai
                              * self.Nai * self.Cao) / (self.KmNa3 + self.Nao * self.Nao * self.Nao))
                            / (self.KmCa + self.Cao)) / sat
            c["vrow_32"] = (((((self.maxINaCa * exp(((((self.gamma - 1.) * F) * V) / R) / T)) * self.Nao
                               * self.Nao * self.Nao) / (self.KmNa3 + self.Nao * self.Nao * self.Nao))
                             / (self.KmCa + self.Cao)) / sat) / 1000.
            del kden
            c["vrow_36"] = V * self.GbCa
            c["vrow_7"] = self.GNa * (V - E_Na)
            w_inf = 1. - 1. / (1. + exp((40. - V) / 17.))
            xr_inf = 1. / (1. + exp((V + 14.1) / -6.5))
            xs_inf = 1. / np.sqrt(1. + exp((V - 19.9) / -12.7))
            c["INaK"] = (((self.maxINaK * f_NaK) / (1. + pow(self.KmNai / self.Nai, 1.5))) * self.Ko) / (self.Ko + self.KmKo)

            def rl(inf, tau):
                return -inf * expm1(-dt / tau), exp(-dt / tau)

            c["d_A"], c["d_B"] = rl(d_inf, tau_d)
            c["f_A"], c["f_B"] = rl(f_inf, tau_f)
            c["h_A"] = (-a_h / (a_h + b_h)) * expm1(-dt * (a_h + b_h))
            c["h_B"] = exp(-dt * (a_h + b_h))
            c["j_A"] = (-a_j / (a_j + b_j)) * expm1(-dt * (a_j + b_j))
            c["j_B"] = exp(-dt * (a_j + b_j))
            c["m_A"] = (-a_m / (a_m + b_m)) * expm1(-dt * (a_m + b_m))
            c["m_B"] = exp(-dt * (a_m + b_m))
            tau_oa = (1. / (aa_oa + bb_oa)) / self.K_Q10
            tau_oi = (1. / (aa_oi + bb_oi)) / self.K_Q10
            tau_ua = (1. / (aa_ua + bb_ua)) / self.K_Q10
            tau_ui = (1. / (aa_ui + bb_ui)) / self.K_Q10
            tau_xr = 1. / (aa_xr + bb_xr)
            tau_xs = 0.5 / (aa_xs + bb_xs)
            c["w_A"], c["w_B"] = rl(w_inf, tau_w)
            c["oa_A"], c["oa_B"] = rl(oa_inf, tau_oa)
            c["oi_A"], c["oi_B"] = rl(oi_inf, tau_oi)
            c["ua_A"], c["ua_B"] = rl(ua_inf, tau_ua)
            c["ui_A"], c["ui_B"] = rl(ui_inf, tau_ui)
            c["xr_A"], c["xr_B"] = rl(xr_inf, tau_xr)
            c["xs_A"], c["xs_B"] = rl(xs_inf, tau_xs)
        self.V_tab = np.stack([c[k] for k in self.V_COLS], axis=1)

        fn = arange_f32(*self.fn_T[:3]).astype(np.float64)
        tau_v = 1.91 + 2.09 / (1. + exp((3.4175e-13 - fn) / 13.67e-16))
        u_inf = 1. / (1. + exp((3.4175e-13 - fn) / 13.67e-16))
        v_inf = 1. - 1. / (1. + exp((6.835e-14 - fn) / 13.67e-16))
        self.fn_tab = np.stack([-u_inf * u_C, -v_inf * expm1(-dt / tau_v), exp(-dt / tau_v)], axis=1)

    def initialize(self, n):
        self.construct_tables()
        self.s = {k: np.full(n, v) for k, v in self.INIT.items()}
        return np.full(n, self.V_init)

    def differentiate(self, V):
        s = self.s
        cv = self.V_COLS.index
        cc = self.CAI_COLS.index
        Vr = interpolate(V, self.V_tab, *self.V_T)
        Cr = interpolate(s["Cai"], self.Cai_tab, *self.Cai_T)
        R, T, F = self.R, self.T, self.F
        E_K = ((R * T) / F) * np.log(self.Ko / s["Ki"])
        ICaL = Vr[:, cv("vrow_29")] * s["d"] * s["f"] * s["f_Ca"]
        IKACh = (self.GACh * (10.0 / (1.0 + 9.13652 / pow(self.ACh, 0.477811)))) * \
            (0.0517 + 0.4516 / (1.0 + np.exp((V + 59.53) / 17.18))) * (V - E_K)
        INa = Vr[:, cv("vrow_7")] * s["m"] * s["m"] * s["m"] * s["h"] * s["j"]
        INaCa = Vr[:, cv("vrow_31")] - s["Cai"] * Vr[:, cv("vrow_32")]
        vrow_13 = self.Gto * (V - E_K)
        vrow_18 = (self.factorGKur * Vr[:, cv("GKur")]) * (V - E_K)
        vrow_21 = (self.GKr * (V - E_K)) / (1. + np.exp((V + 15.) / 22.4))
        vrow_24 = self.GKs * (V - E_K)
        IK1 = (self.GK1 * (V - E_K)) / (1. + np.exp(0.07 * (V + 80.)))
        IKr = vrow_21 * s["xr"]
        IKs = vrow_24 * s["xs"] * s["xs"]
        IKur = vrow_18 * s["ua"] * s["ua"] * s["ua"] * s["ui"]
        IpCa = Cr[:, cc("carow_3")] + Vr[:, cv("vrow_36")]
        Ito = vrow_13 * s["oa"] * s["oa"] * s["oa"] * s["oi"]
        Iion = (INa + IK1 + Ito + IKur + IKr + IKs + ICaL + IpCa + INaCa + Vr[:, cv("IbNa")]
                + Vr[:, cv("INaK")] + IKACh)

        Itr = (self.factorGtr * (s["Ca_up"] - s["Ca_rel"])) / self.tau_tr
        Irel = self.factorGrel * s["u"] * s["u"] * s["v"] * self.k_rel * s["w"] * (s["Ca_rel"] - Cr[:, cc("conCa")])
        dIups = Cr[:, cc("carow_1")] - (self.maxIup / self.maxCaup) * s["Ca_up"]
        diff_Ca_rel = (Itr - Irel) / (1. + (self.C_dCa_rel / (s["Ca_rel"] + self.KmCsqn)) / (s["Ca_rel"] + self.KmCsqn))
        diff_Ki = -((Ito + IKr + IKur + IKs + IK1 + IKACh) - 2.0 * Vr[:, cv("INaK")]) / (F * self.Voli)
        diff_Ca_up = dIups - Itr * self.C_dCaup
        diff_Cai = ((self.C_B1d * ((INaCa + INaCa - IpCa) - ICaL) - self.C_B1e * dIups) + Irel) / Cr[:, cc("carow_2")]
        dt = self.dt
        new = {}
        new["Ca_rel"] = s["Ca_rel"] + diff_Ca_rel * dt
        new["Ca_up"] = s["Ca_up"] + diff_Ca_up * dt
        new["Cai"] = s["Cai"] + diff_Cai * dt
        new["Ki"] = s["Ki"] + diff_Ki * dt
        fn = self.C_Fn1 * Irel - self.C_Fn2 * (ICaL - 0.4 * INaCa)
        Fr = interpolate(fn, self.fn_tab, *self.fn_T)
        new["f_Ca"] = Cr[:, cc("f_Ca_A")] + self.f_Ca_B * s["f_Ca"]
        new["u"] = Fr[:, 0] + self.u_B * s["u"]
        new["v"] = Fr[:, 1] + Fr[:, 2] * s["v"]
        for g in ("d", "f", "h", "j", "m", "oa", "oi", "ua", "ui", "w", "xr", "xs"):
            new[g] = Vr[:, cv(g + "_A")] + Vr[:, cv(g + "_B")] * s[g]
        self.s = new
        return -Iion

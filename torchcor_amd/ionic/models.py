"""The three ionic models of the reference's monodomain path.

Parameter names and defaults follow ionic/ModifiedMS2v.py:6-23,
ionic/TenTusscherPanfilov.py:8-183 and ionic/CourtemancheRamirezNattel.py:7-200.
"""
import ctypes as C

import torch

from .. import _lib as L
from .base import IonicModel


class ModifiedMS2v(IonicModel):
    """Modified Mitchell-Schaeffer, two variables (ModifiedMS2v.py:4-44)."""
    name = "ModifiedMS2v"
    STATES = ("H",)
    INIT = {"H": 1.0}
    PARAMS = tuple(L.MS2V_FIELDS)

    def __init__(self, dt, device=None, dtype=torch.float32):
        super().__init__(dt, device, dtype)
        self.tau_in = 0.1
        self.tau_out = 9.0
        self.tau_open = 100.0
        self.tau_close = 120.0
        self.u_gate = 0.13
        self.u_crit = 0.13
        self.vmin = -80.0
        self.vmax = 20.0

    @property
    def DV(self):
        return self.vmax - self.vmin

    @property
    def V_init(self):
        return self.vmin

    MODEL = L.MODEL_MS2V

    def c_params(self):
        return L.MS2vParams(**self._param_values())

    def c_tables(self):
        return (None, None, None)

    def step(self, io):
        p = L.MS2vParams(**self._param_values())
        L.call("tc_ms2v_step", C.byref(p), L.ptr(self._states), C.byref(io),
               L.dtype_code(self.dtype), L.stream_handle(self.device))


_TTP_PARAMS = dict(
    Bufc=0.2, Bufsr=10., Bufss=0.4, CAPACITANCE=0.185, Cao=2.0, D_CaL_off=0., Fconst=96485.3415,
    GK1=5.405, GNa=14.838, GbCa=0.000592, GbNa=0.00029, GpCa=0.1238, GpK=0.0146, Kbufc=0.001,
    Kbufsr=0.3, Kbufss=0.00025, KmCa=1.38, KmK=1.0, KmNa=40.0, KmNai=87.5, Ko=5.4, KpCa=0.0005,
    Kup=0.00025, Nao=140.0, Rconst=8314.472, T=310.0, Vc=0.016404, Vleak=0.00036, Vmaxup=0.006375,
    Vrel=0.102, Vsr=0.001094, Vss=0.00005468, knaca=1000., knak=2.724, ksat=0.1, n=0.35,
    pKNa=0.03, scl_tau_f=1., vHalfXs=5., xr2_off=0., EC=1.5, Vxfer=0.0038, k1_=0.15, k2_=0.045,
    k3=0.060, k4=0.005, maxsr=2.5, minsr=1.)


class TenTusscherPanfilov(IonicModel):
    """ten Tusscher-Panfilov 2006 (TenTusscherPanfilov.py:6-454)."""
    name = "TenTusscherPanfilov"
    STATES = ("CaSR", "CaSS", "Cai", "D", "F", "F2", "FCaSS", "H", "J", "Ki", "M", "Nai", "R",
              "R_", "S", "Xr1", "Xr2", "Xs")
    INIT = dict(CaSR=1.3, CaSS=0.00007, Cai=0.00007, D=0.0, F=1.0, F2=1.0, FCaSS=1.0, H=0.75,
                J=0.75, Ki=138.3, M=0.0, Nai=7.67, R=0.0, R_=1.0, S=1.0, Xr1=0.0, Xr2=1.0, Xs=0.0)
    V_init = -86.2
    # V table rows: the 27 columns of TenTusscherPanfilov.py:197-318 and a zero
    # (16-byte rows for vector loads)
    V_ROWS, CASS_ROWS, VEK_ROWS = (32000, 28), (999999, 3), (32000, 1)

    def __init__(self, cell_type, dt, device=None, dtype=torch.float32):
        super().__init__(dt, device, dtype)
        self.cell_type = "EPI" if cell_type is None else cell_type
        for k, v in _TTP_PARAMS.items():
            setattr(self, k, v)
        self.GCaL_init = 0.00003980
        self.GKr_init = 0.153
        # keyed on the argument as in TenTusscherPanfilov.py:88-89
        self.GKs_init = 0.392 if cell_type == "EPI" else 0.098 if cell_type == "MCELL" else 0.392
        self.Gto_init = 0.294 if cell_type == "EPI" else 0.294 if cell_type == "MCELL" else 0.073
        self.V_tab = self.CaSS_tab = self.VEk_tab = None

    def _params(self):
        vals = {k: float(getattr(self, k)) for k in L.TTP_FIELDS if not k.startswith("G") or k in
                ("GK1", "GNa", "GbCa", "GbNa", "GpCa", "GpK")}
        vals.update(GCaL=float(self.GCaL_init), GKr=float(self.GKr_init), GKs=float(self.GKs_init),
                    Gto=float(self.Gto_init))
        return L.TTPParams(endo=1 if self.cell_type == "ENDO" else 0, **vals)

    def construct_tables(self):
        mk = lambda shp: torch.empty(shp, dtype=self.dtype, device=self.device)  # noqa: E731
        self.V_tab, self.CaSS_tab, self.VEk_tab = mk(self.V_ROWS), mk(self.CASS_ROWS), mk(self.VEK_ROWS)
        p = self._params()
        L.call("tc_ttp_tables", C.byref(p), float(self.dt), L.ptr(self.V_tab), L.ptr(self.CaSS_tab),
               L.ptr(self.VEk_tab), L.dtype_code(self.dtype), L.stream_handle(self.device))

    MODEL = L.MODEL_TTP

    def c_params(self):
        return self._params()

    def c_tables(self):
        return (self.V_tab, self.CaSS_tab, self.VEk_tab)

    def step(self, io):
        p = self._params()
        L.call("tc_ttp_step", C.byref(p), L.ptr(self._states), L.ptr(self.V_tab),
               L.ptr(self.CaSS_tab), L.ptr(self.VEk_tab), C.byref(io), L.dtype_code(self.dtype),
               L.stream_handle(self.device))


_CRN_PARAMS = dict(
    C_B1a=3.79138232501097e-05, C_B1b=0.0811764705882353, C_B1c=0.00705882352941176,
    C_B1d=0.00537112496043221, C_B1e=11.5, C_Fn1=9.648e-13, C_Fn2=2.5910306809e-13, C_dCa_rel=8.,
    C_dCaup=0.0869565217391304, F=96.4867, K_Q10=3., K_up=0.00092, KmCa=1.38, KmCmdn=0.00238,
    KmCsqn=0.8, KmKo=1.5, KmNa=87.5, KmNa3=669921.875, KmNai=10., KmTrpn=0.0005, Nai=11.2,
    R=8.3143, T=310., Volcell=20100., Voli=13668., Volrel=96.48, Volup=1109.52, gamma=0.35,
    k_rel=30., k_sat=0.1, maxCmdn=0.05, maxCsqn=10., maxTrpn=0.07, tau_f_Ca=2., tau_tr=180.,
    tau_u=8., ACh=0.000001, Cao=1.8, Cm=100., GACh=0., GCaL=0.1238, GK1=0.09, GKr=0.0294,
    GKs=0.129, GNa=7.8, GbCa=0.00113, GbNa=0.000674, Gto=0.1652, Ko=5.4, Nao=140.,
    factorGKur=1., factorGrel=1., factorGtr=1., factorGup=1., factorhGate=0., factormGate=0.,
    factoroaGate=0., factorxrGate=1., maxCaup=15., maxINaCa=1600., maxINaK=0.60, maxIpCa=0.275,
    maxIup=0.005)


class CourtemancheRamirezNattel(IonicModel):
    """Courtemanche-Ramirez-Nattel 1998 (CourtemancheRamirezNattel.py:5-444)."""
    name = "CourtemancheRamirezNattel"
    STATES = ("Ca_rel", "Ca_up", "Cai", "Ki", "d", "f", "f_Ca", "h", "j", "m", "oa", "oi", "u",
              "ua", "ui", "v", "w", "xr", "xs")
    INIT = dict(Ca_rel=1.49, Ca_up=1.49, Cai=1.02e-1, Ki=139.0, d=1.37e-4, f=0.999, f_Ca=0.775,
                h=0.965, j=0.978, m=2.91e-3, oa=3.04e-2, oi=0.999, u=0., ua=4.96e-3, ui=0.999,
                v=1., w=0.999, xr=3.29e-5, xs=1.87e-2)
    V_init = -81.2
    CAI_ROWS, V_ROWS, FN_ROWS = (100000, 5), (4000, 32), (60000, 3)
    PARAMS = tuple(L.CRN_FIELDS)

    def __init__(self, dt, device=None, dtype=torch.float32):
        super().__init__(dt, device, dtype)
        for k, v in _CRN_PARAMS.items():
            setattr(self, k, v)
        self.Cai_tab = self.V_tab = self.fn_tab = None

    def construct_tables(self):
        mk = lambda shp: torch.empty(shp, dtype=self.dtype, device=self.device)  # noqa: E731
        self.Cai_tab, self.V_tab, self.fn_tab = mk(self.CAI_ROWS), mk(self.V_ROWS), mk(self.FN_ROWS)
        p = L.CRNParams(**self._param_values())
        L.call("tc_crn_tables", C.byref(p), float(self.dt), L.ptr(self.Cai_tab), L.ptr(self.V_tab),
               L.ptr(self.fn_tab), L.dtype_code(self.dtype), L.stream_handle(self.device))

    MODEL = L.MODEL_CRN

    def c_params(self):
        return L.CRNParams(**self._param_values())

    def c_tables(self):
        return (self.Cai_tab, self.V_tab, self.fn_tab)

    def step(self, io):
        p = L.CRNParams(**self._param_values())
        L.call("tc_crn_step", C.byref(p), L.ptr(self._states), L.ptr(self.Cai_tab),
               L.ptr(self.V_tab), L.ptr(self.fn_tab), C.byref(io), L.dtype_code(self.dtype),
               L.stream_handle(self.device))

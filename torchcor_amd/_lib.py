"""ctypes binding of libtorchcor_amd.so (the C ABI of include/torchcor_amd.h).

torch is imported first so that the library resolves the HIP runtime torch
already loaded (one runtime per process, shared streams).  A missing library
raises at import: there is no CPU or PyTorch fallback for these operations.
"""
import ctypes as C
from pathlib import Path

import torch  # noqa: F401  (must precede loading the HIP library)

LIB_PATH = Path(__file__).resolve().parent / "libtorchcor_amd.so"

TC_F32, TC_F64 = 0, 1
TC_ERR_ARG, TC_ERR_HIP, TC_ERR_SIZE, TC_ERR_STATE = -1, -2, -3, -4
MAX_STIMULI = 32
# rows of the largest system tc_pcg_solve_static takes (its static scratch)
TC_PCG_STATIC_MAX_ROWS = 1 << 20

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -m torchcor_amd.build` "
        "(hipcc, gfx950). torchcor_amd has no fallback path.")

lib = C.CDLL(str(LIB_PATH))

p = C.c_void_p
i64 = C.c_int64
i32 = C.c_int
dbl = C.c_double


class Sell(C.Structure):
    _fields_ = [("n", i64), ("n_slices", i64), ("slice_ptr", p), ("col", p), ("dcol", p),
                ("pat", p), ("slice_tag", p)]


class Stage(C.Structure):
    _fields_ = [("blocks", C.c_int), ("max_cols", C.c_int), ("col_ptr", C.c_void_p),
                ("own", C.c_void_p), ("cols", C.c_void_p), ("lidx", C.c_void_p),
                ("chunk", C.c_int)]


class Stimulus(C.Structure):
    _fields_ = [("start", dbl), ("duration", dbl), ("period", dbl), ("count", dbl), ("amp", dbl)]


POOL_STIM = 4
STAGE_MAX_COLS = 4096          # TC_STAGE_MAX_COLS: columns of one block's staged set


class IonicIO(C.Structure):
    _fields_ = [("n", i64), ("dt", dbl), ("dt_rhs", dbl), ("Cm", dbl), ("V", p), ("w", p),
                ("dV", p), ("stim_mask", p), ("stim_active", C.c_uint32),
                ("stim_amp", dbl * MAX_STIMULI), ("AT", p), ("RT", p), ("lat_t", dbl),
                ("gate", p), ("v_f64", i32), ("w_f64", i32), ("lanes", i32),
                ("lane_stride", i32), ("lane0", i32), ("pool_slices", p), ("pool_slice_sys", p),
                ("pool_sys_row0", p), ("pool_sys_n", p), ("pool_sys_t", p), ("pool_sys_steps", p),
                ("pool_sys_bits", p), ("pool_sys_stim", p), ("state_stride", i64),
                ("stream_states", i32), ("state_block", i32)]


MS2V_FIELDS = ["tau_in", "tau_out", "tau_open", "tau_close", "u_gate", "u_crit", "vmin", "vmax"]
TTP_FIELDS = ["GCaL", "GKr", "GKs", "Gto", "Bufc", "Bufsr", "Bufss", "CAPACITANCE", "Cao",
              "D_CaL_off", "Fconst", "GK1", "GNa", "GbCa", "GbNa", "GpCa", "GpK", "Kbufc",
              "Kbufsr", "Kbufss", "KmCa", "KmK", "KmNa", "KmNai", "Ko", "KpCa", "Kup", "Nao",
              "Rconst", "T", "Vc", "Vleak", "Vmaxup", "Vrel", "Vsr", "Vss", "knaca", "knak",
              "ksat", "n", "pKNa", "scl_tau_f", "vHalfXs", "xr2_off", "EC", "Vxfer", "k1_",
              "k2_", "k3", "k4", "maxsr", "minsr"]
CRN_FIELDS = ["C_B1c", "C_B1d", "C_B1e", "C_Fn1", "C_Fn2", "C_dCa_rel", "C_dCaup", "F", "K_Q10",
              "K_up", "KmCa", "KmCmdn", "KmCsqn", "KmKo", "KmNa3", "KmNai", "KmTrpn", "Nai", "R",
              "T", "Voli", "gamma", "k_rel", "k_sat", "maxCmdn", "maxTrpn", "tau_f_Ca", "tau_tr",
              "tau_u", "ACh", "Cao", "GACh", "GCaL", "GK1", "GKr", "GKs", "GNa", "GbCa", "GbNa",
              "Gto", "Ko", "Nao", "factorGKur", "factorGrel", "factorGtr", "factorGup",
              "factorhGate", "factormGate", "factoroaGate", "factorxrGate", "maxCaup", "maxINaCa",
              "maxINaK", "maxIpCa", "maxIup"]


class MS2vParams(C.Structure):
    _fields_ = [(f, dbl) for f in MS2V_FIELDS]


class TTPParams(C.Structure):
    _fields_ = [("endo", i32)] + [(f, dbl) for f in TTP_FIELDS]


class CRNParams(C.Structure):
    _fields_ = [(f, dbl) for f in CRN_FIELDS]


class SimDesc(C.Structure):
    _fields_ = [("model", i32), ("params", p), ("states", p), ("tab", p * 3), ("pcg", p),
                ("Mv", p), ("Kv", p), ("dtype", i32), ("n", i64), ("dt", dbl), ("dt_ionic", dbl),
                ("Cm", dbl), ("theta", dbl), ("x_f64", i32), ("w_f64", i32), ("x", p * 2),
                ("xp", p * 2),
                ("w", p), ("AT", p), ("RT", p),
                ("stim_mask", p), ("n_stim", i32), ("stim", Stimulus * MAX_STIMULI),
                ("a_tol", dbl), ("r_tol", dbl), ("max_iter", i32), ("linear_guess", i32),
                ("depth", i32), ("state_stride", i64), ("separate_check", i32), ("k_history", i32),
                ("graph", i32), ("state_block", i32)]


BATCH_MAX_K = 5
BATCH_MAX_GROUPS = 8


class BatchGroup(C.Structure):
    _fields_ = [("model", i32), ("params", p), ("states", p), ("tab", p * 3), ("b0", i32),
                ("count", i32), ("state_stride", i64)]


class BatchDesc(C.Structure):
    _fields_ = [("n", i64), ("nnz", i64), ("B", i32), ("lanes", i32), ("NK", i32),
                ("dtype", i32), ("x_f64", i32), ("w_f64", i32), ("row_ptr", p), ("col", p),
                ("Mv", p), ("Kv", p * BATCH_MAX_K), ("coef", p), ("dt", dbl), ("dt_ionic", dbl),
                ("Cm", dbl), ("theta", dbl), ("x", p * 2), ("xp", p * 2), ("w", p), ("Dinv", p),
                ("AT", p), ("RT", p), ("stim_mask", p), ("n_stim", i32),
                ("stim", Stimulus * MAX_STIMULI), ("n_groups", i32),
                ("groups", BatchGroup * BATCH_MAX_GROUPS), ("a_tol", dbl), ("r_tol", dbl),
                ("max_iter", i32), ("linear_guess", i32), ("depth", i32), ("ws", p),
                ("ws_bytes", C.c_size_t), ("max_blocks", i32), ("graph", i32)]


class ServeDesc(C.Structure):
    _fields_ = [("dtype", i32), ("x_f64", i32), ("w_f64", i32), ("cap_slices", i64),
                ("cap_entries", i64), ("max_sys", i32), ("slice_off", p), ("slice_w", p),
                ("slice_sys", p), ("col", p), ("dcol", p), ("pat", p), ("Mv", p), ("Kv", p),
                ("Av", p), ("Dinv", p), ("x", p * 2), ("xp", p * 2), ("w", p), ("r", p), ("p", p * 2), ("Ap", p),
                ("AT", p), ("RT", p), ("stim_mask", p), ("states", p * 3), ("params", p * 3),
                ("tab", (p * 3) * 3), ("dt", dbl), ("dt_ionic", dbl), ("Cm", dbl),
                ("theta", dbl), ("a_tol", dbl), ("r_tol", dbl), ("max_iter", i32),
                ("linear_guess", i32), ("depth", i32), ("ws", p), ("ws_bytes", C.c_size_t),
                ("k_history", i32), ("graph", i32)]


class ServeSystem(C.Structure):
    _fields_ = [("n", i64), ("n_entries", i64), ("slice_ptr", p), ("col", p), ("dcol", p),
                ("pat", p), ("slice_tag", p), ("Mv", p), ("Kv", p), ("Av", p), ("Dinv", p), ("model", i32), ("n_states", i32),
                ("states", p), ("V0", p), ("stim_mask", p), ("n_stim", i32),
                ("stim", Stimulus * POOL_STIM), ("lat", i32), ("slices_per_block", i32)]


MODEL_MS2V, MODEL_TTP, MODEL_CRN = 0, 1, 2


def _proto(name, *args):
    fn = getattr(lib, name)
    fn.restype = C.c_int
    fn.argtypes = list(args)
    return fn


lib.tc_last_error.restype = C.c_char_p
lib.tc_last_error.argtypes = []
lib.tc_abi_version.restype = C.c_int
lib.tc_abi_version.argtypes = []

SIZE_P = C.POINTER(C.c_size_t)
I64_P = C.POINTER(i64)
INT_P = C.POINTER(C.c_int)
SELL_P = C.POINTER(Sell)
IO_P = C.POINTER(IonicIO)

FUNCS = {
    "tc_pattern_workspace_size": (i64, i32, SIZE_P),
    "tc_pattern_sort": (p, i64, i32, i64, p, C.c_size_t, I64_P, p),
    "tc_pattern_finish": (i64, i32, i64, i64, p, p, p, p, p, p),
    "tc_assemble_p1": (p, p, i64, i32, p, p, p, p, p, p, i64, p, i32, p, p, p),
    "tc_operator_values": (p, p, p, p, i64, i64, dbl, dbl, dbl, dbl, i32, p, p, p, p, p),
    "tc_sell_layout": (p, i64, p, I64_P, p),
    "tc_sell_share": (p, p, i64, p, p, p, p),
    "tc_sell_columns": (p, p, i64, p, p, p, p, p),
    "tc_sell_values": (p, p, i64, p, p, p, i32, p, p, p),
    "tc_sell_deltas": (SELL_P, p, p, p),
    "tc_rcm_order": (i64, p, p, p, I64_P, I64_P),
    "tc_level_order": (i64, p, p, p, p, I64_P, I64_P),
    "tc_table_shape": (C.c_char_p, i32, I64_P, C.POINTER(C.c_int32)),
    "tc_table_read_f64": (C.c_char_p, i32, i32, i32, i64, p),
    "tc_table_read_i64": (C.c_char_p, i32, i32, i32, i64, p),
    "tc_sell_spmv": (SELL_P, p, p, p, i32, p),
    "tc_rhs": (SELL_P, p, p, p, p, dbl, p, i32, p),
    "tc_pcg_workspace_size": (i64, i32, SIZE_P),
    "tc_pcg_create": (C.POINTER(p), SELL_P, p, p, i32, p, C.c_size_t),
    "tc_pcg_destroy": (p,),
    "tc_pcg_solve": (p, p, p, p, p, dbl, dbl, i32, i32, INT_P, p),
    "tc_pcg_set_blocks": (p, i32),
    "tc_pcg_set_stage": (p, C.POINTER(Stage)),
    "tc_sell_stage": (i64, i64, p, p, i32, i32, p, p, p, p, C.POINTER(C.c_int32)),
    "tc_pcg_set_timing": (p, i32),
    "tc_pcg_get_timing": (p, C.POINTER(dbl), I64_P),
    "tc_pcg_solve_static": (SELL_P, p, p, i32, p, p, p, p, dbl, dbl, i32, i32, INT_P, p),
    "tc_ms2v_step": (C.POINTER(MS2vParams), p, IO_P, i32, p),
    "tc_ttp_tables": (C.POINTER(TTPParams), dbl, p, p, p, i32, p),
    "tc_ttp_step": (C.POINTER(TTPParams), p, p, p, p, IO_P, i32, p),
    "tc_crn_tables": (C.POINTER(CRNParams), dbl, p, p, p, i32, p),
    "tc_crn_step": (C.POINTER(CRNParams), p, p, p, p, IO_P, i32, p),
    "tc_lat_update": (p, p, p, i64, dbl, i32, p),
    "tc_stream_copy": (p, p, i64, p),
    "tc_sim_create": (C.POINTER(p), C.POINTER(SimDesc)),
    "tc_sim_destroy": (p,),
    "tc_sim_advance": (p, i64, C.POINTER(C.c_int32), p, i64, i64, p),
    "tc_sim_advance_ring": (p, i64, C.POINTER(C.c_int32), p, p, i32, i64, i64, p),
    "tc_sim_finish": (p, p),
    "tc_sim_status": (p, C.POINTER(dbl), I64_P, INT_P, I64_P),
    "tc_sim_set_timing": (p, i32),
    "tc_sim_get_timing": (p, C.POINTER(dbl), I64_P),
    "tc_sim_graph_stats": (p, I64_P, I64_P),
    "tc_batch_workspace_size": (i64, i32, i32, SIZE_P),
    "tc_batch_create": (C.POINTER(p), C.POINTER(BatchDesc), p),
    "tc_batch_destroy": (p,),
    "tc_batch_advance": (p, i64, C.POINTER(C.c_int32), p),
    "tc_batch_finish": (p, I64_P, p),
    "tc_batch_status": (p, C.POINTER(dbl), I64_P, INT_P),
    "tc_batch_graph_stats": (p, I64_P, I64_P),
    "tc_serve_workspace_size": (i64, i32, SIZE_P),
    "tc_serve_create": (C.POINTER(p), C.POINTER(ServeDesc), p),
    "tc_serve_destroy": (p,),
    "tc_serve_fits": (p, i64, i64, INT_P),
    "tc_serve_add": (p, C.POINTER(ServeSystem), INT_P, p),
    "tc_serve_remove": (p, i32, p, p, p, I64_P, I64_P, p),
    "tc_serve_advance": (p, i64, C.POINTER(C.c_int32), p),
    "tc_serve_stats": (p, I64_P, I64_P),
    "tc_serve_graph_stats": (p, I64_P, I64_P),
}
for _n, _a in FUNCS.items():
    _proto(_n, *_a)

# Every name the header declares, for the export check in the tests.
EXPORTS = ["tc_last_error", "tc_abi_version", *FUNCS]


class TorchcorError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        msg = lib.tc_last_error().decode(errors="replace")
        raise TorchcorError(f"{what or 'torchcor_amd'} failed ({rc}): {msg}")


def call(name, *args):
    check(getattr(lib, name)(*args), name)


def dtype_code(dtype):
    if dtype == torch.float32:
        return TC_F32
    if dtype == torch.float64:
        return TC_F64
    raise TypeError(f"unsupported dtype {dtype} (float32 or float64)")


def ptr(t):
    """Device pointer of a tensor (None -> NULL); requires a GPU tensor."""
    if t is None:
        return None
    if not t.is_cuda:
        raise TorchcorError("torchcor_amd kernels need GPU tensors (no CPU path)")
    return t.data_ptr()


def stream_handle(device=None):
    return torch.cuda.current_stream(device).cuda_stream

"""One mesh's operators on the GPU: CSR pattern, M/K/A values, SELL layout, PCG.

This is the engine behind ``Matrices3D``/``Matrices3DSurface``
(torchcor/core/assemble.py:99-248), ``Monodomain.assemble``
(torchcor/simulator/monodomain.py:83-106) and ``ConjugateGradient``
(torchcor/core/solver.py:5-76).  Every numeric step runs in the HIP library;
this class only owns the tensors and sequences the calls.
"""
import ctypes as C

import torch

from .. import _lib as L


class P1System:
    """CSR pattern of a P1 mesh plus (after ``set_operators``) its working operators.

    Pattern arrays are int32 ``row_ptr`` (n+1) and ``col`` (nnz), sorted by
    (row, col) with explicit zeros kept, identical to the reference's
    coalesced COO.  ``seg_ptr``/``contrib`` map each CSR entry to its element
    contributions in the reference's listing order.
    """

    #: hot passes read int16 column offsets (col - row) when every offset fits
    int16_columns = True
    #: slices whose rows share one stencil are laid out on it (with int16 offsets)
    share_slices = True
    #: operators without shared slices get staged gathers in the time-step
    #: runtime's spmv pass (tc_sell_stage) when every block's column set fits
    stage_gathers = True

    #: force the staged split (blocks, slices per unit) instead of choosing
    #: it (None); for tests of the unit loop on small meshes
    stage_override = None

    @staticmethod
    def stage_split(n):
        """(blocks, chunk) of the staged split for n rows: 8 slices a block
        while that needs at most 1024 blocks (chunk 0: each block stages its
        whole range at once); beyond, at most 1024 blocks (a multiple of 8)
        of a multiple of 8 slices each, staged in units of 8 slices (512
        rows: on an unstructured mesh about 3,000 columns, where units of 16
        slices exceed the 4,096-column stage)."""
        from .. import autotune
        ns = -(-int(n) // 64)
        g8 = autotune.pcg_blocks(n, 8)
        if -(-ns // g8) <= 16:
            return g8, 0
        per = 8 * -(-ns // (8 * 1024))
        G = -(-ns // per)
        up = -(-G // 8) * 8
        if -(-ns // up) == per:
            G = up
        if -(-ns // G) != per:
            return g8, 0
        return G, 8

    @staticmethod
    def sort_window(n):
        """Rows of one staged unit of the staged split of n rows: the window
        in which ``slice_sort`` groups rows by length, so that each unit
        keeps its own nodes."""
        G, chunk = P1System.stage_split(n)
        ns = -(-int(n) // 64)
        return (chunk or -(-ns // G)) * 64

    def __init__(self, nodes, elems, device=None):
        device = torch.device(device) if device is not None else elems.device
        if device.type != "cuda":
            raise L.TorchcorError("P1System needs a GPU device (no CPU path)")
        self.device = device
        #: rows sorted by length inside windows of this many rows (0: not
        #: sorted); the staged split aligned with them is tried first
        self.row_window = 0
        self.elems = torch.as_tensor(elems).to(device=device, dtype=torch.int32).contiguous()
        self.nodes = torch.as_tensor(nodes).to(device=device, dtype=torch.float64).contiguous()
        self.n = int(self.nodes.shape[0])
        self.E, self.k = (int(s) for s in self.elems.shape)
        if self.k not in (3, 4):
            raise ValueError("elements must be triangles (3) or tetrahedra (4)")
        s = L.stream_handle(device)
        nb = C.c_size_t(0)
        L.call("tc_pattern_workspace_size", self.E, self.k, C.byref(nb))
        ws = torch.empty(nb.value, dtype=torch.uint8, device=device)
        nnz = C.c_int64(0)
        L.call("tc_pattern_sort", L.ptr(self.elems), self.E, self.k, self.n, L.ptr(ws), nb.value,
               C.byref(nnz), s)
        self.nnz = int(nnz.value)
        i32 = dict(dtype=torch.int32, device=device)
        self.row_ptr = torch.empty(self.n + 1, **i32)
        self.col = torch.empty(self.nnz, **i32)
        self.seg_ptr = torch.empty(self.nnz + 1, **i32)
        self.contrib = torch.empty(self.E * self.k * self.k, **i32)
        L.call("tc_pattern_finish", self.E, self.k, self.n, self.nnz, L.ptr(ws), L.ptr(self.row_ptr),
               L.ptr(self.col), L.ptr(self.seg_ptr), L.ptr(self.contrib), s)
        torch.cuda.current_stream(device).synchronize()
        del ws
        self._pcg = None

    # ---------------------------------------------------------- assembly
    def assemble(self, sigma=None, sigma_l=None, sigma_t=None, fibres=None, round_f32=True):
        """Global (M, K) float64 values in CSR order (assemble.py:163-248).

        Give either ``sigma`` (E,3,3) or ``sigma_l``, ``sigma_t`` (E) and
        ``fibres`` (E,3).  ``round_f32`` rounds each element contribution to
        float32 first, as the reference's list-based COO path does.
        """
        dev = self.device
        f64 = lambda t: None if t is None else torch.as_tensor(t).to(device=dev, dtype=torch.float64).contiguous()  # noqa: E731
        sig, sl, st, fb = f64(sigma), f64(sigma_l), f64(sigma_t), f64(fibres)
        if sig is None and (sl is None or st is None or fb is None):
            raise ValueError("need sigma, or sigma_l, sigma_t and fibres")
        kk = self.k * self.k
        local = torch.empty(2 * kk * self.E, dtype=torch.float64, device=dev)
        M = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        K = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        L.call("tc_assemble_p1", L.ptr(self.nodes), L.ptr(self.elems), self.E, self.k, L.ptr(sig),
               L.ptr(sl), L.ptr(st), L.ptr(fb), L.ptr(self.seg_ptr), L.ptr(self.contrib),
               self.nnz, L.ptr(local), 1 if round_f32 else 0, L.ptr(M), L.ptr(K),
               L.stream_handle(dev))
        return M, K

    def release_assembly_maps(self):
        """Free the entry->contribution maps once values are final."""
        self.seg_ptr = self.contrib = None

    # --------------------------------------------------------- operators
    def set_operators(self, M, K, Chi, Cm, theta, dt, dtype):
        """Working values Mv, Kv = K/Chi, Av = Cm*Mv + theta*dt*Kv, Dinv in ``dtype``
        (monodomain.py:93-100) plus their SELL-64 copies and a PCG plan."""
        dev = self.device
        self.dtype = dtype
        self.code = L.dtype_code(dtype)
        s = L.stream_handle(dev)
        mk = lambda m: torch.empty(m, dtype=dtype, device=dev)  # noqa: E731
        self.Mv, self.Kv, self.Av, self.Dinv = mk(self.nnz), mk(self.nnz), mk(self.nnz), mk(self.n)
        L.call("tc_operator_values", L.ptr(M), L.ptr(K), L.ptr(self.row_ptr), L.ptr(self.col),
               self.n, self.nnz, float(Chi), float(Cm), float(theta), float(dt), self.code,
               L.ptr(self.Mv), L.ptr(self.Kv), L.ptr(self.Av), L.ptr(self.Dinv), s)
        ns = (self.n + 63) // 64
        self.slice_ptr = torch.empty(ns + 1, dtype=torch.int32, device=dev)
        padded = C.c_int64(0)
        L.call("tc_sell_layout", L.ptr(self.row_ptr), self.n, L.ptr(self.slice_ptr), C.byref(padded), s)
        self.padded = int(padded.value)
        # slices whose rows share one stencil (a structured grid's interior)
        # are laid out on it: the hot passes then read one offset per entry
        # position instead of one per row
        self.pat = self.slice_tag = None
        self.n_shared = 0
        if self.int16_columns and self.share_slices:
            self.pat = torch.zeros(max(self.padded // 64, 1), dtype=torch.int32, device=dev)
            self.slice_tag = torch.empty(ns + 1, dtype=torch.int32, device=dev)
            L.call("tc_sell_share", L.ptr(self.row_ptr), L.ptr(self.col), self.n,
                   L.ptr(self.slice_ptr), L.ptr(self.pat), L.ptr(self.slice_tag), s)
            self.n_shared = int((self.slice_tag[:ns] & 1).sum().item())
            if self.n_shared == 0:
                self.pat = self.slice_tag = None
        self.scol = torch.empty(self.padded, dtype=torch.int32, device=dev)
        L.call("tc_sell_columns", L.ptr(self.row_ptr), L.ptr(self.col), self.n, L.ptr(self.slice_ptr),
               L.ptr(self.scol), L.ptr(self.pat), L.ptr(self.slice_tag), s)
        self.sM, self.sK, self.sA = (self.to_sell(v) for v in (self.Mv, self.Kv, self.Av))
        self.sell = L.Sell(self.n, ns, L.ptr(self.slice_ptr), L.ptr(self.scol), None, None, None)
        # int16 column offsets for the hot passes when the bandwidth allows
        self.dcol = torch.empty(self.padded, dtype=torch.int16, device=dev)
        mx = torch.empty(1, dtype=torch.int32, device=dev)
        L.call("tc_sell_deltas", C.byref(self.sell), L.ptr(self.dcol), L.ptr(mx), s)
        self.bandwidth = int(mx.item())
        if self.int16_columns and self.bandwidth <= 32767:
            self.sell.dcol = L.ptr(self.dcol)
            self.sell.pat = L.ptr(self.pat)
            self.sell.slice_tag = L.ptr(self.slice_tag)
        else:
            self.dcol = None
        self._make_pcg()

    def to_sell(self, vals):
        """Values (nnz, CSR order) in the system's SELL layout."""
        out = torch.empty(self.padded, dtype=vals.dtype, device=self.device)
        L.call("tc_sell_values", L.ptr(self.row_ptr), L.ptr(self.col), self.n, L.ptr(self.slice_ptr),
               L.ptr(vals), L.ptr(out), L.dtype_code(vals.dtype), L.ptr(self.pat), L.ptr(self.slice_tag),
               L.stream_handle(self.device))
        return out

    def _make_pcg(self):
        self._free_pcg()
        nb = C.c_size_t(0)
        L.call("tc_pcg_workspace_size", self.n, self.code, C.byref(nb))
        self.pcg_ws = torch.empty(nb.value, dtype=torch.uint8, device=self.device)
        h = C.c_void_p(0)
        L.call("tc_pcg_create", C.byref(h), C.byref(self.sell), L.ptr(self.sA), L.ptr(self.Dinv),
               self.code, L.ptr(self.pcg_ws), nb.value)
        self._pcg = h
        self._tune_plan(h)
        self.staged = self._stage_plan(h) if self.stage_gathers else False

    def _stage_plan(self, plan):
        """Staged gathers for the plan's spmv pass (operators without shared
        slices): each block's distinct columns, ascending, and every SELL
        entry's 16-bit position in its block's set, built on the host over
        the plan's block split (tc_sell_stage).  When a block of the tuned
        split references more than TC_STAGE_MAX_COLS columns, the split of
        8 slices per block (more, smaller blocks) is tried and, if it fits,
        becomes the plan's; rows sorted by length inside that split's block
        windows (``row_window``) try it first.  Meshes whose blocks hold more
        than 16 slices even at 1024 blocks stage each block's range in units
        of 8 slices (``stage_split``).  False when the operator has shared
        slices or no split fits (the plan then gathers per entry)."""
        import numpy as np
        from .. import autotune
        if self.slice_tag is not None:
            return False
        ns = (self.n + 63) // 64
        sp = self.slice_ptr.cpu().numpy()
        cl = self.scol.cpu().numpy()
        col_ptr = own = None
        split = self.stage_split(self.n)
        cands = [(int(self.pcg_blocks), 0), (autotune.pcg_blocks(self.n, 8), 0)]
        if self.row_window and self.row_window == self.sort_window(self.n):
            cands.insert(0, split)
        if split[1]:
            cands.append(split)
        if self.stage_override is not None:
            cands = [tuple(self.stage_override)]
        for G, chunk in dict.fromkeys(cands):
            per = -(-ns // G)
            units = G * (-(-per // chunk) if chunk else 1)
            col_ptr = np.empty(units + 1, np.int32)
            own = np.empty(units, np.int32)
            cols = np.empty(self.padded, np.int32)
            lidx = np.empty(self.padded, np.uint16)
            mx = C.c_int32(0)
            rc = L.lib.tc_sell_stage(self.n, ns, sp.ctypes.data, cl.ctypes.data, G, chunk,
                                     col_ptr.ctypes.data, own.ctypes.data, cols.ctypes.data,
                                     lidx.ctypes.data, C.byref(mx))
            if rc != L.TC_ERR_SIZE:
                break
        if rc == L.TC_ERR_SIZE:
            return False
        L.check(rc, "tc_sell_stage")
        if G != self.pcg_blocks:
            L.call("tc_pcg_set_blocks", plan, G)
            self.pcg_blocks = G
        dev = self.device
        self.st_col_ptr = torch.from_numpy(col_ptr).to(dev)
        self.st_own = torch.from_numpy(own).to(dev)
        self.st_cols = torch.from_numpy(cols[:int(col_ptr[units])].copy()).to(dev)
        self.st_lidx = torch.from_numpy(lidx.view(np.int16)).to(dev)
        st = L.Stage(G, int(mx.value), L.ptr(self.st_col_ptr), L.ptr(self.st_own),
                     L.ptr(self.st_cols), L.ptr(self.st_lidx), chunk)
        L.call("tc_pcg_set_stage", plan, C.byref(st))
        self.stage_max_cols = int(mx.value)
        self.stage_chunk = chunk
        return True

    def _tune_plan(self, plan):
        """Blocks of a PCG plan's passes from the tuning table (by rows)."""
        from .. import autotune
        spb = autotune.lookup("pcg_slices_per_block", self.n)
        self.pcg_blocks = autotune.pcg_blocks(self.n, spb)
        L.call("tc_pcg_set_blocks", plan, self.pcg_blocks)

    def _free_pcg(self):
        if getattr(self, "_pcg", None):
            L.lib.tc_pcg_destroy(self._pcg)
            self._pcg = None
        if getattr(self, "_pcg64", None):
            L.lib.tc_pcg_destroy(self._pcg64)
            self._pcg64 = None

    def pcg_plan_f64(self):
        """The PCG plan of a float64 solve on the working operator: for a
        float32 system, float64 copies of A's SELL values and of Dinv (both
        exact conversions), as the reference's ConjugateGradient converts A
        to float64 and keeps the Jacobi scaling in A's dtype, so that
        z = Dinv r multiplies by the working-dtype scaling (solver.py:24-25,
        preconditioner.py:9-43).  A float64 system's own plan otherwise."""
        if self.code == L.TC_F64:
            return self._pcg
        if getattr(self, "_pcg64", None) is None:
            self.sA64 = self.sA.to(torch.float64)
            self.Dinv64 = self.Dinv.to(torch.float64)
            nb = C.c_size_t(0)
            L.call("tc_pcg_workspace_size", self.n, L.TC_F64, C.byref(nb))
            self.pcg64_ws = torch.empty(nb.value, dtype=torch.uint8, device=self.device)
            h = C.c_void_p(0)
            L.call("tc_pcg_create", C.byref(h), C.byref(self.sell), L.ptr(self.sA64),
                   L.ptr(self.Dinv64), L.TC_F64, L.ptr(self.pcg64_ws), nb.value)
            self._pcg64 = h
            self._tune_plan(h)
        return self._pcg64

    def __del__(self):
        try:
            self._free_pcg()
        except Exception:
            pass

    # ---------------------------------------------------------- products
    def spmv(self, which, x, out=None):
        vals = {"M": self.sM, "K": self.sK, "A": self.sA}[which]
        out = torch.empty_like(x) if out is None else out
        L.call("tc_sell_spmv", C.byref(self.sell), L.ptr(vals), L.ptr(x), L.ptr(out), self.code,
               L.stream_handle(self.device))
        return out

    def rhs(self, w, u, coef, out):
        L.call("tc_rhs", C.byref(self.sell), L.ptr(self.sM), L.ptr(self.sK), L.ptr(w), L.ptr(u),
               float(coef), L.ptr(out), self.code, L.stream_handle(self.device))
        return out

    def pcg_solve(self, b, x, x_prev, x_next, a_tol, r_tol, max_iter, linear_guess, plan=None):
        """One solve on the working plan (vectors in the working dtype), or
        on ``plan`` (e.g. ``pcg_plan_f64()``, float64 vectors)."""
        it = C.c_int(0)
        L.call("tc_pcg_solve", self._pcg if plan is None else plan, L.ptr(b), L.ptr(x),
               L.ptr(x_prev), L.ptr(x_next), float(a_tol), float(r_tol), int(max_iter),
               1 if linear_guess else 0, C.byref(it), L.stream_handle(self.device))
        return int(it.value)

    def pcg_timing(self, on=None):
        """Per-kernel PCG timing: ``pcg_timing(True)`` starts (and resets) it;
        ``pcg_timing()`` returns {"init"|"spmv"|"update": (ms_total, launches)}."""
        if on is not None:
            L.call("tc_pcg_set_timing", self._pcg, 1 if on else 0)
            return None
        ms = (C.c_double * 3)()
        cnt = (C.c_int64 * 3)()
        L.call("tc_pcg_get_timing", self._pcg, ms, cnt)
        return {k: (ms[i], cnt[i]) for i, k in enumerate(("init", "spmv", "update"))}

    def csr(self, vals):
        """A torch sparse CSR tensor over this pattern."""
        return torch.sparse_csr_tensor(self.row_ptr.to(torch.int64), self.col.to(torch.int64), vals,
                                       size=(self.n, self.n))


def csr_system(row_ptr, col, n, device):
    """A ``P1System`` over a given CSR pattern (no mesh): for solving with an
    arbitrary sparse matrix through the same SELL/PCG kernels."""
    self = P1System.__new__(P1System)
    self.device = torch.device(device)
    self.n = int(n)
    self.row_ptr = row_ptr.to(device=self.device, dtype=torch.int32).contiguous()
    self.col = col.to(device=self.device, dtype=torch.int32).contiguous()
    self.nnz = int(self.col.numel())
    self._pcg = None
    return self

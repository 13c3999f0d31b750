/*
 * torchcor_amd — C ABI of the MI355X (gfx950) monodomain time-step library.
 *
 * Drop-in boundary for the reference's hot path (sanpinnawala/torchcor):
 * P1 FEM assembly into CSR, the Jacobi-PCG on A = Cm*M + theta*dt*K, and the
 * fused ionic updates.  The reference is pure Python/PyTorch, so these entry
 * points replace the PyTorch operations its classes call; each declaration
 * names the reference function it stands in for.
 *
 * Conventions
 *  - Every function returns 0 (TC_OK) or a negative tc_status; the text of
 *    the last error on the calling thread is tc_last_error().
 *  - Array arguments are DEVICE pointers owned by the caller (PyTorch
 *    tensors' data_ptr()), except where a name ends in _host.
 *  - `stream` is a hipStream_t passed as void*; work is enqueued
 *    asynchronously on it.  Only the calls marked "synchronizes" wait.
 *  - `dtype` selects the storage type of value/vector arrays:
 *    TC_F32 (float) or TC_F64 (double).  Reductions are always float64.
 *  - No function allocates device memory in the time-step calls; scratch
 *    space is passed in as a caller-owned workspace sized by the matching
 *    *_workspace_size function.
 */
#ifndef TORCHCOR_AMD_H
#define TORCHCOR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tc_status {
  TC_OK = 0,
  TC_ERR_ARG = -1,     /* invalid argument (null pointer, bad size/dtype) */
  TC_ERR_HIP = -2,     /* HIP runtime or launch error */
  TC_ERR_SIZE = -3,    /* workspace too small or size limit exceeded */
  TC_ERR_STATE = -4    /* object used in the wrong state */
};

enum tc_dtype { TC_F32 = 0, TC_F64 = 1 };

/* Text of the last error on this thread ("" if none). */
const char* tc_last_error(void);
/* ABI version of this library (bumped on any signature or descriptor layout
 * change). */
int tc_abi_version(void);

/* ------------------------------------------------------------------------
 * CSR pattern of the P1 operators.
 * Replaces the COO listing + coalesce of Matrices3D.assemble_matrices
 * (torchcor/core/assemble.py:217-248, and :67-96 for triangles): the entries
 * (elems[e][a], elems[e][b]) for a, b in 0..k-1 listed in the order
 * (a, b, e) are sorted by (row, col) with a stable sort, so duplicates keep
 * that order.  The result is the coalesced pattern with explicit zeros kept.
 * ---------------------------------------------------------------------- */
int tc_pattern_workspace_size(int64_t n_elems, int nodes_per_elem, size_t* bytes_host);
/* Sorts the k*k*E entries into `ws`; returns the number of distinct
 * (row, col) pairs in *nnz_host.  Synchronizes `stream`. */
int tc_pattern_sort(const int32_t* elems, int64_t n_elems, int nodes_per_elem,
                    int64_t n_nodes, void* ws, size_t ws_bytes, int64_t* nnz_host,
                    void* stream);
/* Writes row_ptr (n+1), col_idx (nnz), seg_ptr (nnz+1) and contrib (k*k*E):
 * entry s of the CSR sums the element contributions contrib[seg_ptr[s] ..
 * seg_ptr[s+1]-1], each an id (a*k+b)*E+e.  `ws` is the one given to
 * tc_pattern_sort. */
int tc_pattern_finish(int64_t n_elems, int nodes_per_elem, int64_t n_nodes, int64_t nnz,
                      const void* ws, int32_t* row_ptr, int32_t* col_idx,
                      int32_t* seg_ptr, int32_t* contrib, void* stream);

/* ------------------------------------------------------------------------
 * Element matrices and their sum into the CSR values.
 * Replaces Conductivity.calculate_sigma (core/conductivity.py:49-69) and
 * Matrices3D.construct_local_matrices / Matrices3DSurface.local_stiffness
 * (core/assemble.py:163-215, :99-145) plus the value half of
 * assemble_matrices (:217-248).  Local matrices are formed in float64 from
 * closed-form cofactors; with round_f32 != 0 each contribution is rounded to
 * float32 before the sum, as the reference's Python-list COO path does
 * (assemble.py:233-243).  Sums run in float64 in contrib order.
 * nodes (n,3) f64 in mm; elems (E,k) int32, k = 4 (tet) or 3 (surface tri).
 * The conductivity is either a full tensor per element, sigma (E,3,3) f64,
 * or (sigma == NULL) formed on the fly as st*I + (sl-st) f f^T from
 * sigma_l, sigma_t (E) f64 and fibres (E,3) f64.
 * local_ws holds 2*k*k*E f64.
 * ---------------------------------------------------------------------- */
int tc_assemble_p1(const double* nodes, const int32_t* elems, int64_t n_elems,
                   int nodes_per_elem, const double* sigma, const double* sigma_l,
                   const double* sigma_t, const double* fibres, const int32_t* seg_ptr,
                   const int32_t* contrib,
                   int64_t nnz, double* local_ws, int round_f32, double* M_vals,
                   double* K_vals, void* stream);

/* Working operator values in `dtype` (monodomain.py:93-100,
 * preconditioner.py:9-27): Kv = K/Chi, Mv = M, Av = Cm*Mv + theta*dt*Kv,
 * Dinv = 1/diag(Av). */
int tc_operator_values(const double* M_vals, const double* K_vals, const int32_t* row_ptr,
                       const int32_t* col_idx, int64_t n, int64_t nnz, double Chi,
                       double Cm, double theta, double dt, int dtype, void* Mv, void* Kv,
                       void* Av, void* Dinv, void* stream);

/* ------------------------------------------------------------------------
 * SELL-64 layout: rows in slices of 64 (one wavefront), each slice padded
 * to its longest row and stored column-major so that lane l of slice s reads
 * entry j of row 64*s+l at slice_ptr[s] + 64*j + l.  Padding entries hold
 * col = row and value 0.  The CSR stays the external format; SELL is the
 * internal layout of the SpMV kernels.
 *
 * Shared slices: when every row of a slice has its column offsets
 * (col - row) among those of the slice's longest row, the slice may be laid
 * out on that row's offsets (its pattern): entry j of every row then has
 * offset pat[slice_ptr[s]/64 + j], or is a padding entry (col = row, value
 * 0) where the row lacks that offset.  Each row keeps its entries in
 * column order, so every product sums the same terms in the same order.
 * The per-row columns stay complete; the PCG and RHS passes read the
 * pattern (one value per entry position instead of 64) on the slices
 * flagged in slice_tag.
 * ---------------------------------------------------------------------- */
typedef struct tc_sell {
  int64_t n;              /* rows */
  int64_t n_slices;       /* ceil(n/64) */
  const int32_t* slice_ptr; /* n_slices+1 entry offsets */
  const int32_t* col;     /* padded column indices */
  /* optional (NULL: unused): col - row as int16, from tc_sell_deltas when
   * every offset fits; the PCG and RHS passes then read these instead */
  const int16_t* dcol;
  /* optional, used with dcol (NULL: no shared slices): the patterns
   * (slice_ptr[n_slices]/64 int32 offsets, each within int16) and slice_ptr
   * with bit 0 set on the shared slices (n_slices+1; slice offsets are
   * multiples of 64) */
  const int32_t* pat;
  const int32_t* slice_tag;
} tc_sell;

/* Padded entry count; synchronizes `stream`.  slice_ptr (n_slices+1) is
 * filled as a side effect. */
int tc_sell_layout(const int32_t* row_ptr, int64_t n, int32_t* slice_ptr,
                   int64_t* padded_host, void* stream);
/* Shared-slice detection over the layout: slice s is shared when every
 * row's offsets lie among those of the slice's longest row and fit int16;
 * its offsets then go to pat (slice_ptr[n_slices]/64 int32) and
 * slice_tag[s] = slice_ptr[s] | 1 (else slice_ptr[s]; n_slices+1 entries).
 * Stream-ordered, does not synchronize. */
int tc_sell_share(const int32_t* row_ptr, const int32_t* col_idx, int64_t n,
                  const int32_t* slice_ptr, int32_t* pat, int32_t* slice_tag, void* stream);
/* Columns and values in the layout; with pat/slice_tag from tc_sell_share
 * (NULL: every slice per row) the shared slices are laid out on their
 * pattern (tc_sell_values then needs col_idx to place each value; it may
 * be NULL without a pattern). */
int tc_sell_columns(const int32_t* row_ptr, const int32_t* col_idx, int64_t n,
                    const int32_t* slice_ptr, int32_t* sell_col, const int32_t* pat,
                    const int32_t* slice_tag, void* stream);
int tc_sell_values(const int32_t* row_ptr, const int32_t* col_idx, int64_t n,
                   const int32_t* slice_ptr, const void* csr_vals, void* sell_vals, int dtype,
                   const int32_t* pat, const int32_t* slice_tag, void* stream);
/* dcol[e] = col[e] - row(e) as int16 for every padded entry; *max_abs_dev
 * (one device int32) receives the largest |col - row|, so dcol is usable
 * when it is <= 32767.  Stream-ordered, does not synchronize. */
int tc_sell_deltas(const tc_sell* A, int16_t* dcol, int32_t* max_abs_dev, void* stream);

/* ------------------------------------------------------------------------
 * Node reordering (host).
 * Replaces torchcor.core.reorder.RCM.calculate_rcm_order
 * (torchcor/core/reorder.py:15-38, scipy's reverse_cuthill_mckee over the
 * element pattern): a reverse Cuthill-McKee order of the symmetric pattern
 * row_ptr/col (HOST arrays, n+1 and nnz, diagonal entries allowed), started
 * per connected component from a pseudo-peripheral node.  order_host[i] is
 * the old index of the node placed at position i.  *bw_in_host / *bw_out_host
 * (either may be NULL) receive max |row - col| of the pattern before and
 * after.  No GPU work.
 * ---------------------------------------------------------------------- */
int tc_rcm_order(int64_t n, const int32_t* row_ptr_host, const int32_t* col_host,
                 int32_t* order_host, int64_t* bw_in_host, int64_t* bw_out_host);
/* Locality order for the SELL passes (no counterpart in the reference; the
 * simulator applies it inside, undoing it at the API): the Cuthill-McKee
 * level structure of tc_rcm_order, components and levels in order, each
 * level's nodes along a Z-order curve of xyz_host (n x 3 doubles).  Same
 * arguments and outputs as tc_rcm_order otherwise. */
int tc_level_order(int64_t n, const int32_t* row_ptr_host, const int32_t* col_host,
                   const double* xyz_host, int32_t* order_host, int64_t* bw_in_host,
                   int64_t* bw_out_host);

/* ------------------------------------------------------------------------
 * Mesh and stimulus files (host).
 * Replace the np.loadtxt parsing of MeshReader.read_nodes / read_elems /
 * read_fibres (torchcor/core/mesh.py:26-57) and load_stimulus_region
 * (torchcor/core/stimulation.py:27-36): a whitespace-separated numeric
 * table after `skip_lines` header lines (blank lines and '#' comments
 * ignored, every row as wide as the first), parsed by several host threads.
 * tc_table_shape gives its rows and columns; tc_table_read_* parse columns
 * [col0, col0 + ncols) of all `rows` rows into a row-major HOST array (rows
 * must match the file's count, else TC_ERR_ARG).  No GPU work.
 * ---------------------------------------------------------------------- */
int tc_table_shape(const char* path, int skip_lines, int64_t* rows_host, int32_t* cols_host);
int tc_table_read_f64(const char* path, int skip_lines, int col0, int ncols, int64_t rows,
                      double* out_host);
int tc_table_read_i64(const char* path, int skip_lines, int col0, int ncols, int64_t rows,
                      int64_t* out_host);
/* y = A x over a SELL operator (float64 accumulation), for tests and
 * callers of Monodomain.M/K/A products. */
int tc_sell_spmv(const tc_sell* A, const void* vals, const void* x, void* y, int dtype,
                 void* stream);

/* ------------------------------------------------------------------------
 * Right-hand side (monodomain.py:130-131):
 *   b = M w - coef * (K u),   coef = (1-theta)*dt.
 * M and K share the pattern of A.
 * ---------------------------------------------------------------------- */
int tc_rhs(const tc_sell* A, const void* Mv, const void* Kv, const void* w, const void* u,
           double coef, void* b, int dtype, void* stream);

/* ------------------------------------------------------------------------
 * Jacobi-PCG (core/solver.py:5-76, core/preconditioner.py:29-43).
 * A plan binds the operator (SELL pattern + A values + Dinv) and a
 * caller-owned workspace (scalars, partial sums, r, p, p', Ap).
 * ---------------------------------------------------------------------- */
typedef struct tc_pcg tc_pcg;
int tc_pcg_workspace_size(int64_t n, int dtype, size_t* bytes_host);
int tc_pcg_create(tc_pcg** plan, const tc_sell* A, const void* Av, const void* Dinv,
                  int dtype, void* ws, size_t ws_bytes);
int tc_pcg_destroy(tc_pcg* plan);
/* One solve of ConjugateGradient.solve (solver.py:40-76) with the reference
 * stopping rule: after each update stop when ||Dinv r|| < a_tol or
 * ||Dinv r|| / ||Dinv b|| < r_tol; *n_iter_host counts the iterations that
 * did not stop (0 when the first iteration converges).
 * With linear_guess != 0 the start is x_next = 2x - x_prev
 * (solver.py:35-38) and the solution is left in x_next (the caller then
 * treats x as the new x_prev and x_next as the new x); otherwise x is
 * updated in place and x_prev/x_next are not touched.  Synchronizes
 * `stream` (to return the iteration count). */
int tc_pcg_solve(tc_pcg* plan, const void* b, void* x, const void* x_prev, void* x_next,
                 double a_tol, double r_tol, int max_iter, int linear_guess,
                 int* n_iter_host, void* stream);
/* Per-kernel timing of tc_pcg_solve with HIP events on the solve's stream
 * (for measurement; off by default).  set_timing resets the totals;
 * get_timing returns milliseconds and launch counts for
 * [0] the start pass, [1] the SpMV pass, [2] the update pass, counting only
 * iterations that ran.  No counterpart in the reference. */
/* Blocks of the plan's PCG passes (a tuning choice: it groups the partial
 * sums, so results may differ in the last bits): 1 .. the default
 * (min(ceil(n/512), 1024), rounded down to a multiple of 8), or 0 for the
 * default.  A time-step runtime on the plan may be switched between
 * tc_sim_advance calls (its pipeline is drained then), not during one.
 * No counterpart in the reference. */
int tc_pcg_set_blocks(tc_pcg* plan, int blocks);

/* Staged gathers for operators whose slices do not share a stencil (an
 * unstructured mesh): the spmv pass of block v first forms p' once for
 * every distinct column its rows reference (cols[col_ptr[v] ..
 * col_ptr[v+1]), ascending) into LDS, then reads each SELL entry's p'
 * through its 16-bit position lidx in that set, instead of gathering
 * Dinv, r and p per entry.  Same products in the same order: bit-identical
 * results.  Block v owns slices [v*per, (v+1)*per), per = ceil(n_slices /
 * blocks), as the passes split them.  Replaces the gathers inside the
 * reference's CSR product A @ p (torchcor/core/solver.py:52,58); no
 * counterpart otherwise. */
typedef struct tc_stage {
  int blocks;               /* the plan's block count the stage was built for */
  int max_cols;             /* largest set of one unit (<= TC_STAGE_MAX_COLS) */
  const int32_t* col_ptr;   /* (units+1) device */
  const int32_t* own;       /* (units) position of the unit's first row in its set */
  const int32_t* cols;      /* (col_ptr[units]) device */
  const uint16_t* lidx;     /* (padded entries) device */
  /* Slices per unit: 0 = one unit per block (its whole slice range); else a
   * multiple of 8, at most 16: block v's slice range (per = ceil(n_slices /
   * blocks) slices from v * per) is staged in consecutive units of `chunk`
   * slices, one after the other (units = blocks * ceil(per / chunk), unit
   * v * ceil(per / chunk) + c for block v's c-th), so blocks holding more
   * slices than one set can stage still stage. */
  int chunk;
} tc_stage;
#define TC_STAGE_MAX_COLS 4096
/* Host: the stage of a SELL pattern (slice_ptr, col: host copies, per-row
 * int32 columns) split over `blocks` blocks, each staged in units of
 * `chunk` slices (0: one unit per block; as tc_stage.chunk).  cols_host
 * needs room for slice_ptr[n_slices] entries (an upper bound);
 * col_ptr_host (units+1), own_host (units), lidx_host
 * (slice_ptr[n_slices]).  *max_cols_host receives the largest set;
 * TC_ERR_SIZE when it exceeds TC_STAGE_MAX_COLS (the pattern is then left
 * unstaged). */
int tc_sell_stage(int64_t n, int64_t n_slices, const int32_t* slice_ptr_host,
                  const int32_t* col_host, int blocks, int chunk, int32_t* col_ptr_host,
                  int32_t* own_host, int32_t* cols_host, uint16_t* lidx_host,
                  int32_t* max_cols_host);
/* Attach a stage (device arrays owned by the caller; NULL detaches).  The
 * spmv passes of a time-step runtime on the plan (tc_sim_*) use it while
 * the plan's block count equals stage->blocks and its slices are not
 * shared; may be changed between tc_sim_advance calls. */
int tc_pcg_set_stage(tc_pcg* plan, const tc_stage* stage);
int tc_pcg_set_timing(tc_pcg* plan, int on);
int tc_pcg_get_timing(tc_pcg* plan, double* ms_host, int64_t* count_host);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code that must not call hipMalloc: tight serving loops,
 * harnesses that forbid allocation): the solve of tc_pcg_solve, same
 * kernels, same arguments and results, with no plan and no workspace.  The
 * operator is passed directly (A pattern, Av values, Dinv in `dtype`).  Its
 * scratch memory (control block, partial sums, r, p, p', Ap) is a static
 * device array of the library, 33,611,776 bytes, which every GPU the
 * library runs on holds until the process ends; the largest system it
 * takes has TC_PCG_STATIC_MAX_ROWS rows (larger ones fail with
 * TC_ERR_SIZE).  It allocates nothing (no hipMalloc/hipFree, no pinned
 * host memory).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two host threads, or two streams) race and must be
 * serialised by the caller.  Synchronizes `stream`. */
#define TC_PCG_STATIC_MAX_ROWS 1048576
int tc_pcg_solve_static(const tc_sell* A, const void* Av, const void* Dinv, int dtype,
                        const void* b, void* x, const void* x_prev, void* x_next,
                        double a_tol, double r_tol, int max_iter, int linear_guess,
                        int* n_iter_host, void* stream);

/* ------------------------------------------------------------------------
 * Ionic models.  State arrays are one (S, n) row-major block in `dtype`
 * (state k of node i at states[k*n + i]), in the order listed per model.
 * Tables are built on the device in float64 on the reference's float32
 * grids and stored in `dtype`.
 *
 * One step (the ionic half of Monodomain.step, monodomain.py:116-128, plus
 * the stimulus of Stimuli.apply, stimulation.py:51-63):
 *   dV   = -Iion(V, states)            (model.differentiate)
 *   w    = Cm*V + dt_rhs*(dV/100) + dt_rhs*(Istim/100)   when w != NULL
 *   Istim = sum over active stimuli k with node bit k set of amp[k]
 *   states advance by dt from the old values.
 * dV may be NULL; w may be NULL (then only dV and the states are written).
 * ---------------------------------------------------------------------- */
#define TC_MAX_STIMULI 32
/* One stimulus: on iff t/period <= count and start <= t % period <=
 * start + duration (stimulation.py:51-63), intensity amp (uA/cm^2). */
typedef struct tc_stimulus {
  double start, duration, period, count, amp;
} tc_stimulus;
#define TC_POOL_STIM 4

typedef struct tc_ionic_io {
  int64_t n;
  double dt;                 /* ionic model time step (ms): state advance */
  double dt_rhs;             /* monodomain time step (ms) used in w */
  double Cm;                 /* membrane capacitance for w (0.01) */
  const void* V;             /* (n) membrane potential, mV */
  void* w;                   /* (n) out, or NULL */
  void* dV;                  /* (n) out, or NULL */
  const uint32_t* stim_mask; /* (n) bit k = node in stimulus k, or NULL */
  uint32_t stim_active;      /* bit k = stimulus k is on this step */
  double stim_amp[TC_MAX_STIMULI]; /* intensity of stimulus k (uA/cm^2) */
  /* Activation/repolarisation bookkeeping folded into the same pass
   * (monodomain.py:174-176) for the PREVIOUS step's solution V at time
   * lat_t: AT[(V>0)&(AT==-1)] = lat_t; RT[(AT>0)&(RT==-1)&(V<-70)] = lat_t.
   * Disabled when AT == NULL. */
  void* AT;
  void* RT;
  double lat_t;
  /* When non-NULL and *gate != 0 at kernel start, the kernel does nothing
   * (used by the pipelined time-step runtime). */
  const int* gate;
  /* V is float64 (rounded to the state dtype on load): the float64 PCG
   * iterate of a float32 run. */
  int v_f64;
  /* w is stored in float64 and formed in float64 from the model's dV
   * (float32 runs in precise-RHS mode). */
  int w_f64;
  /* Batch layout (lanes > 0): the n entries are node-instances of one model
   * group of a batch, i = node * lanes + lane; V, w, dV, stim_mask, AT and
   * RT are rows of lane_stride entries per node and entry i lives at
   * node * lane_stride + lane0 + lane.  States stay (S, n).  lanes = 0:
   * one entry per node. */
  int32_t lanes, lane_stride, lane0;
  /* Pool layout (pool_slices != NULL, the served-stream engine): the launch
   * covers the 64-row slices listed in pool_slices (n = 64 * their count);
   * entry i is row pool_slices[i/64]*64 + i%64 of every per-row array, states
   * are (S, state_stride), and the system owning the slice supplies the valid
   * row range, the stimuli on this step (pool_sys_bits, amplitudes from its
   * TC_POOL_STIM-entry table pool_sys_stim) and the LAT/RT time (pool_sys_t,
   * once pool_sys_steps > 0).  stim_mask then holds each row's bits into its
   * system's stimulus table. */
  const int32_t* pool_slices;
  const int32_t* pool_slice_sys;
  const int64_t* pool_sys_row0;
  const int64_t* pool_sys_n;
  const double* pool_sys_t;
  const long long* pool_sys_steps;
  const uint32_t* pool_sys_bits;
  const struct tc_stimulus* pool_sys_stim;
  /* Row stride of the (S, stride) state block: the pool's, or in the plain
   * layout a padded one (0 = n); rows starting on 256-byte boundaries keep
   * every wavefront's state access to whole cache lines. */
  int64_t state_stride;
  /* Load and store the states non-temporally: they then pass through the
   * caches without displacing what is reused across steps.  Worth it when
   * the states share the last-level cache with large operators (the
   * time-step runtime, the served pool); a batch on a small shared mesh is
   * faster without. */
  int32_t stream_states;
  /* Non-zero: the states lie in blocks of 64 entries, entry i's state k at
   * (i / 64) * S * 64 + k * 64 + i % 64 (S = the model's state count;
   * state_stride is then unused; not with lanes).  A wavefront's S state
   * accesses then fall in one contiguous span instead of S rows far apart,
   * which keeps the memory's row buffers open (the time-step runtime). */
  int32_t state_block;
} tc_ionic_io;

/* ModifiedMS2v (ionic/ModifiedMS2v.py:4-44).  States: H. */
typedef struct tc_ms2v_params {
  double tau_in, tau_out, tau_open, tau_close, u_gate, u_crit, vmin, vmax;
} tc_ms2v_params;
int tc_ms2v_step(const tc_ms2v_params* p, void* states, const tc_ionic_io* io, int dtype,
                 void* stream);

/* TenTusscherPanfilov (ionic/TenTusscherPanfilov.py:6-454).  States:
 * CaSR CaSS Cai D F F2 FCaSS H J Ki M Nai R R_ S Xr1 Xr2 Xs.
 * Tables: V (32000 x 28: the 27 columns and a zero, 16-byte rows),
 * CaSS (999999 x 3), VEk (32000 x 1). */
typedef struct tc_ttp_params {
  int endo;                 /* 1: ENDO tau_S branch (TenTusscherPanfilov.py:258) */
  double GCaL, GKr, GKs, Gto;
  double Bufc, Bufsr, Bufss, CAPACITANCE, Cao, D_CaL_off, Fconst, GK1, GNa, GbCa, GbNa,
      GpCa, GpK, Kbufc, Kbufsr, Kbufss, KmCa, KmK, KmNa, KmNai, Ko, KpCa, Kup, Nao, Rconst,
      T, Vc, Vleak, Vmaxup, Vrel, Vsr, Vss, knaca, knak, ksat, n, pKNa, scl_tau_f, vHalfXs,
      xr2_off, EC, Vxfer, k1_, k2_, k3, k4, maxsr, minsr;
} tc_ttp_params;
int tc_ttp_tables(const tc_ttp_params* p, double dt, void* V_tab, void* CaSS_tab,
                  void* VEk_tab, int dtype, void* stream);
int tc_ttp_step(const tc_ttp_params* p, void* states, const void* V_tab,
                const void* CaSS_tab, const void* VEk_tab, const tc_ionic_io* io, int dtype,
                void* stream);

/* CourtemancheRamirezNattel (ionic/CourtemancheRamirezNattel.py:5-444).
 * States: Ca_rel Ca_up Cai Ki d f f_Ca h j m oa oi u ua ui v w xr xs.
 * Tables: Cai (100000 x 5), V (4000 x 32), fn (60000 x 3). */
typedef struct tc_crn_params {
  double C_B1c, C_B1d, C_B1e, C_Fn1, C_Fn2, C_dCa_rel, C_dCaup, F, K_Q10, K_up, KmCa,
      KmCmdn, KmCsqn, KmKo, KmNa3, KmNai, KmTrpn, Nai, R, T, Voli, gamma, k_rel, k_sat,
      maxCmdn, maxTrpn, tau_f_Ca, tau_tr, tau_u, ACh, Cao, GACh, GCaL, GK1, GKr, GKs, GNa,
      GbCa, GbNa, Gto, Ko, Nao, factorGKur, factorGrel, factorGtr, factorGup,
      factorhGate, factormGate, factoroaGate, factorxrGate, maxCaup, maxINaCa, maxINaK,
      maxIpCa, maxIup;
} tc_crn_params;
int tc_crn_tables(const tc_crn_params* p, double dt, void* Cai_tab, void* V_tab,
                  void* fn_tab, int dtype, void* stream);
int tc_crn_step(const tc_crn_params* p, void* states, const void* Cai_tab,
                const void* V_tab, const void* fn_tab, const tc_ionic_io* io, int dtype,
                void* stream);

/* ------------------------------------------------------------------------
 * Time-step runtime: the loop of Monodomain.solve (monodomain.py:141-196)
 * in native code.  Each step enqueues the fused ionic kernel, the fused
 * RHS + PCG start pass, a speculative number of PCG iterations (the most the
 * recent steps needed) and a one-block stopping check; the host runs up to
 * `depth` steps ahead and reads each step's status from pinned memory.  A
 * step that needs more iterations than were launched sets a device stall
 * flag that turns every later launch into a no-op; the host then launches
 * the missing iterations and re-issues the later steps, so results are those
 * of the plain loop.  Time is accumulated as a double (t += dt), stimulus
 * windows and LAT/RT use that t, as in the reference.
 * ---------------------------------------------------------------------- */
enum tc_model { TC_MODEL_MS2V = 0, TC_MODEL_TTP = 1, TC_MODEL_CRN = 2 };
typedef struct tc_sim_desc {
  int model;                 /* tc_model */
  const void* params;        /* tc_ms2v_params / tc_ttp_params / tc_crn_params */
  void* states;              /* (S, n) state block */
  const void* tab[3];        /* model tables in the order of its *_step call */
  tc_pcg* pcg;               /* plan over A: SELL pattern, A values, Dinv */
  const void* Mv;            /* SELL values of M */
  const void* Kv;            /* SELL values of K (already divided by Chi) */
  int dtype;
  int64_t n;
  double dt, dt_ionic, Cm, theta;
  int x_f64;                 /* iterate stored in float64 (float32 runs) */
  int w_f64;                 /* precise RHS: w in float64, b never rounded, K u
                                with the float64 iterate (needs x_f64) */
  void* x[2];                /* (n) iterate buffers (float64 when x_f64), current first */
  void* xp[2];               /* (n) x_prev buffers in dtype, current first */
  void* w;                   /* (n) scratch vector (float64 when w_f64) */
  void* AT;                  /* (n) activation times or NULL */
  void* RT;                  /* (n) repolarisation times */
  const uint32_t* stim_mask; /* (n) stimulus bits or NULL */
  int n_stim;
  tc_stimulus stim[TC_MAX_STIMULI];
  double a_tol, r_tol;
  int max_iter, linear_guess;
  int depth;                 /* steps enqueued ahead of the status read, >= 1 */
  int64_t state_stride;      /* row stride of the state block (0 = n) */
  int separate_check;        /* 0: each step's stopping test runs in its last update pass;
                                1: in a one-block pass of its own (same results) */
  int k_history;             /* K (PCG iterations launched per step before its status is
                                read) = the most any of the last k_history steps needed;
                                0 = 4, at most 256 (same results for any value) */
  int graph;                 /* 1: each step's RHS + PCG launches (fixed by the buffer
                                parity and the iteration count) are replayed as a captured
                                HIP graph; same results.  Steps with timing events launch
                                eagerly. */
  int state_block;           /* 1: states in blocks of 64 nodes (tc_ionic_io.state_block),
                                a (ceil(n/64), S, 64) array; state_stride unused */
} tc_sim_desc;
typedef struct tc_sim tc_sim;
int tc_sim_create(tc_sim** sim, const tc_sim_desc* desc);
int tc_sim_destroy(tc_sim* sim);
/* Run nsteps more steps.  iters_host (nsteps) receives each step's PCG
 * iteration count (reference convention).  When frames != NULL, after every
 * step whose global index is a multiple of ts_per_frame the solution (in
 * dtype) is copied to frames + k*n for the next frame k < max_frames.  Synchronizes
 * `stream` before returning. */
int tc_sim_advance(tc_sim* sim, int64_t nsteps, int32_t* iters_host, void* frames,
                   int64_t ts_per_frame, int64_t max_frames, void* stream);
/* As tc_sim_advance, with the frames in host memory: frames_host (pinned,
 * max_frames x n in dtype) receives frame k through device slot
 * k % ring_slots of `ring` (ring_slots x n in dtype, 1..16 slots), copied by
 * DMA on a copy stream of the runtime while the next steps run; a slot is
 * rewritten only after its previous copy ended.  Both streams are
 * synchronized before returning.  Replaces the reference's list of device
 * snapshots (torchcor/simulator/monodomain.py:162,186-187) when the frames
 * of a run should not stay in device memory. */
int tc_sim_advance_ring(tc_sim* sim, int64_t nsteps, int32_t* iters_host, void* frames_host,
                        void* ring, int ring_slots, int64_t ts_per_frame, int64_t max_frames,
                        void* stream);
/* Apply the LAT/RT test to the last step's solution (it is otherwise folded
 * into the next step's ionic pass).  Idempotent. */
int tc_sim_finish(tc_sim* sim, void* stream);
/* t (double), steps done, index 0..1 of desc.x holding the solution, frames
 * written. */
int tc_sim_status(tc_sim* sim, double* t_host, int64_t* steps_host, int* x_index_host,
                  int64_t* frames_host);
/* Per-kernel timing of the runtime with HIP events around each launch of
 * every `period`-th step (0 = off, the default; for measurement).  get
 * returns total ms and launches over the timed steps for
 * [0] ionic, [1] rhs+pcg start, [2] pcg spmv, [3] pcg update, [4] check,
 * counting only PCG iterations that ran. */
int tc_sim_set_timing(tc_sim* sim, int period);
int tc_sim_get_timing(tc_sim* sim, double* ms_host, int64_t* count_host);
/* Graph mode: graphs recorded and replayed so far.  No counterpart in the
 * reference (its loop launches every operation eagerly,
 * torchcor/simulator/monodomain.py:164-176). */
int tc_sim_graph_stats(const tc_sim* sim, int64_t* captures_host, int64_t* replays_host);

/* ------------------------------------------------------------------------
 * Batched simulations on one shared mesh (no counterpart in the reference:
 * many Monodomain runs, monodomain.py:141-226, advanced together).
 *
 * B <= 64 instances share the mesh and differ in conductivity scale, fibre
 * angle, stimulus sites and ionic model.  Their stiffness matrices are
 * combinations of NK shared component matrices (same CSR pattern as M):
 *   K_b = sum_k coef[k * lanes + b] K_k          (K_k already divided by Chi)
 *   A_b = Cm M + theta dt K_b, formed in float64 from the stored values.
 * Per-instance vectors are instance-minor: entry (node i, instance b) at
 * i * lanes + b, lanes = the smallest power of two >= B.  One wavefront
 * covers 64/lanes rows and every instance of them, so the pattern and the
 * shared values are read once for all instances (multi-right-hand-side
 * Jacobi-PCG, per-instance scalars and stopping tests: an instance that has
 * met the reference rule stops while the others iterate).
 * Model groups: instances b0 .. b0+count-1 run one ionic model whose states
 * are (S, n*count) with entry node * count + (b - b0).
 * ---------------------------------------------------------------------- */
#define TC_BATCH_MAX 64
#define TC_BATCH_MAX_K 5
#define TC_BATCH_MAX_GROUPS 8
typedef struct tc_batch_group {
  int model;                 /* tc_model */
  const void* params;        /* tc_*_params */
  void* states;              /* (S, state_stride) block of n * count columns */
  const void* tab[3];        /* model tables */
  int b0, count;
  int64_t state_stride;      /* row stride of the state block (0 = n * count) */
} tc_batch_group;
typedef struct tc_batch_desc {
  int64_t n, nnz;
  int B, lanes, NK;
  int dtype;
  int x_f64, w_f64;          /* as in tc_sim_desc */
  const int32_t* row_ptr;    /* (n+1) CSR pattern */
  const int32_t* col;        /* (nnz) */
  const void* Mv;            /* (nnz) M values, dtype */
  const void* Kv[TC_BATCH_MAX_K]; /* (nnz) component values, dtype */
  const double* coef;        /* (NK, lanes) device */
  double dt, dt_ionic, Cm, theta;
  void* x[2];                /* (n, lanes) iterate buffers, current first */
  void* xp[2];               /* (n, lanes) x_prev buffers in dtype */
  void* w;                   /* (n, lanes) scratch (float64 when w_f64) */
  void* Dinv;                /* (n, lanes) dtype, filled by tc_batch_create */
  void* AT;                  /* (n, lanes) or NULL */
  void* RT;
  const uint32_t* stim_mask; /* (n, lanes) bit k: stimulus k applies there */
  int n_stim;
  tc_stimulus stim[TC_MAX_STIMULI];
  int n_groups;
  tc_batch_group groups[TC_BATCH_MAX_GROUPS];
  double a_tol, r_tol;
  int max_iter, linear_guess;
  int depth;                 /* steps enqueued ahead of the status read */
  void* ws;                  /* workspace of tc_batch_workspace_size bytes */
  size_t ws_bytes;
  int max_blocks;            /* cap on the blocks of the PCG passes, 8..1024, or 0 for
                                1024 (a tuning choice; it groups the partial sums) */
  int graph;                 /* 1: each step's RHS + PCG launches (fixed by the buffer
                                parity and the iteration count) are replayed as a captured
                                HIP graph; same results */
} tc_batch_desc;
typedef struct tc_batch tc_batch;
int tc_batch_workspace_size(int64_t n, int lanes, int dtype, size_t* bytes_host);
/* Validates the description, fills Dinv, clears the control block. */
int tc_batch_create(tc_batch** batch, const tc_batch_desc* desc, void* stream);
int tc_batch_destroy(tc_batch* batch);
/* nsteps time steps; iters[k] (host) = the largest n_iter over the
 * instances in step k.  Synchronizes at the end. */
int tc_batch_advance(tc_batch* batch, int64_t nsteps, int32_t* iters_host, void* stream);
/* Graph mode: graphs recorded and replayed so far.  No counterpart in the
 * reference (its loop launches every operation eagerly,
 * torchcor/simulator/monodomain.py:164-176). */
int tc_batch_graph_stats(const tc_batch* batch, int64_t* captures_host, int64_t* replays_host);
/* LAT/RT test of the last solution; per-instance sums of n_iter (B). */
int tc_batch_finish(tc_batch* batch, int64_t* total_iters_host, void* stream);
/* t, steps done, index 0..1 of desc.x holding the solution. */
int tc_batch_status(tc_batch* batch, double* t_host, int64_t* steps_host, int* x_index_host);

/* ------------------------------------------------------------------------
 * Served stream (no counterpart in the reference: independent Monodomain
 * runs, monodomain.py:141-226, on different meshes, admitted and retired
 * while the others run).
 *
 * A device pool holds every admitted system: a contiguous range of 64-row
 * slices (rows = nodes padded to whole slices) and of SELL entries in
 * pool-wide arrays whose column indices are pool rows (block-offset batched
 * SELL), pool-wide per-row vectors and one ionic state block per model.  All
 * live systems advance together, one time step per engine step, with their
 * own PCG scalars, stopping tests, clocks and stimuli.  Admission and
 * retirement happen between tc_serve_advance calls.
 * ---------------------------------------------------------------------- */
typedef struct tc_serve_desc {
  int dtype, x_f64, w_f64;   /* as tc_sim_desc */
  int64_t cap_slices;        /* pool rows = 64 * cap_slices */
  int64_t cap_entries;       /* pool SELL entries */
  int max_sys;               /* live systems at most (1..1024) */
  int32_t* slice_off;        /* (cap_slices) first entry of each slice */
  int32_t* slice_w;          /* (cap_slices) padded width */
  int32_t* slice_sys;        /* (cap_slices) owning system */
  int32_t* col;              /* (cap_entries) pool rows */
  int16_t* dcol;             /* (cap_entries) col - row, or NULL (int32 only) */
  int32_t* pat;              /* (cap_entries/64) shared-slice patterns by entry
                                offset / 64, or NULL (per-row columns only) */
  void* Mv;                  /* (cap_entries) dtype: M, K/Chi, A values */
  void* Kv;
  void* Av;
  void* Dinv;                /* (rows) dtype */
  void* x[2];                /* (rows) iterate (float64 when x_f64) */
  void* xp[2];               /* (rows) dtype */
  void* w;                   /* (rows) float64 when w_f64 */
  void* r;                   /* (rows) dtype */
  void* p[2];
  void* Ap;
  void* AT;                  /* (rows) dtype, or NULL */
  void* RT;
  uint32_t* stim_mask;       /* (rows) bits into the row's system stimuli */
  void* states[3];           /* per tc_model: (cap_slices, S, 64) dtype (each slice's
                                S states of its 64 rows together), NULL if unused */
  const void* params[3];     /* per tc_model: tc_*_params */
  const void* tab[3][3];     /* per tc_model: tables */
  double dt, dt_ionic, Cm, theta, a_tol, r_tol;
  int max_iter, linear_guess, depth;
  void* ws;                  /* tc_serve_workspace_size bytes */
  size_t ws_bytes;
  int k_history;             /* as tc_sim_desc: K = the most any of the last k_history
                                engine steps needed (0 = 4, at most 256) */
  int graph;                 /* 1: each engine step's ionic, RHS + PCG launches (fixed by
                                the live systems, the buffer parity and the iteration
                                count) are replayed as a captured HIP graph, recorded anew
                                after every admission or retirement; same results */
} tc_serve_desc;
typedef struct tc_serve_system {
  int64_t n;                 /* nodes */
  int64_t n_entries;         /* padded SELL entries (slice_ptr[ceil(n/64)]) */
  const int32_t* slice_ptr;  /* the system's SELL-64 operator (tc_sell_*) */
  const int32_t* col;        /* its columns (0 .. n-1) */
  const int16_t* dcol;       /* its col - row as int16 (tc_sell_deltas), or NULL */
  const int32_t* pat;        /* its shared-slice patterns and tagged slice table */
  const int32_t* slice_tag;  /* (tc_sell_share), or NULL */
  const void* Mv;            /* SELL values of M, K/Chi, A and Dinv, dtype */
  const void* Kv;
  const void* Av;
  const void* Dinv;
  int model;                 /* tc_model */
  int n_states;
  const void* states;        /* (n_states, n) initial states */
  const void* V0;            /* (n) initial potential */
  const uint32_t* stim_mask; /* (n) bit k: stimulus k applies */
  int n_stim;
  tc_stimulus stim[TC_POOL_STIM];
  int lat;                   /* track AT/RT */
  int slices_per_block;      /* 64-row slices per block of this system's passes: 8..64, or 0
                                for the default 16 (a tuning choice; a system's results depend
                                on it, never on the other systems in the pool) */
} tc_serve_system;
typedef struct tc_serve tc_serve;
int tc_serve_workspace_size(int64_t cap_slices, int max_sys, size_t* bytes_host);
int tc_serve_create(tc_serve** serve, const tc_serve_desc* desc, void* stream);
int tc_serve_destroy(tc_serve* serve);
/* *fits_host = 1 when a system of that size can be admitted now. */
int tc_serve_fits(tc_serve* serve, int64_t n_slices, int64_t n_entries, int* fits_host);
/* Copies the system into the pool (synchronizes); *slot_host = its id. */
int tc_serve_add(tc_serve* serve, const tc_serve_system* sys, int* slot_host, void* stream);
/* Final LAT/RT test, copies V (and AT, RT) of the system to (n) device
 * buffers, returns its n_iter sum and steps done, frees its ranges. */
int tc_serve_remove(tc_serve* serve, int slot, void* V_out, void* AT_out, void* RT_out,
                    int64_t* total_iters_host, int64_t* steps_host, void* stream);
/* nsteps engine steps of every live system; iters[k] (host) = the largest
 * n_iter among them in step k.  Synchronizes at the end. */
int tc_serve_advance(tc_serve* serve, int64_t nsteps, int32_t* iters_host, void* stream);
/* Engine steps since creation, and how many of them needed more PCG
 * iterations than the pipeline had launched (each drains the pipeline). */
int tc_serve_stats(const tc_serve* serve, int64_t* steps_host, int64_t* stalls_host);
/* Graph mode: graphs recorded and replayed so far.  No counterpart in the
 * reference. */
int tc_serve_graph_stats(const tc_serve* serve, int64_t* captures_host, int64_t* replays_host);

/* Device copy of nbytes (a multiple of 16, 16-byte aligned buffers), one
 * 16-byte load and store per thread: the bandwidth reference of the
 * benchmark.  No counterpart in the reference. */
int tc_stream_copy(const void* src, void* dst, int64_t nbytes, void* stream);

/* LAT/RT update alone (monodomain.py:174-176), for the last step of a run. */
int tc_lat_update(const void* u, void* AT, void* RT, int64_t n, double t, int dtype,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TORCHCOR_AMD_H */

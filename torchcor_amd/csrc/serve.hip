// Served stream: many monodomain runs on different meshes, admitted and
// retired while the others keep running (continuous batching).
//
// Device pool.  Every admitted system owns a contiguous range of 64-row
// slices (its rows, padded to whole slices) and a contiguous range of SELL
// entries in pool-wide arrays: a block-offset batched SELL whose column
// indices are pool rows.  Per-row vectors (iterate, x_prev, w, r, p, Ap,
// Dinv, AT, RT, stimulus bits) and per-model ionic state blocks are pool-wide
// too, so a system's state stays resident from admission to retirement.
// Ranges are handed out first-fit; nothing moves while a system is live.
//
// Passes run over a list of the live systems' slices (system by system): one
// wavefront per slice, one thread per row, the slice's system giving the
// PCG scalars, the stopping test and the valid row range.  A block never
// spans two systems; each block writes its partial sums, and every block of
// the next pass adds its own system's block partials in block order before
// it starts (no separate fold launch; the grouping depends only on the
// system's size, so a run is bit-identical alone or in any mix).  A system
// that has met the stopping rule is inactive for the rest of the step.
// Ionic kernels take per-model slice lists (pool mode of tc_ionic_io).
//
// The time loop is the pipelined one of the single-run runtime: ionic ->
// RHS + PCG start -> K speculative iterations -> check, several steps ahead;
// the check pass advances every live system's clock and stimulus bits and
// publishes the step status in pinned host memory; a step that needs more
// iterations stalls the pipeline until the host has finished it.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "pcg_kernels.h"

namespace tc {
namespace {

constexpr int kSB = 512;            // threads per block of the passes
constexpr int kSW = kSB / 64;
constexpr int kSRing = 64;
constexpr int kSHistoryDefault = 4, kSHistoryMax = 256;

// Per-system device state (arrays of max_sys), plus the pipeline gate.
struct SvCtl {
  int stall, iter;
  int n_active;  // systems still iterating in the current step
  int pad;
  double a_tol, r_tol;
  double* rz2;          // r.z per system, two iterations (2 * max_sys)
  double* bnorm;
  double* alpha;
  double* beta;
  int* active;
  int* n_iter;
  long long* total;
  double* t;            // time of the last finished step (t += dt, as the reference)
  long long* steps;     // steps finished
  int64_t* row0;        // first pool row
  int64_t* n;           // rows (nodes)
  uint32_t* bits;       // stimuli on in the next step
  int* nstim;
  int* use16;           // the system's entries have int16 column offsets
  tc_stimulus* stim;    // TC_POOL_STIM per system
};

// Block -> contiguous range of the slice list; blocks of one XCD adjacent.
struct LRange {
  int64_t l0, l1;
};
__device__ inline LRange block_list(int64_t nl) {
  const int G = gridDim.x;
  const int b = blockIdx.x;
  int vb = b;
  if (G % kXcd == 0) vb = (b % kXcd) * (G / kXcd) + b / kXcd;
  const int64_t per = (nl + G - 1) / G;
  LRange r;
  r.l0 = (int64_t)vb * per;
  r.l1 = r.l0 + per < nl ? r.l0 + per : nl;
  return r;
}


// Blocks never span two systems: a system of s slices gets
// min(ceil(s / kSPer), kSBlkMax) blocks of equal slice counts, so its
// partial sums group the same way whatever else is in the pool (a run gives
// bit-identical results alone or in a mix), and a fold reads at most
// kSBlkMax partials (one per lane).
constexpr int kSPer = 16;       // default slices per block
constexpr int kSPerMin = 8;     // smallest per-system choice (tc_serve_system.slices_per_block)
constexpr int kSPerMax = 64;
constexpr int kSBlkMax = 1024;

// One block's work, formed on the host when the lists are rebuilt: a
// contiguous range of the slice list, all of one system, with the system
// fields the passes need (one scalar load instead of a chain of lookups).
struct SvBlk {
  int32_t l0, l1;       // list range
  int32_t q, sys;       // live-system index (-1: padding block) and slot
  int64_t row_end;      // first pool row past the system
  int32_t use16;        // the system's entries have int16 column offsets
  int32_t vb0, vb1;     // the system's block range
  int32_t row0;         // first pool row of the system
};

// One entry of the slice list: the slice, its SELL position and width, and
// whether it is laid out on a shared pattern (at pat[off / 64]).
struct SvEnt {
  int32_t sl, off, w, shared;
};

// Live system q: its slot and block range.
struct SvSys {
  int32_t slot, vb0, vb1, pad;
};

struct SvPool {
  const SvBlk* blk;         // per block
  const SvEnt* ent;         // per list entry
  const SvSys* sysd;        // per live system
  const int32_t* list;      // slices of the live systems, system by system
  int64_t nl;
  const int32_t* slice_off; // first entry of each slice
  const int32_t* slice_w;   // padded width of each slice
  const int32_t* slice_sys;
  const int32_t* col;       // pool rows
  const int16_t* dcol;      // col - row, for systems with use16
  const int32_t* pat;       // shared-slice patterns by entry offset / 64, or NULL
};

// Column source of list entry e of a block: per-row columns, or (int16
// systems) the slice's shared pattern clamped to the system's own rows.
__device__ inline Cols<int32_t> sv_cols(const int32_t* cols, const int32_t*, const SvEnt& e,
                                        int lane, const SvBlk&) {
  return per_row(cols + e.off + lane);
}

__device__ inline Cols<int16_t> sv_cols(const int16_t* cols, const int32_t* pat, const SvEnt& e,
                                        int lane, const SvBlk& bk) {
  return Cols<int16_t>{cols + e.off + lane, e.shared ? pat + (e.off >> 6) : nullptr, bk.row0,
                       (int32_t)(bk.row_end - 1)};
}

// Partial sums are per block (one system each) at part[v * nv + k]; system
// q's total is the sum over its blocks in block order, which every consumer
// evaluates itself (one wavefront, lanes striding over the blocks).
struct BlockView {
  int vb;
  int64_t l0, l1;  // list range
  int q;           // system index, -1 for a padding block
};


// Sum of the block partials [v0, v1] (at most kSBlkMax blocks, nv values
// each) with every thread of the block, in the same order in every block:
// each thread issues its loads (clamped index, then a select) before adding
// them in order.  Result in every thread.
__device__ inline void fold_range_block(const double* __restrict__ P, int nv, int v0, int v1,
                                        double& a, double& b) {
  __shared__ double fs[kSW][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int R = kSBlkMax / kSB;
  double x0[R], x1[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int v = v0 + (int)threadIdx.x + q * kSB;
    const int vc = v <= v1 ? v : v1;
    const double t0 = P[(int64_t)vc * nv];
    const double t1 = nv == 2 ? P[(int64_t)vc * nv + 1] : 0.0;
    x0[q] = v <= v1 ? t0 : 0.0;
    x1[q] = v <= v1 ? t1 : 0.0;
  }
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    s0 += x0[q];
    s1 += x1[q];
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (lane == 0) {
    fs[wv][0] = s0;
    fs[wv][1] = s1;
  }
  __syncthreads();
  a = 0.0;
  b = 0.0;
  for (int w = 0; w < kSW; ++w) {
    a += fs[w][0];
    b += fs[w][1];
  }
  __syncthreads();
}

// Position of this block in the block order (blocks of one XCD adjacent).
__device__ inline int sv_vblock() {
  const int G = gridDim.x, b = blockIdx.x;
  return G % kXcd == 0 ? (b % kXcd) * (G / kXcd) + b / kXcd : b;
}

// Block partials: each lane's sums over the slices of its wavefront, a
// wavefront sum, then the wavefronts in order.
__device__ inline void store_block_partials(const BlockView& v, double r0, double r1, int nv,
                                            double* P) {
  __shared__ double ws[kSW][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  r0 = wave_sum(r0);
  if (nv == 2) r1 = wave_sum(r1);
  if (lane == 0) {
    ws[wv][0] = r0;
    ws[wv][1] = r1;
  }
  __syncthreads();
  if (threadIdx.x == 0 && v.q >= 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < kSW; ++w) {
      s0 += ws[w][0];
      s1 += ws[w][1];
    }
    P[(int64_t)v.vb * nv] = s0;
    if (nv == 2) P[(int64_t)v.vb * nv + 1] = s1;
  }
}

// Rows of one block of the fused RHS + PCG start pass.  The first chunk of
// each wavefront's first slice is fetched before the barrier that publishes
// the stall flag (*s_stop), so the flag's latency overlaps those loads.
template <class T, class XT, class WT, bool LG, class CI>
__device__ inline void sv_rhs_rows(const SvBlk& bk, const SvEnt* __restrict__ ent,
                                   const CI* __restrict__ cols, const int32_t* __restrict__ pat,
                                   const T* __restrict__ Mv,
                                   const T* __restrict__ Kv, const T* __restrict__ Dinv,
                                   const WT* __restrict__ w, const XT* __restrict__ x,
                                   const T* __restrict__ xp, XT* __restrict__ xnext,
                                   T* __restrict__ xpnext, T* __restrict__ r, double coef,
                                   double Cm, double tdt, const int* s_stop, double& red0,
                                   double& red1) {
  const int lane = threadIdx.x & 63, wv = wave_id();
  int64_t li = bk.l0 + wv;
  RhsChunk<T, XT, WT, LG, CI> ch;
  SvEnt e{0, 0, 0, 0};
  if (li < bk.l1) {
    e = ent[li];
    if (e.w > 0)
      ch.fetch(sv_cols(cols, pat, e, lane, bk), Mv + e.off + lane, Kv + e.off + lane, 0, e.w,
               (int64_t)e.sl * 64 + lane, x, w, xp);
  }
  __syncthreads();
  if (*s_stop) return;
  bool first = true;
  for (; li < bk.l1; li += kSW) {
    if (!first) e = ent[li];
    const int64_t i = (int64_t)e.sl * 64 + lane;
    const Cols<CI> c = sv_cols(cols, pat, e, lane, bk);
    const T* m = Mv + e.off + lane;
    const T* kk = Kv + e.off + lane;
    double am = 0.0, ak = 0.0, mx = 0.0, kx = 0.0;
    for (int j0 = 0; j0 < e.w; j0 += kChunk) {
      if (!(first && j0 == 0)) ch.fetch(c, m, kk, j0, e.w, i, x, w, xp);
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        if (j0 + u >= e.w) break;
        double x0;
        if (LG) x0 = (double)(XT)(2.0 * (double)ch.xj[u] - (double)ch.pj[u]);
        else x0 = (double)ch.xj[u];
        const double mj = (double)ch.mv[u], kj = (double)ch.kv[u];
        am += mj * (double)ch.wj[u];
        if (sizeof(WT) == 8) ak += kj * (double)ch.xj[u];
        else ak += kj * (double)(T)ch.xj[u];
        mx += mj * x0;
        kx += kj * x0;
      }
    }
    first = false;
    if (i < bk.row_end) {
      if (LG) {
        xnext[i] = (XT)(2.0 * (double)x[i] - (double)xp[i]);
        xpnext[i] = (T)x[i];
      }
      double bi = am - coef * ak;
      if (sizeof(WT) != 8) bi = (double)(T)bi;
      const double d = (double)Dinv[i];
      const T ri = (T)(bi - ax0(Cm, mx, tdt, kx));
      r[i] = ri;
      red0 += (double)ri * (d * (double)ri);
      const double db = d * bi;
      red1 += db * db;
    }
  }
}

// Fused RHS + PCG start over the slice list (as k_rhs_init); block partials
// of r.z and (Dinv b)^2.
template <class T, class XT, class WT, bool LG>
__global__ void __launch_bounds__(kSB)
k_sv_rhs_init(SvPool pl, const SvEnt* __restrict__ ent, const int32_t* __restrict__ pat,
              const int16_t* __restrict__ dcol, const int32_t* __restrict__ col, int nsys,
              const T* __restrict__ Mv, const T* __restrict__ Kv,
              const T* __restrict__ Dinv, const WT* __restrict__ w, const XT* __restrict__ x,
              const T* __restrict__ xp, XT* __restrict__ xnext, T* __restrict__ xpnext,
              T* __restrict__ r, double coef, double Cm, double tdt, SvCtl* ctl,
              double* P0) {
  __shared__ int s_stop;
  const int vb = sv_vblock();
  const SvBlk bk = pl.blk[vb];
  if (threadIdx.x == 0) s_stop = ctl->stall;
  double red0 = 0.0, red1 = 0.0;
  if (bk.use16)
    sv_rhs_rows<T, XT, WT, LG, int16_t>(bk, ent, dcol, pat, Mv, Kv, Dinv, w, x, xp, xnext,
                                        xpnext, r, coef, Cm, tdt, &s_stop, red0, red1);
  else
    sv_rhs_rows<T, XT, WT, LG, int32_t>(bk, ent, col, nullptr, Mv, Kv, Dinv, w, x, xp, xnext,
                                        xpnext, r, coef, Cm, tdt, &s_stop, red0, red1);
  if (s_stop) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->n_active = nsys;
  store_block_partials(BlockView{vb, bk.l0, bk.l1, bk.q}, red0, red1, 2, P0);
}

// Rows of one block of the spmv pass of iteration k.  First the scalars of
// the block's system: k = 0 takes r.z and ||Dinv b|| from the start pass;
// k > 0 applies the stopping test of iteration k-1 (solver.py:57-74) to the
// last update's sums, then beta.  The system's first block records the
// outcome.  Then p' = beta p + Dinv r on the fly, Ap', block partials of
// p'.Ap'.  The flags, the partial sums and the first chunk of each
// wavefront's first slice are all loaded before the first wait.
template <class T, class CI>
__device__ inline void sv_spmv_block(const SvBlk& bk, int vb, const SvEnt* __restrict__ ent,
                                     const int32_t* __restrict__ pat,
                                     const CI* __restrict__ cols, const T* __restrict__ Av,
                                     const T* __restrict__ Dinv, const T* __restrict__ r,
                                     const T* __restrict__ pin, T* __restrict__ pout,
                                     T* __restrict__ Ap, int k, int max_sys, SvCtl* ctl,
                                     double* rz2, double* bnorm, int* active, int* n_iter,
                                     const double* __restrict__ P0,
                                     const double* __restrict__ P2, double* P1) {
  __shared__ int s_stop, s_pre;
  __shared__ double s_sc[4];  // ||Dinv b||, previous r.z, a_tol, r_tol
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int sys = bk.sys;
  // one read of the flags for the whole block (another block may clear the
  // system's flag during this pass; the branches below must be uniform):
  // thread 0 stages them in LDS before the fold's barrier, and every thread
  // then takes the same decisions from the same values
  if (threadIdx.x == 0) {
    s_stop = ctl->stall != 0 || (k > 0 && ctl->n_active == 0);
    s_pre = bk.q >= 0 ? active[sys] : 0;
    double bn = 1.0, rzp = 1.0;
    if (k > 0 && bk.q >= 0) {
      bn = bnorm[sys];
      rzp = rz2[((k - 1) & 1) * max_sys + sys];
    }
    s_sc[0] = bn;
    s_sc[1] = rzp;
    s_sc[2] = ctl->a_tol;
    s_sc[3] = ctl->r_tol;
  }
  int64_t li = bk.l0 + wv;
  SpmvChunk<T, CI> ch;
  SvEnt e{0, 0, 0, 0};
  if (li < bk.l1) {
    e = ent[li];
    if (e.w > 0)
      ch.fetch(sv_cols(cols, pat, e, lane, bk), Av + e.off + lane, 0, e.w, (int64_t)e.sl * 64 + lane,
               Dinv, r, pin, k);
  }
  double a = 0.0, b = 0.0;
  if (bk.q >= 0) fold_range_block(k == 0 ? P0 : P2, 2, bk.vb0, bk.vb1, a, b);
  else __syncthreads();
  if (s_stop || bk.q < 0) return;
  const bool lead = bk.vb0 == vb && threadIdx.x == 0;  // the system's first block records
  double beta = 0.0;
  if (k == 0) {
    if (lead) {
      rz2[sys] = a;
      bnorm[sys] = sqrt(b);
      active[sys] = 1;
      n_iter[sys] = 0;
    }
  } else {
    if (!s_pre) return;
    const double zn = sqrt(a);
    if (zn < s_sc[2] || zn / s_sc[0] < s_sc[3]) {
      if (lead) {
        active[sys] = 0;
        n_iter[sys] = k - 1;
        atomicSub(&ctl->n_active, 1);
      }
      return;
    }
    beta = b / s_sc[1];
    if (lead) {
      rz2[(k & 1) * max_sys + sys] = b;
      n_iter[sys] = k;
    }
  }
  double red = 0.0;
  bool first = true;
  for (; li < bk.l1; li += kSW) {
    if (!first) e = ent[li];
    const int64_t i = (int64_t)e.sl * 64 + lane;
    const Cols<CI> c = sv_cols(cols, pat, e, lane, bk);
    const T* av = Av + e.off + lane;
    double acc = 0.0;
    for (int j0 = 0; j0 < e.w; j0 += kChunk) {
      if (!(first && j0 == 0)) ch.fetch(c, av, j0, e.w, i, Dinv, r, pin, k);
      acc = ch.add(acc, j0, e.w, k, beta);
    }
    first = false;
    if (i < bk.row_end) {
      T pi;
      if (k == 0) pi = (T)((double)Dinv[i] * (double)r[i]);
      else pi = (T)pdir(beta, (double)pin[i], (double)Dinv[i], (double)r[i]);
      const T api = (T)acc;
      pout[i] = pi;
      Ap[i] = api;
      red += (double)pi * (double)api;
    }
  }
  store_block_partials(BlockView{vb, bk.l0, bk.l1, bk.q}, red, 0.0, 1, P1);
}

template <class T>
__global__ void __launch_bounds__(kSB)
k_sv_spmv(SvPool pl, const SvEnt* __restrict__ ent, const int32_t* __restrict__ pat,
          const int16_t* __restrict__ dcol, const int32_t* __restrict__ col,
          const T* __restrict__ Av, const T* __restrict__ Dinv,
          const T* __restrict__ r, const T* __restrict__ pin, T* __restrict__ pout,
          T* __restrict__ Ap, int k, int max_sys, SvCtl* ctl, double* rz2, double* bnorm,
          int* active, int* n_iter, const double* __restrict__ P0,
          const double* __restrict__ P2, double* P1) {
  const int vb = sv_vblock();
  const SvBlk bk = pl.blk[vb];
  if (bk.use16)
    sv_spmv_block<T, int16_t>(bk, vb, ent, pat, dcol, Av, Dinv, r, pin, pout, Ap, k, max_sys,
                              ctl, rz2, bnorm, active, n_iter, P0, P2, P1);
  else
    sv_spmv_block<T, int32_t>(bk, vb, ent, nullptr, col, Av, Dinv, r, pin, pout, Ap, k, max_sys,
                              ctl, rz2, bnorm, active, n_iter, P0, P2, P1);
}

// update pass of iteration k: alpha = rz / p'.Ap' per system, x += alpha p',
// r -= alpha Ap', block partials of z.z and r.z.  The first slice's operands,
// the flags, r.z and the partial sums load before the first wait.
template <class T, class XT>
__device__ inline void sv_update_block(const SvBlk& bk, int vb, const SvEnt* __restrict__ ent,
                                       const T* __restrict__ Dinv, const T* __restrict__ p,
                                       const T* __restrict__ Ap, XT* __restrict__ x,
                                       T* __restrict__ r, int k, int max_sys, SvCtl* ctl,
                                       const double* rz2, const int* active,
                                       const double* __restrict__ P1, double* P2) {
  __shared__ int s_stop, s_pre;
  const int lane = threadIdx.x & 63, wv = wave_id();
  if (threadIdx.x == 0) {
    s_stop = ctl->stall != 0 || ctl->n_active == 0;
    s_pre = bk.q >= 0 ? active[bk.sys] : 0;
  }
  const double rzk = bk.q >= 0 ? rz2[(k & 1) * max_sys + bk.sys] : 1.0;
  int64_t li = bk.l0 + wv;
  int64_t i = 0;
  XT xv = XT(0);
  T pv = T(0), av = T(0), rv = T(0), dv = T(0);
  if (li < bk.l1) {
    i = (int64_t)ent[li].sl * 64 + lane;
    if (i < bk.row_end) {
      xv = x[i];
      pv = p[i];
      av = Ap[i];
      rv = r[i];
      dv = Dinv[i];
    }
  }
  double a = 0.0, b = 0.0;
  if (bk.q >= 0) fold_range_block(P1, 1, bk.vb0, bk.vb1, a, b);
  else __syncthreads();
  if (s_stop || bk.q < 0 || !s_pre) return;
  const double alpha = rzk / a;
  double red0 = 0.0, red1 = 0.0;
  while (li < bk.l1) {
    if (i < bk.row_end) {
      const XT xi = (XT)((double)xv + alpha * (double)pv);
      const T ri = (T)((double)rv - alpha * (double)av);
      x[i] = xi;
      r[i] = ri;
      const double z = (double)dv * (double)ri;
      red0 += z * z;
      red1 += (double)ri * z;
    }
    li += kSW;
    if (li < bk.l1) {
      i = (int64_t)ent[li].sl * 64 + lane;
      if (i < bk.row_end) {
        xv = x[i];
        pv = p[i];
        av = Ap[i];
        rv = r[i];
        dv = Dinv[i];
      }
    }
  }
  store_block_partials(BlockView{vb, bk.l0, bk.l1, bk.q}, red0, red1, 2, P2);
}

template <class T, class XT>
__global__ void __launch_bounds__(kSB)
k_sv_update(SvPool pl, const SvEnt* __restrict__ ent, const T* __restrict__ Dinv,
            const T* __restrict__ p,
            const T* __restrict__ Ap, XT* __restrict__ x, T* __restrict__ r, int k, int max_sys,
            SvCtl* ctl, const double* rz2, const int* active, const double* __restrict__ P1,
            double* P2) {
  const int vb = sv_vblock();
  sv_update_block<T, XT>(pl.blk[vb], vb, ent, Dinv, p, Ap, x, r, k, max_sys, ctl, rz2, active,
                         P1, P2);
}

// The list entries' slice positions (after the lists change).
__global__ void k_sv_ent(int64_t nl, const int32_t* __restrict__ list,
                         const int32_t* __restrict__ slice_off,
                         const int32_t* __restrict__ slice_w, SvEnt* ent) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nl) return;
  const int32_t sl = list[li];
  const int32_t o = slice_off[sl];
  ent[li] = SvEnt{sl, o & ~1, slice_w[sl], o & 1};
}

// End of a launch chunk after k_end iterations.  The stopping test of the
// last iteration for every system still active; systems still iterating
// below max_iter stall the pipeline; otherwise the step is final: every live
// system's n_iter sum, clock (t += dt) and next-step stimulus bits advance,
// and the status goes to pinned host memory.  hc is the host copy of *ctl:
// only its per-system array pointers and tolerances, fixed when the pool is
// created, are read from it (a kernel-argument load instead of a load of the
// pointer from device memory before every array access).
//
// One wavefront per system over ceil(nsys / 16) blocks (one block looping
// over at most kSCheckOneMax systems).  Each wavefront
// stores its system's results (SvChk) and each block its aggregates (stall,
// largest n_iter and iteration count) at device scope; every block waits
// until its stores are complete and counts itself; the block that counts
// last publishes the status and commits the per-system results, so the
// pass costs one round of per-system latency whatever the number of
// systems.  No block waits for another.
constexpr int kSCheck = 1024;     // threads of a check block
constexpr int kSCheckWaves = kSCheck / 64;
constexpr int kSCheckSys = 1024;  // systems of one check pass (max_sys is checked against it)
constexpr int kSCheckBlocks = kSCheckSys / kSCheckWaves;
// Up to this many systems one block runs the pass, its wavefronts taking
// the systems in turn: two rounds of per-system latency cost less than the
// device-scope completion count of several blocks.
constexpr int kSCheckOneMax = 2 * kSCheckWaves;
static_assert(kSCheckBlocks <= 64, "one wavefront folds the block aggregates");

struct SvChk {
  long long tot, stp;
  double t;
  int slot, act, ni;
  uint32_t bits;
};

struct SvChkCtl {
  int any[kSCheckBlocks], nmax[kSCheckBlocks], imax[kSCheckBlocks];
  unsigned done;
};

template <class U>
__device__ inline void st_dev(U* p, U v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class U>
__device__ inline U ld_dev(const U* p) {
  return __hip_atomic_load(const_cast<U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool ONE>
__global__ void __launch_bounds__(kSCheck)
k_sv_check(SvPool pl, int k_end, int nsys, int max_iter, double dt, SvCtl* ctl, SvCtl hc,
           const double* __restrict__ P2, StatusRec* status, long long seq, SvChk* chk,
           SvChkCtl* cc) {
  static_assert((TC_POOL_STIM & (TC_POOL_STIM - 1)) == 0 && TC_POOL_STIM <= 64, "stimulus lanes");
  __shared__ int stall0, any_stall, nmax, imax, last;
  __shared__ SvChk s_chk[kSCheckOneMax];  // the results when one block covers every system
  constexpr bool one = ONE;  // a single block covering every system
  if (threadIdx.x == 0) {
    stall0 = ctl->stall;
    any_stall = 0;
    nmax = 0;
    imax = 0;
  }
  __syncthreads();
  // A gated pass (the pipeline stalled before it) returns in every block:
  // the flag is the same in all of them, since only the last block of this
  // pass writes it, after every block has read it and counted itself.
  if (stall0) return;
  const int lane = threadIdx.x & 63, wv = wave_id();
  // phase 1, one wavefront per system: the stopping test of the last
  // iteration and the system's clock and next-step stimulus bits (lane k
  // tests stimulus k).  Once the system's descriptor is in, every other load
  // of it (flags, norms, clock, stimulus table and block partials; each
  // lane's at once, at clamped in-range addresses) is issued before the
  // first wait.
  const int q_end = one ? nsys : min(nsys, (int)(blockIdx.x + 1) * kSCheckWaves);
  for (int q = blockIdx.x * kSCheckWaves + wv; q < q_end; q += kSCheckWaves) {
    const SvSys y = pl.sysd[q];
    const int s = y.slot;
    const int act = hc.active[s];
    const double bn = hc.bnorm[s];
    const int ni_old = hc.n_iter[s];
    const long long tot = hc.total[s];
    const double t_old = hc.t[s];
    const long long stp = hc.steps[s];
    const int nst = hc.nstim[s];
    const tc_stimulus sk = hc.stim[(int64_t)s * TC_POOL_STIM + (lane & (TC_POOL_STIM - 1))];
    // The block partials of the system, loaded as fold_range_block's
    // threads hold them: thread t = w * 64 + lane of a kSB-thread block
    // holds blocks vb0 + t + j * kSB, j < kSBlkMax / kSB.
    constexpr int R = kSBlkMax / kSB;
    double x0[kSW][R];
#pragma unroll
    for (int w = 0; w < kSW; ++w)
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int v = y.vb0 + w * 64 + lane + j * kSB;
        const double t = P2[(int64_t)(v <= y.vb1 ? v : y.vb1) * 2];
        x0[w][j] = v <= y.vb1 ? t : 0.0;
      }
    // the clock after this step (t += dt, as the reference) and the
    // stimuli on in the next one (the test of stimulus_bits, per lane)
    const double t = t_old + dt;
    const double tn = t + dt;
    bool on = false;
    if (lane < nst) {
      const double p_t = fmod(tn, sk.period);
      const double t_t = tn / sk.period;
      on = t_t <= sk.count && p_t >= sk.start && p_t <= sk.start + sk.duration;
    }
    const uint32_t bits = (uint32_t)__ballot(on);
    // ||z||^2 summed in exactly the order of fold_range_block (each
    // thread's blocks in order, the wavefront sum, then the wavefronts in
    // order), so the stopping test at the end of a launch chunk sees the
    // bits the next spmv pass would see: where a chunk ends (which depends
    // on the other systems in the pool) cannot change a system's result.
    double a = 0.0;
#pragma unroll
    for (int w = 0; w < kSW; ++w) {
      double sw = 0.0;
#pragma unroll
      for (int j = 0; j < R; ++j) sw += x0[w][j];
      a += wave_sum(sw);
    }
    if (lane == 0) {
      int nact = act, ni = ni_old;
      if (act) {
        const double zn = sqrt(a);
        if (zn < hc.a_tol || zn / bn < hc.r_tol) {
          nact = 0;
          ni = k_end - 1;
        } else {
          ni = k_end;
          if (k_end < max_iter) atomicOr(&any_stall, 1);
          else nact = 0;  // max_iter reached
        }
      }
      if constexpr (one) {
        s_chk[q] = SvChk{tot + (long long)ni, stp + 1, t, s, nact | (act << 1), ni, bits};
      } else {
        SvChk* c = chk + q;
        st_dev(&c->slot, s);
        st_dev(&c->act, nact | (act << 1));
        st_dev(&c->ni, ni);
        st_dev(&c->bits, bits);
        st_dev(&c->tot, tot + (long long)ni);
        st_dev(&c->t, t);
        st_dev(&c->stp, stp + 1);
      }
      atomicMax(&nmax, ni);
      atomicMax(&imax, nact ? k_end : (ni + 1 < k_end ? ni + 1 : k_end));
    }
  }
  // every wavefront's stores complete before its block counts itself
  if constexpr (!one) __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if constexpr (one) {
    // one block covers every system: it is the last, with the results in
    // its own memory
    if (threadIdx.x == 0) {
      if (any_stall) ctl->stall = 1;
      if (status)
        put_status(status, (uint32_t)nmax, (uint32_t)imax | (any_stall ? kStStalled : 0u), seq);
    }
    if (threadIdx.x < nsys) {
      const SvChk& c = s_chk[threadIdx.x];
      if (c.act >> 1) {
        hc.active[c.slot] = c.act & 1;
        hc.n_iter[c.slot] = c.ni;
      }
      if (!any_stall) {
        hc.total[c.slot] = c.tot;
        hc.t[c.slot] = c.t;
        hc.steps[c.slot] = c.stp;
        hc.bits[c.slot] = c.bits;
      }
    }
    return;
  }
  if (threadIdx.x == 0) {
    st_dev(&cc->any[blockIdx.x], any_stall);
    st_dev(&cc->nmax[blockIdx.x], nmax);
    st_dev(&cc->imax[blockIdx.x], imax);
    __builtin_amdgcn_s_waitcnt(0);
    last = __hip_atomic_fetch_add(&cc->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
           gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  // phase 2 in the block that counted last: the step status (the next
  // kernels and host copies are ordered after this kernel on the stream)
  // and, one thread per system, the flags, iteration counts and, when no
  // system stalls, the clock and bits
  if (wv == 0) {
    const bool in = lane < (int)gridDim.x;
    int any = in ? ld_dev(&cc->any[lane]) : 0;
    int nm = in ? ld_dev(&cc->nmax[lane]) : 0;
    int im = in ? ld_dev(&cc->imax[lane]) : 0;
    for (int o = 32; o > 0; o >>= 1) {
      any |= __shfl_xor(any, o);
      nm = max(nm, __shfl_xor(nm, o));
      im = max(im, __shfl_xor(im, o));
    }
    if (lane == 0) {
      st_dev(&cc->done, 0u);
      any_stall = any;
      if (any) ctl->stall = 1;
      if (status)
        put_status(status, (uint32_t)nm, (uint32_t)im | (any ? kStStalled : 0u), seq);
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nsys; r += blockDim.x) {
    const SvChk* c = chk + r;
    const int s = ld_dev(&c->slot);
    const int a = ld_dev(&c->act);
    if (a >> 1) {
      hc.active[s] = a & 1;
      hc.n_iter[s] = ld_dev(&c->ni);
    }
    if (!any_stall) {
      hc.total[s] = ld_dev(&c->tot);
      hc.t[s] = ld_dev(&c->t);
      hc.steps[s] = ld_dev(&c->stp);
      hc.bits[s] = ld_dev(&c->bits);
    }
  }
}

// Copy a system's SELL operator and initial state into its pool ranges.
template <class T, class XT>
__global__ void k_sv_place_rows(int64_t row0, int64_t n, int64_t nrows, const T* __restrict__ V0,
                                const T* __restrict__ Dinv_s, const uint32_t* __restrict__ mask_s,
                                XT* x0, XT* x1, T* xp0, T* xp1, void* w, int w_f64, T* Dinv,
                                T* r, T* p0, T* p1, T* Ap, T* AT, T* RT, uint32_t* mask) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrows) return;
  const int64_t i = row0 + q;
  const bool ok = q < n;
  const T v = ok ? V0[q] : T(0);
  x0[i] = (XT)v;
  x1[i] = (XT)v;
  xp0[i] = v;
  xp1[i] = v;
  if (w_f64) ((double*)w)[i] = 0.0;
  else ((T*)w)[i] = T(0);
  Dinv[i] = ok ? Dinv_s[q] : T(0);
  r[i] = p0[i] = p1[i] = Ap[i] = T(0);
  if (AT) {
    AT[i] = T(-1);
    RT[i] = T(-1);
  }
  mask[i] = (ok && mask_s) ? mask_s[q] : 0u;
}

template <class T>
__global__ void k_sv_place_states(int64_t row0, int64_t n, int64_t nrows, int S,
                                  const T* __restrict__ st_s, T* st) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrows) return;
  // the pool's states lie in blocks of one slice: row's state k at
  // (row / 64) * S * 64 + k * 64 + row % 64
  const int64_t row = row0 + q;
  T* b = st + (row >> 6) * S * 64 + (row & 63);
  for (int k = 0; k < S; ++k) b[k * 64] = q < n ? st_s[k * n + q] : T(0);
}

template <class T>
__global__ void k_sv_place_entries(int64_t ne, int64_t entry0, int64_t row0,
                                   const int32_t* __restrict__ col_s,
                                   const int16_t* __restrict__ dcol_s, const T* __restrict__ M_s,
                                   const T* __restrict__ K_s, const T* __restrict__ A_s,
                                   int32_t* col, int16_t* dcol, T* M, T* K, T* A) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  col[entry0 + e] = (int32_t)(row0 + col_s[e]);
  if (dcol && dcol_s) dcol[entry0 + e] = dcol_s[e];
  M[entry0 + e] = M_s[e];
  K[entry0 + e] = K_s[e];
  A[entry0 + e] = A_s[e];
}

__global__ void k_sv_place_slices(int64_t ns, int64_t slice0, int64_t entry0, int sys,
                                  const int32_t* __restrict__ sptr_s,
                                  const int32_t* __restrict__ tag_s, int32_t* slice_off,
                                  int32_t* slice_w, int32_t* slice_sys) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  // bit 0: the slice is laid out on a shared pattern (offsets are multiples of 64)
  slice_off[slice0 + s] = (int32_t)(entry0 + sptr_s[s]) | (tag_s ? (tag_s[s] & 1) : 0);
  slice_w[slice0 + s] = (sptr_s[s + 1] - sptr_s[s]) / 64;
  slice_sys[slice0 + s] = sys;
}

template <class T, class XT>
__global__ void k_sv_extract(int64_t row0, int64_t n, const XT* __restrict__ x,
                             const T* __restrict__ AT, const T* __restrict__ RT, T* V_out,
                             T* AT_out, T* RT_out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  V_out[q] = (T)x[row0 + q];
  if (AT_out) {
    AT_out[q] = AT[row0 + q];
    RT_out[q] = RT[row0 + q];
  }
}

// Final LAT/RT test of a retiring system (its last step's solution).
template <class T, class XT>
__global__ void k_sv_lat_final(int64_t row0, int64_t n, const XT* __restrict__ x, T* AT, T* RT,
                               double t) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int64_t i = row0 + q;
  const T V = (T)x[i];
  T at = AT[i];
  if (V > T(0) && at == T(-1)) {
    at = (T)t;
    AT[i] = at;
  }
  if (at > T(0) && V < T(-70) && RT[i] == T(-1)) RT[i] = (T)t;
}

// First-fit allocator of [0, cap).
struct RangeAlloc {
  std::map<int64_t, int64_t> free_;  // start -> length
  void reset(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  int64_t take(int64_t len) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        const int64_t s = it->first, rest = it->second - len;
        free_.erase(it);
        if (rest > 0) free_[s + len] = rest;
        return s;
      }
    }
    return -1;
  }
  void give(int64_t s, int64_t len) {
    auto it = free_.emplace(s, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t largest() const {
    int64_t m = 0;
    for (auto& kv : free_) m = kv.second > m ? kv.second : m;
    return m;
  }
};

struct SysHost {
  bool live = false;
  int model = 0;
  int64_t slice0 = 0, ns = 0, entry0 = 0, ne = 0, n = 0;
  bool lat = false;
  bool use16 = false;
  int spb = kSPer;  // slices per block of this system's passes
};

struct SPlan {
  int slot, K;
  long long seq;
};

}  // namespace
}  // namespace tc

struct tc_serve {
  tc_serve_desc d;
  union {
    tc_ms2v_params ms2v;
    tc_ttp_params ttp;
    tc_crn_params crn;
  } mp[3];
  tc::SvCtl hctl;       // host copy of the device pointers
  tc::SvCtl* ctl;       // device
  double* part[3];      // block partials: start (2), spmv (1), update (2)
  int32_t* list;        // device slice list (all live systems)
  int32_t* mlist[3];    // device slice lists per model
  tc::SvBlk* blk_desc;  // device: per block
  tc::SvEnt* ent;       // device: per list entry
  tc::SvSys* sysd;      // device: per live system
  tc::SvChk* chk;       // device: the check pass's per-system results
  tc::SvChkCtl* chkctl; // device: its per-block aggregates and completion count
  int64_t nl, nml[3];
  int nsys, G;
  std::vector<tc::SysHost> sys;
  tc::RangeAlloc slices, entries;
  int xc, pc;
  tc::StatusRec* host_status;  // pinned, written by the check pass
  tc::StatusRec* dev_view;     // device address of host_status
  tc::StepStatus status[tc::kSRing];  // records as read by sv_wait
  long long seq;
  std::deque<int> recent;
  int history;  // engine steps whose iteration counts set K
  int64_t steps;
  int64_t stalls = 0;    // steps that needed more iterations than were launched
  tc::GraphCache graphs; // graph mode: each engine step's fixed launches
};

namespace tc {
namespace {

struct SvLayout {
  size_t ctl, arrays, chk, chkctl, part[3], list, mlist[3], blk_desc, ent, sysd, total;
};

SvLayout sv_layout(int64_t cap_slices, int max_sys) {
  SvLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align_up(bytes, 256);
    return o;
  };
  const size_t ms = (size_t)max_sys;
  L.ctl = take(sizeof(SvCtl));
  // rz2 (2 per system) bnorm alpha beta t (double), total steps (ll), row0 n (i64),
  // active n_iter nstim (int), bits (u32), stim
  L.arrays = take(ms * (6 * 8 + 2 * 8 + 2 * 8 + 4 * 4 + 4) + ms * TC_POOL_STIM * sizeof(tc_stimulus));
  // the check pass's per-system results and per-block aggregates (zeroed
  // with the arrays: its completion count starts at 0)
  L.chk = take(ms * sizeof(SvChk));
  L.chkctl = take(sizeof(SvChkCtl));
  // blocks (and partial slots): sum over systems of ceil(slices / spb), padded to 8
  const size_t slots = (size_t)cap_slices / kSPerMin + ms + 16;
  L.part[0] = take(slots * 2 * sizeof(double));
  L.part[1] = take(slots * sizeof(double));
  L.part[2] = take(slots * 2 * sizeof(double));
  L.list = take((size_t)cap_slices * 4);
  for (int m = 0; m < 3; ++m) L.mlist[m] = take((size_t)cap_slices * 4);
  L.blk_desc = take(slots * sizeof(SvBlk));
  L.ent = take((size_t)cap_slices * sizeof(SvEnt));
  L.sysd = take(ms * sizeof(SvSys));
  L.total = off;
  return L;
}

SvPool sv_pool(const tc_serve* S) {
  SvPool p;
  p.blk = S->blk_desc;
  p.ent = S->ent;
  p.sysd = S->sysd;
  p.list = S->list;
  p.nl = S->nl;
  p.slice_off = S->d.slice_off;
  p.slice_w = S->d.slice_w;
  p.slice_sys = S->d.slice_sys;
  p.col = S->d.col;
  p.dcol = S->d.dcol;
  p.pat = S->d.pat;
  return p;
}

// Rebuild the slice lists, the blocks (spb slices of one system each) and
// the per-system block ranges (pipeline drained).
int sv_rebuild(tc_serve* S, hipStream_t s) {
  S->graphs.clear();  // their launches carry the old lists, blocks and counts
  std::vector<int32_t> list, ml[3], sl, vb0, vb1, bl0, bq;
  for (int q = 0; q < (int)S->sys.size(); ++q) {
    const SysHost& h = S->sys[q];
    if (!h.live) continue;
    const int qi = (int)sl.size();
    sl.push_back(q);
    vb0.push_back((int32_t)bq.size());
    int64_t nb = ceil_div(h.ns, h.spb);
    if (nb > kSBlkMax) nb = kSBlkMax;
    const int64_t per = ceil_div(h.ns, nb);
    for (int64_t k = 0; k < h.ns; ++k) {
      if (k % per == 0) {
        bl0.push_back((int32_t)list.size());
        bq.push_back(qi);
      }
      list.push_back((int32_t)(h.slice0 + k));
      ml[h.model].push_back((int32_t)(h.slice0 + k));
    }
    vb1.push_back((int32_t)bq.size() - 1);
  }
  // padding blocks (empty) up to a multiple of the XCD count
  while (bq.empty() || bq.size() % kXcd != 0) {
    bl0.push_back((int32_t)list.size());
    bq.push_back(-1);
  }
  bl0.push_back((int32_t)list.size());
  S->nl = (int64_t)list.size();
  S->nsys = (int)sl.size();
  S->G = (int)bq.size();
  for (int m = 0; m < 3; ++m) S->nml[m] = (int64_t)ml[m].size();
  auto up = [&](int32_t* dst, const std::vector<int32_t>& v) -> int {
    if (!v.empty())
      TC_HIP(hipMemcpyAsync(dst, v.data(), v.size() * 4, hipMemcpyHostToDevice, s));
    return TC_OK;
  };
  std::vector<SvBlk> bd(bq.size());
  for (size_t v = 0; v < bq.size(); ++v) {
    SvBlk& k = bd[v];
    k.l0 = bl0[v];
    k.l1 = bl0[v + 1];
    k.q = bq[v];
    k.sys = k.q >= 0 ? sl[k.q] : 0;
    const SysHost* h = k.q >= 0 ? &S->sys[k.sys] : nullptr;
    k.row_end = h ? h->slice0 * 64 + h->n : 0;
    k.use16 = h && h->use16 ? 1 : 0;
    k.vb0 = k.q >= 0 ? vb0[k.q] : 0;
    k.vb1 = k.q >= 0 ? vb1[k.q] : -1;
    k.row0 = h ? (int32_t)(h->slice0 * 64) : 0;
  }
  int rc;
  if ((rc = up(S->list, list))) return rc;
  for (int m = 0; m < 3; ++m)
    if ((rc = up(S->mlist[m], ml[m]))) return rc;
  TC_HIP(hipMemcpyAsync(S->blk_desc, bd.data(), bd.size() * sizeof(SvBlk), hipMemcpyHostToDevice,
                        s));
  std::vector<SvSys> sd(sl.size());
  for (size_t q = 0; q < sl.size(); ++q) sd[q] = SvSys{sl[q], vb0[q], vb1[q], 0};
  if (!sd.empty())
    TC_HIP(hipMemcpyAsync(S->sysd, sd.data(), sd.size() * sizeof(SvSys), hipMemcpyHostToDevice,
                          s));
  if (S->nl > 0) {
    hipLaunchKernelGGL(k_sv_ent, dim3((int)ceil_div(S->nl, 256)), dim3(256), 0, s, S->nl, S->list,
                       S->d.slice_off, S->d.slice_w, S->ent);
    TC_LAUNCH();
  }
  TC_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
  return TC_OK;
}

tc_ionic_io sv_ionic_io(tc_serve* S, int m) {
  const tc_serve_desc& d = S->d;
  tc_ionic_io io{};
  io.n = S->nml[m] * 64;
  io.dt = d.dt_ionic;
  io.dt_rhs = d.dt;
  io.Cm = d.Cm;
  io.V = d.x[S->xc];
  io.v_f64 = (d.x_f64 && d.dtype == TC_F32) ? 1 : 0;
  io.w_f64 = (d.w_f64 && d.dtype == TC_F32) ? 1 : 0;
  io.w = d.w;
  io.stim_mask = d.stim_mask;
  io.AT = d.AT;
  io.RT = d.RT;
  io.gate = &S->ctl->stall;
  io.stream_states = 1;
  io.pool_slices = S->mlist[m];
  io.pool_slice_sys = d.slice_sys;
  io.pool_sys_row0 = S->hctl.row0;
  io.pool_sys_n = S->hctl.n;
  io.pool_sys_t = S->hctl.t;
  io.pool_sys_steps = S->hctl.steps;
  io.pool_sys_bits = S->hctl.bits;
  io.pool_sys_stim = S->hctl.stim;
  io.state_stride = d.cap_slices * 64;   // the per-row arrays' extent
  io.state_block = 1;                     // states in blocks of one slice
  return io;
}

// The ionic step of every live system: the three models' slice lists in one
// launch.
int sv_ionic(tc_serve* S, hipStream_t s) {
  const tc_serve_desc& d = S->d;
  tc_ionic_io io[3];
  void* st[3];
  for (int m = 0; m < 3; ++m) {
    io[m] = sv_ionic_io(S, m);
    st[m] = d.states[m];
  }
  const void* ttab[3] = {d.tab[TC_MODEL_TTP][0], d.tab[TC_MODEL_TTP][1], d.tab[TC_MODEL_TTP][2]};
  const void* ctab[3] = {d.tab[TC_MODEL_CRN][0], d.tab[TC_MODEL_CRN][1], d.tab[TC_MODEL_CRN][2]};
  return pool_ionic_step(S->nml[TC_MODEL_MS2V] ? &S->mp[TC_MODEL_MS2V].ms2v : nullptr,
                         S->nml[TC_MODEL_TTP] ? &S->mp[TC_MODEL_TTP].ttp : nullptr,
                         S->nml[TC_MODEL_CRN] ? &S->mp[TC_MODEL_CRN].crn : nullptr, st, ttab,
                         ctab, io, d.dtype, s);
}

template <class T, class XT>
int sv_iters(tc_serve* S, int k0, int k1, XT* xs, hipStream_t s) {
  const tc_serve_desc& d = S->d;
  const SvPool pl = sv_pool(S);
  const dim3 g(S->G), blk(kSB);
  T* pp[2] = {(T*)d.p[0], (T*)d.p[1]};
  for (int k = k0; k < k1; ++k) {
    T* pin = pp[k & 1];
    T* pout = pp[(k + 1) & 1];
    hipLaunchKernelGGL(k_sv_spmv<T>, g, blk, 0, s, pl, (const SvEnt*)pl.ent,
                       (const int32_t*)pl.pat, (const int16_t*)pl.dcol, (const int32_t*)pl.col,
                       (const T*)d.Av, (const T*)d.Dinv,
                       (const T*)d.r, (const T*)pin, pout, (T*)d.Ap, k, d.max_sys, S->ctl,
                       S->hctl.rz2, S->hctl.bnorm, S->hctl.active, S->hctl.n_iter,
                       (const double*)S->part[0], (const double*)S->part[2], S->part[1]);
    TC_LAUNCH();
    hipLaunchKernelGGL((k_sv_update<T, XT>), g, blk, 0, s, pl, (const SvEnt*)pl.ent,
                       (const T*)d.Dinv, (const T*)pout,
                       (const T*)d.Ap, xs, (T*)d.r, k, d.max_sys, S->ctl,
                       (const double*)S->hctl.rz2, (const int*)S->hctl.active,
                       (const double*)S->part[1], S->part[2]);
    TC_LAUNCH();
  }
  return TC_OK;
}

int sv_check(tc_serve* S, int k_end, int slot, long long seq, hipStream_t s) {
  const int blocks = S->nsys > kSCheckOneMax ? (S->nsys + kSCheckWaves - 1) / kSCheckWaves : 1;
  if (blocks == 1)
    hipLaunchKernelGGL(k_sv_check<true>, dim3(1), dim3(kSCheck), 0, s, sv_pool(S), k_end,
                       S->nsys, S->d.max_iter, S->d.dt, S->ctl, S->hctl,
                       (const double*)S->part[2], S->dev_view + slot, seq, S->chk, S->chkctl);
  else
    hipLaunchKernelGGL(k_sv_check<false>, dim3(blocks), dim3(kSCheck), 0, s, sv_pool(S), k_end,
                       S->nsys, S->d.max_iter, S->d.dt, S->ctl, S->hctl,
                       (const double*)S->part[2], S->dev_view + slot, seq, S->chk, S->chkctl);
  TC_LAUNCH();
  return TC_OK;
}

int sv_wait(tc_serve* S, int slot, long long seq) {
  const auto t0 = std::chrono::steady_clock::now();
  for (long spins = 0; !status_take(S->host_status + slot, seq, S->status + slot); ++spins) {
    if ((spins & 1023) == 1023) {
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120)) {
        set_error("serve runtime: no status for record %lld after 120 s", seq);
        return TC_ERR_STATE;
      }
      std::this_thread::yield();
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return TC_OK;
}

// The ionic pass, the RHS + PCG start pass and iterations 0 .. K-1 of an
// engine step whose solution is in x[S->xc] (x_prev in xp[S->pc]).
template <class T, class XT, class WT>
int sv_step_seq(tc_serve* S, int K, hipStream_t s) {
  const tc_serve_desc& d = S->d;
  {
    int rc = sv_ionic(S, s);
    if (rc) return rc;
  }
  const SvPool pl = sv_pool(S);
  const dim3 g(S->G), blk(kSB);
  const int xc = S->xc, pc = S->pc;
  const double coef = (1.0 - d.theta) * d.dt, tdt = d.theta * d.dt;
  XT* xs;
  if (d.linear_guess) {
    hipLaunchKernelGGL((k_sv_rhs_init<T, XT, WT, true>), g, blk, 0, s, pl, (const SvEnt*)pl.ent,
                       (const int32_t*)pl.pat, (const int16_t*)pl.dcol, (const int32_t*)pl.col,
                       S->nsys,
                       (const T*)d.Mv, (const T*)d.Kv, (const T*)d.Dinv, (const WT*)d.w,
                       (const XT*)d.x[xc], (const T*)d.xp[pc], (XT*)d.x[1 - xc],
                       (T*)d.xp[1 - pc], (T*)d.r, coef, d.Cm, tdt, S->ctl, S->part[0]);
    xs = (XT*)d.x[1 - xc];
  } else {
    hipLaunchKernelGGL((k_sv_rhs_init<T, XT, WT, false>), g, blk, 0, s, pl, (const SvEnt*)pl.ent,
                       (const int32_t*)pl.pat, (const int16_t*)pl.dcol, (const int32_t*)pl.col,
                       S->nsys,
                       (const T*)d.Mv, (const T*)d.Kv, (const T*)d.Dinv, (const WT*)d.w,
                       (const XT*)d.x[xc], (const T*)nullptr, (XT*)nullptr, (T*)nullptr,
                       (T*)d.r, coef, d.Cm, tdt, S->ctl, S->part[0]);
    xs = (XT*)d.x[xc];
  }
  TC_LAUNCH();
  return sv_iters<T, XT>(S, 0, K, xs, s);
}

template <class T, class XT, class WT>
int sv_enqueue(tc_serve* S, SPlan& plan, hipStream_t s) {
  const tc_serve_desc& d = S->d;
  const int64_t step = S->steps + 1;
  const int slot = (int)(step % kSRing);
  int K = 1;
  for (int v : S->recent) K = v > K ? v : K;
  if (K > d.max_iter) K = d.max_iter;
  const int xc = S->xc, pc = S->pc;
  // the launches fixed by the live systems, the buffer parity and K: a
  // graph replay in graph mode
  int rc;
  if (d.graph)
    rc = graph_replay(S->graphs, (long long)((xc << 1) | pc) | ((long long)K << 2), 0, s,
                      [&](hipStream_t cs) { return sv_step_seq<T, XT, WT>(S, K, cs); });
  else
    rc = sv_step_seq<T, XT, WT>(S, K, s);
  if (rc) return rc;
  plan.seq = ++S->seq;
  plan.slot = slot;
  plan.K = K;
  rc = sv_check(S, K, slot, plan.seq, s);
  if (rc) return rc;
  if (d.linear_guess) {
    S->xc = 1 - xc;
    S->pc = 1 - pc;
  }
  S->steps = step;
  return TC_OK;
}

template <class T, class XT>
int sv_resolve(tc_serve* S, const SPlan& plan, XT* xs, hipStream_t s) {
  StepStatus* st = S->status + plan.slot;
  int it = plan.K;
  int chunk = it > 0 ? it : 1;
  while (st->stalled) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, &S->ctl->stall, 0);
    TC_LAUNCH();
    const int c = chunk < S->d.max_iter - it ? chunk : S->d.max_iter - it;
    int rc = sv_iters<T, XT>(S, it, it + c, xs, s);
    if (rc) return rc;
    it += c;
    const long long seq = ++S->seq;
    rc = sv_check(S, it, plan.slot, seq, s);
    if (rc) return rc;
    rc = sv_wait(S, plan.slot, seq);
    if (rc) return rc;
    chunk = chunk * 2 < 64 ? chunk * 2 : 64;
  }
  return TC_OK;
}

struct SvHostPos {
  int64_t steps;
  int xc, pc;
};

template <class T, class XT, class WT>
int sv_advance(tc_serve* S, int64_t nsteps, int32_t* iters, hipStream_t s) {
  if (S->nsys == 0) {
    for (int64_t k = 0; k < nsteps; ++k) iters[k] = 0;
    return TC_OK;
  }
  const int depth = S->d.depth < 1 ? 1 : (S->d.depth > kSRing / 2 ? kSRing / 2 : S->d.depth);
  struct Inflight {
    SPlan plan;
    SvHostPos before;
  };
  std::deque<Inflight> inflight;
  int64_t enq = 0, done = 0;
  while (done < nsteps) {
    while ((int)inflight.size() < depth && enq < nsteps) {
      Inflight f;
      f.before = {S->steps, S->xc, S->pc};
      int rc = sv_enqueue<T, XT, WT>(S, f.plan, s);
      if (rc) return rc;
      inflight.push_back(f);
      ++enq;
    }
    Inflight f = inflight.front();
    int rc = sv_wait(S, f.plan.slot, f.plan.seq);
    if (rc) return rc;
    StepStatus* st = S->status + f.plan.slot;
    if (st->stalled) {
      ++S->stalls;
      // finish this step, drop the later (gated) ones and resume after it
      XT* xs = (XT*)S->d.x[S->d.linear_guess ? 1 - f.before.xc : f.before.xc];
      rc = sv_resolve<T, XT>(S, f.plan, xs, s);
      if (rc) return rc;
      enq -= (int64_t)inflight.size() - 1;
      inflight.clear();
      S->steps = f.before.steps + 1;
      S->xc = S->d.linear_guess ? 1 - f.before.xc : f.before.xc;
      S->pc = S->d.linear_guess ? 1 - f.before.pc : f.before.pc;
    } else {
      inflight.pop_front();
    }
    iters[done] = st->n_iter;
    S->recent.push_back(st->iter > 0 ? st->iter : 1);
    if ((int)S->recent.size() > S->history) S->recent.pop_front();
    ++done;
  }
  TC_HIP(hipStreamSynchronize(s));
  return TC_OK;
}

}  // namespace
}  // namespace tc

using namespace tc;

extern "C" int tc_serve_workspace_size(int64_t cap_slices, int max_sys, size_t* bytes) {
  TC_CHECK(bytes && cap_slices > 0 && max_sys > 0 && max_sys <= kSCheckSys, TC_ERR_ARG,
           "tc_serve_workspace_size: bad arguments (max_sys 1..%d)", kSCheckSys);
  *bytes = sv_layout(cap_slices, max_sys).total;
  return TC_OK;
}

extern "C" int tc_serve_create(tc_serve** out, const tc_serve_desc* d, void* stream) {
  TC_CHECK(out && d && d->cap_slices > 0 && d->cap_entries > 0 && d->max_sys > 0 &&
               d->max_sys <= kSCheckSys,
           TC_ERR_ARG, "tc_serve_create: bad arguments (max_sys 1..%d)", kSCheckSys);
  TC_CHECK(d->dtype == TC_F32 || d->dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", d->dtype);
  TC_CHECK(!d->w_f64 || d->x_f64 || d->dtype == TC_F64, TC_ERR_ARG, "w_f64 needs x_f64");
  TC_CHECK(d->cap_slices * 64 < (int64_t)INT32_MAX && d->cap_entries < (int64_t)INT32_MAX,
           TC_ERR_SIZE, "pool too large for int32 indices");
  TC_CHECK(d->max_iter >= 1 && d->max_iter <= kMaxIterLimit, TC_ERR_ARG,
           "max_iter %d outside 1..%d", d->max_iter, kMaxIterLimit);
  TC_CHECK(d->k_history >= 0 && d->k_history <= kSHistoryMax, TC_ERR_ARG,
           "k_history %d outside 0..%d", d->k_history, kSHistoryMax);
  TC_CHECK(d->slice_off && d->slice_w && d->slice_sys && d->col && d->Mv && d->Kv && d->Av &&
               d->Dinv && d->x[0] && d->x[1] && d->xp[0] && d->xp[1] && d->w && d->r &&
               d->p[0] && d->p[1] && d->Ap && d->stim_mask,
           TC_ERR_ARG, "tc_serve_create: missing pool array");
  size_t need = 0;
  int rc = tc_serve_workspace_size(d->cap_slices, d->max_sys, &need);
  if (rc) return rc;
  TC_CHECK(d->ws && d->ws_bytes >= need, TC_ERR_SIZE, "serve workspace too small (%zu < %zu)",
           d->ws_bytes, need);
  tc_serve* S = new tc_serve();
  S->d = *d;
  S->history = d->k_history > 0 ? d->k_history : kSHistoryDefault;
  for (int m = 0; m < 3; ++m) {
    if (!d->params[m]) continue;
    if (m == TC_MODEL_MS2V) S->mp[m].ms2v = *(const tc_ms2v_params*)d->params[m];
    if (m == TC_MODEL_TTP) S->mp[m].ttp = *(const tc_ttp_params*)d->params[m];
    if (m == TC_MODEL_CRN) S->mp[m].crn = *(const tc_crn_params*)d->params[m];
    S->d.params[m] = nullptr;
  }
  const SvLayout L = sv_layout(d->cap_slices, d->max_sys);
  char* b = (char*)d->ws;
  S->ctl = (SvCtl*)(b + L.ctl);
  {
    char* a = b + L.arrays;
    const size_t ms = (size_t)d->max_sys;
    SvCtl& h = S->hctl;
    h.stall = 0;
    h.iter = 0;
    h.n_active = 0;
    h.pad = 0;
    h.a_tol = d->a_tol;
    h.r_tol = d->r_tol;
    h.rz2 = (double*)a; a += 2 * ms * 8;
    h.bnorm = (double*)a; a += ms * 8;
    h.alpha = (double*)a; a += ms * 8;
    h.beta = (double*)a; a += ms * 8;
    h.t = (double*)a; a += ms * 8;
    h.total = (long long*)a; a += ms * 8;
    h.steps = (long long*)a; a += ms * 8;
    h.row0 = (int64_t*)a; a += ms * 8;
    h.n = (int64_t*)a; a += ms * 8;
    h.stim = (tc_stimulus*)a; a += ms * TC_POOL_STIM * sizeof(tc_stimulus);
    h.active = (int*)a; a += ms * 4;
    h.n_iter = (int*)a; a += ms * 4;
    h.nstim = (int*)a; a += ms * 4;
    h.use16 = (int*)a; a += ms * 4;
    h.bits = (uint32_t*)a;
  }
  for (int k = 0; k < 3; ++k) S->part[k] = (double*)(b + L.part[k]);
  S->list = (int32_t*)(b + L.list);
  for (int m = 0; m < 3; ++m) S->mlist[m] = (int32_t*)(b + L.mlist[m]);
  S->blk_desc = (SvBlk*)(b + L.blk_desc);
  S->ent = (SvEnt*)(b + L.ent);
  S->sysd = (SvSys*)(b + L.sysd);
  S->chk = (SvChk*)(b + L.chk);
  S->chkctl = (SvChkCtl*)(b + L.chkctl);
  S->nl = 0;
  S->nsys = 0;
  S->G = kXcd;
  for (int m = 0; m < 3; ++m) S->nml[m] = 0;
  S->sys.assign(d->max_sys, SysHost());
  S->slices.reset(d->cap_slices);
  S->entries.reset(d->cap_entries);
  S->xc = S->pc = 0;
  S->seq = 0;
  S->steps = 0;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipHostMalloc((void**)&S->host_status, sizeof(StatusRec) * kSRing,
                               hipHostMallocMapped);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&S->dev_view, S->host_status, 0);
  if (e == hipSuccess) {
    status_clear(S->host_status, kSRing);
    e = hipMemcpyAsync(S->ctl, &S->hctl, sizeof(SvCtl), hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) e = hipMemsetAsync(b + L.arrays, 0, L.part[0] - L.arrays, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error("tc_serve_create: %s", hipGetErrorString(e));
    tc_serve_destroy(S);
    return TC_ERR_HIP;
  }
  *out = S;
  return TC_OK;
}

extern "C" int tc_serve_destroy(tc_serve* S) {
  if (!S) return TC_OK;
  S->graphs.release();
  if (S->host_status) (void)hipHostFree(S->host_status);
  delete S;
  return TC_OK;
}

extern "C" int tc_serve_fits(tc_serve* S, int64_t n_slices, int64_t n_entries, int* fits) {
  TC_CHECK(S && fits, TC_ERR_ARG, "tc_serve_fits: bad arguments");
  bool slot = false;
  for (auto& h : S->sys) slot = slot || !h.live;
  *fits = slot && S->slices.largest() >= n_slices && S->entries.largest() >= n_entries;
  return TC_OK;
}

extern "C" int tc_serve_add(tc_serve* S, const tc_serve_system* y, int* slot_out, void* stream) {
  TC_CHECK(S && y && slot_out && y->n > 0 && y->slice_ptr && y->col && y->Mv && y->Kv && y->Av &&
               y->Dinv && y->V0 && y->states,
           TC_ERR_ARG, "tc_serve_add: bad arguments");
  TC_CHECK(y->model >= 0 && y->model < 3 && S->d.states[y->model], TC_ERR_ARG,
           "tc_serve_add: model %d has no state pool", y->model);
  TC_CHECK(y->n_stim >= 0 && y->n_stim <= TC_POOL_STIM && (y->n_stim == 0 || y->stim_mask),
           TC_ERR_ARG, "tc_serve_add: at most %d stimuli, with a mask", TC_POOL_STIM);
  TC_CHECK(!y->lat || S->d.AT, TC_ERR_ARG, "tc_serve_add: LAT needs AT/RT pool arrays");
  TC_CHECK(y->slices_per_block == 0 ||
               (y->slices_per_block >= kSPerMin && y->slices_per_block <= kSPerMax),
           TC_ERR_ARG, "tc_serve_add: slices_per_block %d outside %d..%d (0: default %d)",
           y->slices_per_block, kSPerMin, kSPerMax, kSPer);
  const tc_serve_desc& d = S->d;
  hipStream_t s = (hipStream_t)stream;
  TC_HIP(hipStreamSynchronize(s));
  int slot = -1;
  for (int q = 0; q < (int)S->sys.size(); ++q)
    if (!S->sys[q].live) {
      slot = q;
      break;
    }
  TC_CHECK(slot >= 0, TC_ERR_SIZE, "tc_serve_add: all %d system slots are in use", d.max_sys);
  const int64_t ns = ceil_div(y->n, 64);
  const int64_t s0 = S->slices.take(ns);
  TC_CHECK(s0 >= 0, TC_ERR_SIZE, "tc_serve_add: no room for %lld slices", (long long)ns);
  const int64_t e0 = S->entries.take(y->n_entries);
  if (e0 < 0) {
    S->slices.give(s0, ns);
    set_error("tc_serve_add: no room for %lld entries", (long long)y->n_entries);
    return TC_ERR_SIZE;
  }
  SysHost& h = S->sys[slot];
  h.live = true;
  h.model = y->model;
  h.slice0 = s0;
  h.ns = ns;
  h.entry0 = e0;
  h.ne = y->n_entries;
  h.n = y->n;
  h.lat = y->lat != 0;
  h.spb = y->slices_per_block > 0 ? y->slices_per_block : kSPer;
  const int64_t row0 = s0 * 64, nrows = ns * 64;
  const int f64 = d.dtype == TC_F64;
  const int xf = d.x_f64 || f64;
  // slices on shared patterns: int16 systems whose patterns the pool keeps
  const bool shared = y->dcol && d.dcol && d.pat && y->pat && y->slice_tag && e0 % 64 == 0;
  hipLaunchKernelGGL(k_sv_place_slices, dim3((int)ceil_div(ns, 256)), dim3(256), 0, s, ns, s0, e0,
                     slot, y->slice_ptr, shared ? y->slice_tag : nullptr, (int32_t*)d.slice_off,
                     (int32_t*)d.slice_w, (int32_t*)d.slice_sys);
  TC_LAUNCH();
  if (shared)
    TC_HIP(hipMemcpyAsync(d.pat + e0 / 64, y->pat, (size_t)(y->n_entries / 64) * sizeof(int32_t),
                          hipMemcpyDeviceToDevice, s));
  const dim3 ge((int)ceil_div(y->n_entries, 256)), gr((int)ceil_div(nrows, 256)), b(256);
  if (f64) {
    hipLaunchKernelGGL(k_sv_place_entries<double>, ge, b, 0, s, y->n_entries, e0, row0, y->col,
                       y->dcol, (const double*)y->Mv, (const double*)y->Kv, (const double*)y->Av,
                       (int32_t*)d.col, d.dcol, (double*)d.Mv, (double*)d.Kv, (double*)d.Av);
    hipLaunchKernelGGL((k_sv_place_rows<double, double>), gr, b, 0, s, row0, y->n, nrows,
                       (const double*)y->V0, (const double*)y->Dinv, y->stim_mask,
                       (double*)d.x[0], (double*)d.x[1], (double*)d.xp[0], (double*)d.xp[1],
                       d.w, 0, (double*)d.Dinv, (double*)d.r, (double*)d.p[0],
                       (double*)d.p[1], (double*)d.Ap, (double*)d.AT, (double*)d.RT,
                       (uint32_t*)d.stim_mask);
    hipLaunchKernelGGL(k_sv_place_states<double>, gr, b, 0, s, row0, y->n, nrows, y->n_states,
                       (const double*)y->states, (double*)d.states[y->model]);
  } else {
    hipLaunchKernelGGL(k_sv_place_entries<float>, ge, b, 0, s, y->n_entries, e0, row0, y->col,
                       y->dcol, (const float*)y->Mv, (const float*)y->Kv, (const float*)y->Av,
                       (int32_t*)d.col, d.dcol, (float*)d.Mv, (float*)d.Kv, (float*)d.Av);
    if (xf)
      hipLaunchKernelGGL((k_sv_place_rows<float, double>), gr, b, 0, s, row0, y->n, nrows,
                         (const float*)y->V0, (const float*)y->Dinv, y->stim_mask,
                         (double*)d.x[0], (double*)d.x[1], (float*)d.xp[0], (float*)d.xp[1],
                         d.w, d.w_f64, (float*)d.Dinv, (float*)d.r, (float*)d.p[0],
                         (float*)d.p[1], (float*)d.Ap, (float*)d.AT, (float*)d.RT,
                         (uint32_t*)d.stim_mask);
    else
      hipLaunchKernelGGL((k_sv_place_rows<float, float>), gr, b, 0, s, row0, y->n, nrows,
                         (const float*)y->V0, (const float*)y->Dinv, y->stim_mask,
                         (float*)d.x[0], (float*)d.x[1], (float*)d.xp[0], (float*)d.xp[1], d.w,
                         0, (float*)d.Dinv, (float*)d.r, (float*)d.p[0], (float*)d.p[1],
                         (float*)d.Ap, (float*)d.AT, (float*)d.RT, (uint32_t*)d.stim_mask);
    hipLaunchKernelGGL(k_sv_place_states<float>, gr, b, 0, s, row0, y->n, nrows, y->n_states,
                       (const float*)y->states, (float*)d.states[y->model]);
  }
  TC_LAUNCH();
  // per-system scalars
  const SvCtl& hc = S->hctl;
  const double zero = 0.0;
  const long long zl = 0;
  const int zi = 0;
  const uint32_t bits0 = stimulus_bits(y->stim, y->n_stim, d.dt);
  tc_stimulus st[TC_POOL_STIM] = {};
  for (int k = 0; k < y->n_stim; ++k) st[k] = y->stim[k];
  TC_HIP(hipMemcpyAsync(hc.t + slot, &zero, 8, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.steps + slot, &zl, 8, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.total + slot, &zl, 8, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.row0 + slot, &row0, 8, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.n + slot, &y->n, 8, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.active + slot, &zi, 4, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.n_iter + slot, &zi, 4, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.nstim + slot, &y->n_stim, 4, hipMemcpyHostToDevice, s));
  const int use16 = (y->dcol && d.dcol) ? 1 : 0;
  h.use16 = use16 != 0;
  TC_HIP(hipMemcpyAsync(hc.use16 + slot, &use16, 4, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.bits + slot, &bits0, 4, hipMemcpyHostToDevice, s));
  TC_HIP(hipMemcpyAsync(hc.stim + (int64_t)slot * TC_POOL_STIM, st, sizeof(st),
                        hipMemcpyHostToDevice, s));
  // the new system starts in the current x / x_prev buffers
  int rc = sv_rebuild(S, s);
  if (rc) return rc;
  *slot_out = slot;
  return TC_OK;
}

extern "C" int tc_serve_remove(tc_serve* S, int slot, void* V_out, void* AT_out, void* RT_out,
                               int64_t* total_iters, int64_t* steps, void* stream) {
  TC_CHECK(S && slot >= 0 && slot < (int)S->sys.size() && S->sys[slot].live && V_out,
           TC_ERR_ARG, "tc_serve_remove: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const tc_serve_desc& d = S->d;
  SysHost& h = S->sys[slot];
  TC_HIP(hipStreamSynchronize(s));
  long long st_steps = 0, tot = 0;
  double t = 0.0;
  TC_HIP(hipMemcpy(&st_steps, S->hctl.steps + slot, 8, hipMemcpyDeviceToHost));
  TC_HIP(hipMemcpy(&tot, S->hctl.total + slot, 8, hipMemcpyDeviceToHost));
  TC_HIP(hipMemcpy(&t, S->hctl.t + slot, 8, hipMemcpyDeviceToHost));
  const int64_t row0 = h.slice0 * 64;
  const dim3 g((int)ceil_div(h.n, 256)), b(256);
  const void* x = d.x[S->xc];
  const bool lat = h.lat && AT_out && RT_out;
  if (d.dtype == TC_F64) {
    if (lat && st_steps > 0)
      hipLaunchKernelGGL((k_sv_lat_final<double, double>), g, b, 0, s, row0, h.n,
                         (const double*)x, (double*)d.AT, (double*)d.RT, t);
    hipLaunchKernelGGL((k_sv_extract<double, double>), g, b, 0, s, row0, h.n, (const double*)x,
                       (const double*)d.AT, (const double*)d.RT, (double*)V_out,
                       lat ? (double*)AT_out : nullptr, lat ? (double*)RT_out : nullptr);
  } else if (d.x_f64) {
    if (lat && st_steps > 0)
      hipLaunchKernelGGL((k_sv_lat_final<float, double>), g, b, 0, s, row0, h.n,
                         (const double*)x, (float*)d.AT, (float*)d.RT, t);
    hipLaunchKernelGGL((k_sv_extract<float, double>), g, b, 0, s, row0, h.n, (const double*)x,
                       (const float*)d.AT, (const float*)d.RT, (float*)V_out,
                       lat ? (float*)AT_out : nullptr, lat ? (float*)RT_out : nullptr);
  } else {
    if (lat && st_steps > 0)
      hipLaunchKernelGGL((k_sv_lat_final<float, float>), g, b, 0, s, row0, h.n, (const float*)x,
                         (float*)d.AT, (float*)d.RT, t);
    hipLaunchKernelGGL((k_sv_extract<float, float>), g, b, 0, s, row0, h.n, (const float*)x,
                       (const float*)d.AT, (const float*)d.RT, (float*)V_out,
                       lat ? (float*)AT_out : nullptr, lat ? (float*)RT_out : nullptr);
  }
  TC_LAUNCH();
  if (total_iters) *total_iters = tot;
  if (steps) *steps = st_steps;
  S->slices.give(h.slice0, h.ns);
  S->entries.give(h.entry0, h.ne);
  h = SysHost();
  return sv_rebuild(S, s);
}

extern "C" int tc_serve_stats(const tc_serve* S, int64_t* steps, int64_t* stalls) {
  TC_CHECK(S && steps && stalls, TC_ERR_ARG, "tc_serve_stats: bad arguments");
  *steps = S->steps;
  *stalls = S->stalls;
  return TC_OK;
}

extern "C" int tc_serve_graph_stats(const tc_serve* S, int64_t* captures, int64_t* replays) {
  TC_CHECK(S && captures && replays, TC_ERR_ARG, "tc_serve_graph_stats: bad arguments");
  *captures = S->graphs.captures;
  *replays = S->graphs.replays;
  return TC_OK;
}

extern "C" int tc_serve_advance(tc_serve* S, int64_t nsteps, int32_t* iters, void* stream) {
  TC_CHECK(S && iters && nsteps >= 0, TC_ERR_ARG, "tc_serve_advance: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (S->d.dtype == TC_F64) return sv_advance<double, double, double>(S, nsteps, iters, s);
  if (S->d.x_f64 && S->d.w_f64) return sv_advance<float, double, double>(S, nsteps, iters, s);
  if (S->d.x_f64) return sv_advance<float, double, float>(S, nsteps, iters, s);
  return sv_advance<float, float, float>(S, nsteps, iters, s);
}

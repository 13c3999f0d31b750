// Batched monodomain runs on one shared mesh: multi-right-hand-side
// Jacobi-PCG over instance-minor vectors, and the pipelined time loop that
// advances all instances together.
//
// Layout: entry (node i, instance b) of every per-instance vector lives at
// i * lanes + b.  A wavefront owns 64/lanes consecutive rows; lane l works on
// row l / lanes of them for instance l % lanes.  The CSR pattern and the
// shared values (M and the NK stiffness components) are loaded once per
// wavefront for all of its instances, and each neighbour gather is a
// coalesced row of `lanes` consecutive entries.
//
// Reductions: every pass block writes per-instance partial sums
// partials[b * blocks + block]; a fold kernel with one block per instance adds
// them in block order (deterministic) and updates that instance's scalars.
// Instances that met the stopping rule become inactive: their lanes skip the
// updates and the folds leave their scalars alone, so each instance follows
// exactly the iteration of a run of its own.
//
// The time loop mirrors the single-run runtime: per step the host enqueues
// the ionic kernel of every model group, the fused RHS + PCG start pass, K
// speculative iterations and a one-block check that publishes the step
// status in pinned host memory; a step that needs more iterations raises
// the stall flag, every later kernel then returns at once and the host
// finishes that step and re-issues the later ones.
#include <atomic>
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "pcg_kernels.h"

namespace tc {
namespace {

constexpr int kLanes = TC_BATCH_MAX;  // 64
constexpr int kBB = 512;              // threads per block of the passes
constexpr int kBW = kBB / 64;         // waves per block
constexpr int kBMaxBlocks = 1024;
constexpr int kBRing = 64;
constexpr int kBHistory = 16;

struct BatCtl {
  int stall;  // gate of every later launch of a pipelined run
  int iter;   // iterations executed in the current step
  int pad[2];
  double a_tol, r_tol;
  double rz[kLanes], bnorm[kLanes], alpha[kLanes], beta[kLanes];
  int active[kLanes];  // instance still iterating in this step
  int n_iter[kLanes];  // reference n_iter of this step
  long long total[kLanes];
};

template <class T>
struct BatOp {
  int64_t n;
  int lanes, lg, B;  // lanes = 1 << lg
  const int32_t* rp;
  const int32_t* col;
  const T* M;
  const T* K[TC_BATCH_MAX_K];
  const double* coef;  // (NK, lanes)
  double Cm, tdt, ck;  // ck = (1 - theta) dt
};

struct RowRange {
  int64_t g0, g1;
};

// Row entries per chunk: the column and shared-value loads of a chunk are
// issued before its gathers (several loads in flight per wavefront).
constexpr int kBChunk = 4;
// Entries per gather group of the 64-lane passes (all gathers of a group in
// flight together).
constexpr int kBSeg = 8;


// Row groups (64/lanes rows each, one per wavefront pass) of this block:
// one contiguous range per block, blocks of one XCD adjacent.
__device__ inline RowRange block_rows(int64_t ngroups) {
  const int G = gridDim.x;
  const int b = blockIdx.x;
  int vb = b;
  if (G % kXcd == 0) vb = (b % kXcd) * (G / kXcd) + b / kXcd;
  const int64_t per = (ngroups + G - 1) / G;
  RowRange r;
  r.g0 = (int64_t)vb * per;
  r.g1 = r.g0 + per < ngroups ? r.g0 + per : ngroups;
  return r;
}

// Block gate: the pipeline is stalled, or (need_active) no instance is
// still iterating.
__device__ inline bool bat_halted(const BatCtl* ctl, int B, bool need_active) {
  __shared__ int h;
  if (threadIdx.x < 64) {
    const int a = (int)threadIdx.x < B ? ctl->active[threadIdx.x] : 0;
    const unsigned long long any = __ballot(a);
    if (threadIdx.x == 0) h = (ctl->stall != 0) || (need_active && any == 0ull);
  }
  __syncthreads();
  return h != 0;
}

// Per-instance block sums of NV values -> part[k * G * lanes + b * G + block]:
// instance-major, so the fold of one instance reads G consecutive values.
template <int NV>
__device__ inline void bat_partials(double (&v)[NV], double* part, int lanes) {
  __shared__ double sh[NV][kBW][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int off = lanes; off < 64; off <<= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (lane < lanes) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][wv][lane] = v[k];
  }
  __syncthreads();
  if ((int)threadIdx.x < lanes) {
    const int64_t GL = (int64_t)gridDim.x * lanes;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < kBW; ++w) s += sh[k][w][threadIdx.x];
      part[k * GL + (int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
  }
}

// Value of lane `src` (wave-uniform) in every lane.
__device__ inline int32_t bcast(int32_t v, int src) { return __builtin_amdgcn_readlane(v, src); }
__device__ inline float bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ inline double bcast(double v, int src) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One row segment of up to 64 entries held across the wave (lane t: entry
// j + t): column, M and the NK component values.
template <class T, int NK>
struct RowSeg {
  int32_t col;
  T m;
  T k[NK];
  __device__ inline void load(const BatOp<T>& op, int32_t j, int len, int lane) {
    const int32_t jj = j + (lane < len ? lane : len - 1);
    col = op.col[jj];
    m = op.M[jj];
#pragma unroll
    for (int q = 0; q < NK; ++q) k[q] = op.K[q][jj];
  }
};

// a_j of instance b: Cm M_j + sum_k tc[k] K_k,j (tc = tdt * coef).
template <class T, int NK>
__device__ inline double a_val(const BatOp<T>& op, const double (&tcf)[NK], int32_t j) {
  double a = op.Cm * (double)op.M[j];
#pragma unroll
  for (int k = 0; k < NK; ++k) a += tcf[k] * (double)op.K[k][j];
  return a;
}

template <class T, int NK>
__device__ inline double k_val(const BatOp<T>& op, const double (&cf)[NK], int32_t j) {
  double a = 0.0;
#pragma unroll
  for (int k = 0; k < NK; ++k) a += cf[k] * (double)op.K[k][j];
  return a;
}

// Dinv = 1 / diag(A_b) for every (node, lane).
template <class T, int NK>
__global__ void k_bat_dinv(BatOp<T> op, T* __restrict__ Dinv) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= op.n * op.lanes) return;
  const int64_t i = e >> op.lg;
  const int b = (int)(e & (op.lanes - 1));
  double tcf[NK];
#pragma unroll
  for (int k = 0; k < NK; ++k) tcf[k] = op.tdt * op.coef[k * op.lanes + b];
  double d = 0.0;
  for (int32_t j = op.rp[i]; j < op.rp[i + 1]; ++j)
    if (op.col[j] == (int32_t)i) d = a_val<T, NK>(op, tcf, j);
  Dinv[e] = (T)(1.0 / d);
}

// Fused RHS + PCG start (the batched form of k_rhs_init):
//   b = M w - ck K_b u,  x0 = 2x - x_prev (LG) or x,  r = b - A_b x0,
// partial sums of r.z and (Dinv b)^2 per instance.
template <class T, class XT, class WT, bool LG, int NK, bool WU>
__global__ void __launch_bounds__(kBB)
k_bat_rhs_init(BatOp<T> op, const WT* __restrict__ w, const XT* __restrict__ x,
               const T* __restrict__ xp, XT* __restrict__ xnext, T* __restrict__ xpnext,
               T* __restrict__ r, const T* __restrict__ Dinv, BatCtl* ctl, double* part) {
  if (bat_halted(ctl, op.B, false)) return;
  const int lanes = op.lanes, rpw = 64 >> op.lg;
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int b = lane & (lanes - 1), rsub = lane >> op.lg;
  double cf[NK];
#pragma unroll
  for (int k = 0; k < NK; ++k) cf[k] = op.coef[k * lanes + b];
  const int64_t ng = (op.n + rpw - 1) / rpw;
  const RowRange rr = block_rows(ng);
  double red[2] = {0.0, 0.0};
  for (int64_t g = rr.g0 + wv; g < rr.g1; g += kBW) {
    const int64_t i = WU ? g : g * rpw + rsub;
    if (!WU && i >= op.n) continue;
    const int32_t j0 = op.rp[i], j1 = op.rp[i + 1];
    double am = 0.0, ak = 0.0, mx = 0.0, kx = 0.0;
    if (WU) {
      for (int32_t js = j0; js < j1; js += 64) {
        const int len = j1 - js < 64 ? j1 - js : 64;
        RowSeg<T, NK> rs;
        rs.load(op, js, len, lane);
        for (int u0 = 0; u0 < len; u0 += kBSeg) {
          XT xj[kBSeg];
          WT wj[kBSeg];
          T pj[kBSeg];
#pragma unroll
          for (int u = 0; u < kBSeg; ++u) {
            const int uu = u0 + u < len ? u0 + u : len - 1;
            const int64_t e = ((int64_t)bcast(rs.col, uu) << 6) + lane;
            xj[u] = x[e];
            wj[u] = w[e];
            if (LG) pj[u] = xp[e];
          }
#pragma unroll
          for (int u = 0; u < kBSeg; ++u) {
            if (u0 + u >= len) break;
            const double mj = (double)bcast(rs.m, u0 + u);
            double kj = 0.0;
#pragma unroll
            for (int q = 0; q < NK; ++q) kj += cf[q] * (double)bcast(rs.k[q], u0 + u);
            double x0;
            if (LG) x0 = (double)(XT)(2.0 * (double)xj[u] - (double)pj[u]);
            else x0 = (double)xj[u];
            am += mj * (double)wj[u];
            if (sizeof(WT) == 8) ak += kj * (double)xj[u];
            else ak += kj * (double)(T)xj[u];
            mx += mj * x0;
            kx += kj * x0;
          }
        }
      }
    } else
    for (int32_t jc = j0; jc < j1; jc += kBChunk) {
      int32_t cj[kBChunk];
      double mv[kBChunk], kv[kBChunk];
      XT xj[kBChunk];
      WT wj[kBChunk];
      T pj[kBChunk];
#pragma unroll
      for (int u = 0; u < kBChunk; ++u) {
        const int32_t jj = jc + u < j1 ? jc + u : j1 - 1;
        cj[u] = op.col[jj];
        mv[u] = (double)op.M[jj];
        kv[u] = k_val<T, NK>(op, cf, jj);
      }
#pragma unroll
      for (int u = 0; u < kBChunk; ++u) {
        const int64_t e = (int64_t)cj[u] * lanes + b;
        xj[u] = x[e];
        wj[u] = w[e];
        if (LG) pj[u] = xp[e];
      }
#pragma unroll
      for (int u = 0; u < kBChunk; ++u) {
        if (jc + u >= j1) break;
        double x0;
        if (LG) x0 = (double)(XT)(2.0 * (double)xj[u] - (double)pj[u]);
        else x0 = (double)xj[u];
        am += mv[u] * (double)wj[u];
        if (sizeof(WT) == 8) ak += kv[u] * (double)xj[u];
        else ak += kv[u] * (double)(T)xj[u];
        mx += mv[u] * x0;
        kx += kv[u] * x0;
      }
    }
    const int64_t e = i * lanes + b;
    if (LG) {
      xnext[e] = (XT)(2.0 * (double)x[e] - (double)xp[e]);
      xpnext[e] = (T)x[e];
    }
    double bi = am - op.ck * ak;
    if (sizeof(WT) != 8) bi = (double)(T)bi;
    const double d = (double)Dinv[e];
    const T ri = (T)(bi - ax0(op.Cm, mx, op.tdt, kx));
    r[e] = ri;
    red[0] += (double)ri * (d * (double)ri);
    const double db = d * bi;
    red[1] += db * db;
  }
  bat_partials<2>(red, part, lanes);
}

// p' = beta p + Dinv r (k == 0: Dinv r) for the active instances,
// elementwise, so that the spmv pass gathers one vector.
template <class T>
__global__ void __launch_bounds__(256)
k_bat_pform(int64_t n, int lanes, int B, const T* __restrict__ Dinv, const T* __restrict__ r,
            const T* __restrict__ pin, T* __restrict__ pout, int k, BatCtl* ctl) {
  if (bat_halted(ctl, B, true)) return;
  const int64_t m = n * lanes;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e & (lanes - 1));
    if (b >= B || !ctl->active[b]) continue;
    if (k == 0) pout[e] = (T)((double)Dinv[e] * (double)r[e]);
    else pout[e] = (T)pdir(ctl->beta[b], (double)pin[e], (double)Dinv[e], (double)r[e]);
  }
}

// spmv pass of iteration k: Ap' = A_b p' (p' from k_bat_pform), partial
// sums of p'.Ap' (active instances only).
template <class T, int NK, bool WU>
__global__ void __launch_bounds__(kBB)
k_bat_spmv(BatOp<T> op, const int32_t* __restrict__ rp, const T* __restrict__ p,
           T* __restrict__ Ap, int k, BatCtl* ctl, double* part) {
  if (bat_halted(ctl, op.B, true)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->iter = k + 1;
  const int lanes = op.lanes, rpw = 64 >> op.lg;
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int b = lane & (lanes - 1), rsub = lane >> op.lg;
  const bool act = b < op.B && ctl->active[b] != 0;
  double tcf[NK];
#pragma unroll
  for (int q = 0; q < NK; ++q) tcf[q] = op.tdt * op.coef[q * lanes + b];
  const int64_t ng = (op.n + rpw - 1) / rpw;
  const RowRange rr = block_rows(ng);
  double red[1] = {0.0};
  if (WU) {
    // 64 lanes = 64 instances of one row: the row's pattern and shared
    // values are read once, spread over the lanes, and broadcast.
    for (int64_t g = rr.g0 + wv; g < rr.g1; g += kBW) {
      const int64_t i = g;
      const int32_t j0 = rp[i], j1 = rp[i + 1];
      double acc = 0.0;
      for (int32_t js = j0; js < j1; js += 64) {
        const int len = j1 - js < 64 ? j1 - js : 64;
        RowSeg<T, NK> rs;
        rs.load(op, js, len, lane);
        for (int u0 = 0; u0 < len; u0 += kBSeg) {
          T pj[kBSeg];
#pragma unroll
          for (int u = 0; u < kBSeg; ++u) {
            const int uu = u0 + u < len ? u0 + u : len - 1;
            pj[u] = p[((int64_t)bcast(rs.col, uu) << 6) + lane];
          }
#pragma unroll
          for (int u = 0; u < kBSeg; ++u) {
            const int uu = u0 + u < len ? u0 + u : len - 1;
            double a = op.Cm * (double)bcast(rs.m, uu);
#pragma unroll
            for (int q = 0; q < NK; ++q) a += tcf[q] * (double)bcast(rs.k[q], uu);
            const double t = acc + a * (double)pj[u];
            acc = u0 + u < len ? t : acc;
          }
        }
      }
      if (act) {
        const int64_t e = (i << 6) + lane;
        const T pi = p[e];
        const T api = (T)acc;
        Ap[e] = api;
        red[0] += (double)pi * (double)api;
      }
    }
  } else if (act) {
    for (int64_t g = rr.g0 + wv; g < rr.g1; g += kBW) {
      const int64_t i = g * rpw + rsub;
      if (i >= op.n) continue;
      const int32_t j0 = rp[i], j1 = rp[i + 1];
      double acc = 0.0;
      for (int32_t jc = j0; jc < j1; jc += kBChunk) {
        int32_t cj[kBChunk];
        double av[kBChunk];
        T pj[kBChunk];
#pragma unroll
        for (int u = 0; u < kBChunk; ++u) {
          const int32_t jj = jc + u < j1 ? jc + u : j1 - 1;
          cj[u] = op.col[jj];
          av[u] = a_val<T, NK>(op, tcf, jj);
        }
#pragma unroll
        for (int u = 0; u < kBChunk; ++u) pj[u] = p[(int64_t)cj[u] * lanes + b];
#pragma unroll
        for (int u = 0; u < kBChunk; ++u) {
          const double t = acc + av[u] * (double)pj[u];
          acc = jc + u < j1 ? t : acc;
        }
      }
      const int64_t e = i * lanes + b;
      const T pi = p[e];
      const T api = (T)acc;
      Ap[e] = api;
      red[0] += (double)pi * (double)api;
    }
  }
  bat_partials<1>(red, part, lanes);
}

// update pass: x += alpha p', r -= alpha Ap', partial sums of z.z and r.z.
template <class T, class XT>
__global__ void __launch_bounds__(kBB)
k_bat_update(int64_t n, int lanes, int lg, int B, const T* __restrict__ Dinv,
             const T* __restrict__ p, const T* __restrict__ Ap, XT* __restrict__ x,
             T* __restrict__ r, BatCtl* ctl, double* part) {
  if (bat_halted(ctl, B, true)) return;
  const int rpw = 64 >> lg;
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int b = lane & (lanes - 1), rsub = lane >> lg;
  const bool act = b < B && ctl->active[b] != 0;
  const double alpha = act ? ctl->alpha[b] : 0.0;
  const int64_t ng = (n + rpw - 1) / rpw;
  const RowRange rr = block_rows(ng);
  double red[2] = {0.0, 0.0};
  if (act) {
    for (int64_t g = rr.g0 + wv; g < rr.g1; g += kBW) {
      const int64_t i = g * rpw + rsub;
      if (i >= n) continue;
      const int64_t e = i * lanes + b;
      const XT xi = (XT)((double)x[e] + alpha * (double)p[e]);
      const T ri = (T)((double)r[e] - alpha * (double)Ap[e]);
      x[e] = xi;
      r[e] = ri;
      const double z = (double)Dinv[e] * (double)ri;
      red[0] += z * z;
      red[1] += (double)ri * z;
    }
  }
  bat_partials<2>(red, part, lanes);
}

// Per-instance folds, one block per instance:
//   kind 0: start of a step: rz = sum r.z, bnorm = ||Dinv b||, instance active
//   kind 1: alpha = rz / sum p'.Ap'
//   kind 2: stopping test of iteration k (solver.py:57-74), else beta, rz.
__global__ void __launch_bounds__(kBB)
k_bat_fold(int kind, int k, int G, int lanes, int B, BatCtl* ctl, const double* part) {
  const int b = blockIdx.x;
  if (b >= B) return;
  __shared__ int go;
  if (threadIdx.x == 0) go = ctl->stall == 0 && (kind == 0 || ctl->active[b] != 0);
  // the partials load (clamped index, then a select) while the flags arrive
  const int64_t GL = (int64_t)G * lanes;
  constexpr int R = (kBMaxBlocks + kBB - 1) / kBB;
  double x0[R], x1[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int g = (int)threadIdx.x + q * kBB;
    const int gc = g < G ? g : G - 1;
    const double t0 = part[(int64_t)b * G + gc];
    const double t1 = kind != 1 ? part[GL + (int64_t)b * G + gc] : 0.0;
    x0[q] = g < G ? t0 : 0.0;
    x1[q] = g < G ? t1 : 0.0;
  }
  __syncthreads();
  if (!go) return;
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int q = 0; q < R; ++q) {
    v[0] += x0[q];
    v[1] += x1[q];
  }
  block_sum<2>(v);
  if (threadIdx.x != 0) return;
  if (kind == 0) {
    ctl->rz[b] = v[0];
    ctl->bnorm[b] = sqrt(v[1]);
    ctl->active[b] = 1;
    ctl->n_iter[b] = 0;
    ctl->beta[b] = 0.0;
    if (b == 0) ctl->iter = 0;
  } else if (kind == 1) {
    ctl->alpha[b] = ctl->rz[b] / v[0];
  } else {
    const double zn = sqrt(v[0]);
    if (zn < ctl->a_tol || zn / ctl->bnorm[b] < ctl->r_tol) {
      ctl->active[b] = 0;
      ctl->n_iter[b] = k;
    } else {
      ctl->beta[b] = v[1] / ctl->rz[b];
      ctl->rz[b] = v[1];
      ctl->n_iter[b] = k + 1;
    }
  }
}

// One wavefront: end of a launch chunk after k_end iterations.  Instances
// still iterating below max_iter stall the pipeline; otherwise the step is
// final: per-instance n_iter sums advance and the status (largest n_iter)
// goes to pinned host memory as a StatusRec.
__global__ void __launch_bounds__(64)
k_bat_check(int k_end, int B, int max_iter, BatCtl* ctl, StatusRec* status, long long seq) {
  if (ctl->stall) return;
  const int b = threadIdx.x;
  const int act = b < B ? ctl->active[b] : 0;
  const int stalled = act && k_end < max_iter;
  const unsigned long long any_stall = __ballot(stalled);
  if (act && !stalled) ctl->active[b] = 0;  // max_iter reached: n_iter == max_iter
  int ni = b < B ? ctl->n_iter[b] : 0;
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(ni, off, 64);
    ni = o > ni ? o : ni;
  }
  if (!any_stall && b < B) ctl->total[b] += ctl->n_iter[b];
  if (b == 0) {
    if (any_stall) ctl->stall = 1;
    if (status)
      put_status(status, (uint32_t)ni, (uint32_t)ctl->iter | (any_stall ? kStStalled : 0u), seq);
  }
}

__global__ void k_bat_init_ctl(BatCtl* ctl, double a_tol, double r_tol) {
  const int b = threadIdx.x;
  if (b == 0) {
    ctl->stall = 0;
    ctl->iter = 0;
    ctl->a_tol = a_tol;
    ctl->r_tol = r_tol;
  }
  if (b < kLanes) {
    ctl->active[b] = 0;
    ctl->n_iter[b] = 0;
    ctl->total[b] = 0;
    ctl->rz[b] = ctl->bnorm[b] = ctl->alpha[b] = ctl->beta[b] = 0.0;
  }
}

// Blocks of the passes: one wavefront per row group, at most `cap` blocks
// (the workspace is sized for kBMaxBlocks).
inline int bat_blocks(int64_t n, int lanes, int cap = kBMaxBlocks) {
  const int64_t rpw = 64 / lanes;
  const int64_t ng = ceil_div(n, rpw);
  int64_t g = ceil_div(ng, kBW);
  if (g > cap) g = cap;
  if (g >= kXcd) g -= g % kXcd;
  return (int)(g < 1 ? 1 : g);
}

int log2_exact(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return (1 << l) == v ? l : -1;
}

struct BHost {
  int64_t steps = 0;
  double t = 0.0;
  int xc = 0, pc = 0;
  bool lat_pending = false;
  double lat_t = 0.0;
};

struct BPlan {
  int64_t step;
  int slot, K;
  long long seq;
  BHost after;
};

}  // namespace
}  // namespace tc

struct tc_batch {
  tc_batch_desc d;
  union {
    tc_ms2v_params ms2v;
    tc_ttp_params ttp;
    tc_crn_params crn;
  } gp[TC_BATCH_MAX_GROUPS];
  int G, Gp, lg;  // blocks of the row passes and of the elementwise p' pass
  tc::BatCtl* ctl;
  double* part[3];  // start (2), spmv (1), update (2); each block of G*lanes
  void* r;
  void* p[2];
  void* Ap;
  tc::StatusRec* host_status;  // pinned, written by the check pass
  tc::StatusRec* dev_view;     // device address of host_status
  tc::StepStatus status[tc::kBRing];  // records as read by bat_wait
  long long seq;
  tc::BHost vs;
  std::deque<int> recent;
  tc::GraphCache graphs;  // graph mode: the step's RHS + PCG launches
};

namespace tc {
namespace {

struct WsLayout {
  size_t ctl, part[3], r, p0, p1, Ap, total;
};

WsLayout ws_layout(int64_t n, int lanes, int dtype) {
  const int G = bat_blocks(n, lanes);
  const size_t GL = (size_t)G * lanes * sizeof(double);
  const size_t vb = align_up((size_t)n * lanes * (dtype == TC_F64 ? 8 : 4), 256);
  WsLayout L;
  size_t off = 0;
  L.ctl = off; off += align_up(sizeof(BatCtl), 256);
  L.part[0] = off; off += align_up(2 * GL, 256);
  L.part[1] = off; off += align_up(GL, 256);
  L.part[2] = off; off += align_up(2 * GL, 256);
  L.r = off; off += vb;
  L.p0 = off; off += vb;
  L.p1 = off; off += vb;
  L.Ap = off; off += vb;
  L.total = off;
  return L;
}

template <class T>
BatOp<T> make_op(const tc_batch* S) {
  const tc_batch_desc& d = S->d;
  BatOp<T> op;
  op.n = d.n;
  op.lanes = d.lanes;
  op.lg = S->lg;
  op.B = d.B;
  op.rp = d.row_ptr;
  op.col = d.col;
  op.M = (const T*)d.Mv;
  for (int k = 0; k < TC_BATCH_MAX_K; ++k) op.K[k] = k < d.NK ? (const T*)d.Kv[k] : nullptr;
  op.coef = d.coef;
  op.Cm = d.Cm;
  op.tdt = d.theta * d.dt;
  op.ck = (1.0 - d.theta) * d.dt;
  return op;
}

int group_ionic(tc_batch* S, int g, const tc_ionic_io* io, hipStream_t s) {
  const tc_batch_group& G = S->d.groups[g];
  switch (G.model) {
    case TC_MODEL_MS2V:
      return ms2v_step(&S->gp[g].ms2v, G.states, io, S->d.dtype, s, nullptr, nullptr);
    case TC_MODEL_TTP:
      return ttp_step(&S->gp[g].ttp, G.states, G.tab[0], G.tab[1], G.tab[2], io, S->d.dtype, s,
                      nullptr, nullptr);
    case TC_MODEL_CRN:
      return crn_step(&S->gp[g].crn, G.states, G.tab[0], G.tab[1], G.tab[2], io, S->d.dtype, s,
                      nullptr, nullptr);
  }
  set_error("batch group %d: unknown model %d", g, G.model);
  return TC_ERR_ARG;
}

template <class T, class XT, int NK>
int bat_iters(tc_batch* S, int k0, int k1, XT* xs, hipStream_t s) {
  const tc_batch_desc& d = S->d;
  const BatOp<T> op = make_op<T>(S);
  const dim3 g(S->G), blk(kBB), fb(kBB);
  T* pp[2] = {(T*)S->p[0], (T*)S->p[1]};
  for (int k = k0; k < k1; ++k) {
    T* pin = pp[k & 1];
    T* pout = pp[(k + 1) & 1];
    hipLaunchKernelGGL(k_bat_pform<T>, dim3(S->Gp), dim3(256), 0, s, d.n, d.lanes, d.B,
                       (const T*)d.Dinv, (const T*)S->r, (const T*)pin, pout, k, S->ctl);
    TC_LAUNCH();
    if (d.lanes == 64)
      hipLaunchKernelGGL((k_bat_spmv<T, NK, true>), g, blk, 0, s, op, (const int32_t*)op.rp,
                         (const T*)pout,
                         (T*)S->Ap, k, S->ctl, S->part[1]);
    else
      hipLaunchKernelGGL((k_bat_spmv<T, NK, false>), g, blk, 0, s, op, (const int32_t*)op.rp,
                         (const T*)pout,
                         (T*)S->Ap, k, S->ctl, S->part[1]);
    TC_LAUNCH();
    hipLaunchKernelGGL(k_bat_fold, dim3(d.B), fb, 0, s, 1, k, S->G, d.lanes, d.B, S->ctl,
                       (const double*)S->part[1]);
    TC_LAUNCH();
    hipLaunchKernelGGL((k_bat_update<T, XT>), g, blk, 0, s, d.n, d.lanes, S->lg, d.B,
                       (const T*)d.Dinv, (const T*)pout, (const T*)S->Ap, xs, (T*)S->r, S->ctl,
                       S->part[2]);
    TC_LAUNCH();
    hipLaunchKernelGGL(k_bat_fold, dim3(d.B), fb, 0, s, 2, k, S->G, d.lanes, d.B, S->ctl,
                       (const double*)S->part[2]);
    TC_LAUNCH();
  }
  return TC_OK;
}

int bat_check(tc_batch* S, int k_end, int slot, long long seq, hipStream_t s) {
  hipLaunchKernelGGL(k_bat_check, dim3(1), dim3(64), 0, s, k_end, S->d.B, S->d.max_iter, S->ctl,
                     S->dev_view + slot, seq);
  TC_LAUNCH();
  return TC_OK;
}

int bat_wait(tc_batch* S, int slot, long long seq) {
  const auto t0 = std::chrono::steady_clock::now();
  for (long spins = 0; !status_take(S->host_status + slot, seq, S->status + slot); ++spins) {
    if ((spins & 1023) == 1023) {
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120)) {
        set_error("batch runtime: no status for record %lld after 120 s", seq);
        return TC_ERR_STATE;
      }
      std::this_thread::yield();
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return TC_OK;
}

template <class T, class XT, class WT, bool LG, int NK>
void bat_rhs_launch(tc_batch* S, const BatOp<T>& op, int xc, int pc, hipStream_t s) {
  const tc_batch_desc& d = S->d;
  const WT* w = (const WT*)d.w;
  const XT* x = (const XT*)d.x[xc];
  const T* xp = LG ? (const T*)d.xp[pc] : nullptr;
  XT* xn = LG ? (XT*)d.x[1 - xc] : nullptr;
  T* xpn = LG ? (T*)d.xp[1 - pc] : nullptr;
  const dim3 g(S->G), blk(kBB);
  if (d.lanes == 64)
    hipLaunchKernelGGL((k_bat_rhs_init<T, XT, WT, LG, NK, true>), g, blk, 0, s, op, w, x, xp, xn,
                       xpn, (T*)S->r, (const T*)d.Dinv, S->ctl, S->part[0]);
  else
    hipLaunchKernelGGL((k_bat_rhs_init<T, XT, WT, LG, NK, false>), g, blk, 0, s, op, w, x, xp,
                       xn, xpn, (T*)S->r, (const T*)d.Dinv, S->ctl, S->part[0]);
}

// The RHS + PCG start pass, its fold and iterations 0 .. K-1 of a step whose
// solution is in x[xc] (x_prev in xp[pc]).
template <class T, class XT, class WT, int NK>
int bat_pcg_seq(tc_batch* S, int xc, int pc, int K, hipStream_t s) {
  const tc_batch_desc& d = S->d;
  const BatOp<T> op = make_op<T>(S);
  XT* xs;
  if (d.linear_guess) {
    bat_rhs_launch<T, XT, WT, true, NK>(S, op, xc, pc, s);
    xs = (XT*)d.x[1 - xc];
  } else {
    bat_rhs_launch<T, XT, WT, false, NK>(S, op, xc, pc, s);
    xs = (XT*)d.x[xc];
  }
  TC_LAUNCH();
  hipLaunchKernelGGL(k_bat_fold, dim3(d.B), dim3(kBB), 0, s, 0, 0, S->G, d.lanes, d.B, S->ctl,
                     (const double*)S->part[0]);
  TC_LAUNCH();
  return bat_iters<T, XT, NK>(S, 0, K, xs, s);
}

template <class T, class XT, class WT, int NK>
int bat_enqueue(tc_batch* S, BHost& ps, BPlan& plan, hipStream_t s) {
  const tc_batch_desc& d = S->d;
  const int64_t step = ps.steps + 1;
  const double t = ps.t + d.dt;
  const int slot = (int)(step % kBRing);
  int K = 1;
  for (int v : S->recent) K = v > K ? v : K;
  if (K > d.max_iter) K = d.max_iter;
  const int xc = ps.xc, pc = ps.pc;
  // 1. ionic, one launch per model group
  const uint32_t bits = stimulus_bits(d.stim, d.n_stim, t);
  for (int g = 0; g < d.n_groups; ++g) {
    const tc_batch_group& G = d.groups[g];
    tc_ionic_io io{};
    io.n = d.n * G.count;
    io.dt = d.dt_ionic;
    io.dt_rhs = d.dt;
    io.Cm = d.Cm;
    io.V = d.x[xc];
    io.v_f64 = (sizeof(XT) == 8 && sizeof(T) == 4) ? 1 : 0;
    io.w_f64 = (sizeof(WT) == 8 && sizeof(T) == 4) ? 1 : 0;
    io.w = d.w;
    io.stim_mask = bits ? d.stim_mask : nullptr;
    io.stim_active = bits;
    for (int k = 0; k < d.n_stim; ++k) io.stim_amp[k] = d.stim[k].amp;
    if (d.AT && ps.lat_pending) {
      io.AT = d.AT;
      io.RT = d.RT;
      io.lat_t = ps.lat_t;
    }
    io.gate = &S->ctl->stall;
    io.lanes = G.count;
    io.lane_stride = d.lanes;
    io.lane0 = G.b0;
    io.state_stride = G.state_stride;
    int rc = group_ionic(S, g, &io, s);
    if (rc) return rc;
  }
  // 2. RHS + PCG start, 3. speculative iterations: one launch sequence
  // fixed by the buffer parity and K, replayed as a graph in graph mode
  int rc;
  if (d.graph)
    rc = graph_replay(S->graphs, (long long)((xc << 1) | pc) | ((long long)K << 2), 0, s,
                      [&](hipStream_t cs) { return bat_pcg_seq<T, XT, WT, NK>(S, xc, pc, K, cs); });
  else
    rc = bat_pcg_seq<T, XT, WT, NK>(S, xc, pc, K, s);
  if (rc) return rc;
  // 4. check
  plan.seq = ++S->seq;
  rc = bat_check(S, K, slot, plan.seq, s);
  if (rc) return rc;
  BHost nx = ps;
  nx.steps = step;
  nx.t = t;
  if (d.linear_guess) {
    nx.xc = 1 - xc;
    nx.pc = 1 - pc;
  }
  nx.lat_pending = true;
  nx.lat_t = t;
  plan.step = step;
  plan.slot = slot;
  plan.K = K;
  plan.after = nx;
  ps = nx;
  return TC_OK;
}

template <class T, class XT, int NK>
int bat_resolve(tc_batch* S, const BPlan& plan, hipStream_t s) {
  const tc_batch_desc& d = S->d;
  StepStatus* st = S->status + plan.slot;
  XT* xs = (XT*)d.x[plan.after.xc];
  int it = plan.K;
  int chunk = it > 0 ? it : 1;
  while (st->stalled) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, &S->ctl->stall, 0);
    TC_LAUNCH();
    const int c = chunk < d.max_iter - it ? chunk : d.max_iter - it;
    int rc = bat_iters<T, XT, NK>(S, it, it + c, xs, s);
    if (rc) return rc;
    it += c;
    const long long seq = ++S->seq;
    rc = bat_check(S, it, plan.slot, seq, s);
    if (rc) return rc;
    rc = bat_wait(S, plan.slot, seq);
    if (rc) return rc;
    chunk = chunk * 2 < 64 ? chunk * 2 : 64;
  }
  return TC_OK;
}

template <class T, class XT, class WT, int NK>
int bat_advance(tc_batch* S, int64_t nsteps, int32_t* iters, hipStream_t s) {
  const int depth = S->d.depth < 1 ? 1 : (S->d.depth > kBRing / 2 ? kBRing / 2 : S->d.depth);
  std::deque<BPlan> inflight;
  BHost ps = S->vs;
  int64_t enq = 0, done = 0;
  while (done < nsteps) {
    while ((int)inflight.size() < depth && enq < nsteps) {
      BPlan plan;
      int rc = bat_enqueue<T, XT, WT, NK>(S, ps, plan, s);
      if (rc) return rc;
      inflight.push_back(plan);
      ++enq;
    }
    BPlan plan = inflight.front();
    int rc = bat_wait(S, plan.slot, plan.seq);
    if (rc) return rc;
    StepStatus* st = S->status + plan.slot;
    if (st->stalled) {
      rc = bat_resolve<T, XT, NK>(S, plan, s);
      if (rc) return rc;
      enq -= (int64_t)inflight.size() - 1;
      inflight.clear();
      ps = plan.after;
    } else {
      inflight.pop_front();
    }
    iters[done] = st->n_iter;
    S->recent.push_back(st->iter > 0 ? st->iter : 1);
    if ((int)S->recent.size() > kBHistory) S->recent.pop_front();
    S->vs = plan.after;
    ++done;
  }
  TC_HIP(hipStreamSynchronize(s));
  return TC_OK;
}

template <class T, class XT, class WT>
int bat_advance_nk(tc_batch* S, int64_t nsteps, int32_t* iters, hipStream_t s) {
  switch (S->d.NK) {
    case 1: return bat_advance<T, XT, WT, 1>(S, nsteps, iters, s);
    case 3: return bat_advance<T, XT, WT, 3>(S, nsteps, iters, s);
    case 5: return bat_advance<T, XT, WT, 5>(S, nsteps, iters, s);
  }
  set_error("batch: NK must be 1, 3 or 5 (got %d)", S->d.NK);
  return TC_ERR_ARG;
}

template <class T, int NK>
int bat_dinv(tc_batch* S, hipStream_t s) {
  const BatOp<T> op = make_op<T>(S);
  const int64_t tot = S->d.n * S->d.lanes;
  hipLaunchKernelGGL((k_bat_dinv<T, NK>), dim3((int)ceil_div(tot, 256)), dim3(256), 0, s, op,
                     (T*)S->d.Dinv);
  TC_LAUNCH();
  return TC_OK;
}

template <class T>
int bat_dinv_nk(tc_batch* S, hipStream_t s) {
  switch (S->d.NK) {
    case 1: return bat_dinv<T, 1>(S, s);
    case 3: return bat_dinv<T, 3>(S, s);
    case 5: return bat_dinv<T, 5>(S, s);
  }
  set_error("batch: NK must be 1, 3 or 5 (got %d)", S->d.NK);
  return TC_ERR_ARG;
}

}  // namespace
}  // namespace tc

using namespace tc;

extern "C" int tc_batch_workspace_size(int64_t n, int lanes, int dtype, size_t* bytes) {
  TC_CHECK(bytes && n > 0 && log2_exact(lanes) >= 0 && lanes <= kLanes, TC_ERR_ARG,
           "tc_batch_workspace_size: bad arguments");
  TC_CHECK(dtype == TC_F32 || dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", dtype);
  *bytes = ws_layout(n, lanes, dtype).total;
  return TC_OK;
}

extern "C" int tc_batch_create(tc_batch** out, const tc_batch_desc* d, void* stream) {
  TC_CHECK(out && d && d->n > 0 && d->row_ptr && d->col && d->Mv && d->coef && d->w && d->Dinv,
           TC_ERR_ARG, "tc_batch_create: bad arguments");
  TC_CHECK(d->B >= 1 && d->B <= kLanes && d->lanes >= d->B && d->lanes <= kLanes &&
               log2_exact(d->lanes) >= 0,
           TC_ERR_ARG, "tc_batch_create: need 1 <= B <= lanes <= 64, lanes a power of two");
  TC_CHECK(d->NK == 1 || d->NK == 3 || d->NK == 5, TC_ERR_ARG, "NK must be 1, 3 or 5");
  for (int k = 0; k < d->NK; ++k) TC_CHECK(d->Kv[k], TC_ERR_ARG, "missing K component %d", k);
  TC_CHECK(d->dtype == TC_F32 || d->dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", d->dtype);
  TC_CHECK(!d->w_f64 || d->x_f64 || d->dtype == TC_F64, TC_ERR_ARG, "w_f64 needs x_f64");
  TC_CHECK(d->x[0] && d->x[1] && (!d->linear_guess || (d->xp[0] && d->xp[1])), TC_ERR_ARG,
           "tc_batch_create: need two x buffers (and two x_prev buffers with the linear guess)");
  TC_CHECK(d->n_stim >= 0 && d->n_stim <= TC_MAX_STIMULI, TC_ERR_ARG, "bad stimulus count");
  TC_CHECK(d->n_stim == 0 || d->stim_mask, TC_ERR_ARG, "stimuli need a mask");
  TC_CHECK(!d->AT || d->RT, TC_ERR_ARG, "AT without RT");
  TC_CHECK(d->max_iter >= 0 && d->max_iter <= kMaxIterLimit, TC_ERR_ARG,
           "max_iter %d outside 0..%d", d->max_iter, kMaxIterLimit);
  TC_CHECK(d->n_groups >= 1 && d->n_groups <= TC_BATCH_MAX_GROUPS, TC_ERR_ARG,
           "bad group count %d", d->n_groups);
  TC_CHECK(d->n * d->lanes < (int64_t)INT32_MAX * 4, TC_ERR_SIZE, "batch too large");
  {
    // every instance in exactly one group
    unsigned long long seen = 0;
    for (int g = 0; g < d->n_groups; ++g) {
      const tc_batch_group& G = d->groups[g];
      TC_CHECK(G.states && G.params && G.count >= 1 && G.b0 >= 0 && G.b0 + G.count <= d->B,
               TC_ERR_ARG, "batch group %d: bad instance range or buffers", g);
      for (int b = G.b0; b < G.b0 + G.count; ++b) {
        TC_CHECK(!(seen >> b & 1ull), TC_ERR_ARG, "instance %d is in two model groups", b);
        seen |= 1ull << b;
      }
    }
    const unsigned long long all = d->B == 64 ? ~0ull : (1ull << d->B) - 1;
    TC_CHECK(seen == all, TC_ERR_ARG, "the model groups do not cover all %d instances", d->B);
  }
  TC_CHECK(d->max_blocks == 0 || (d->max_blocks >= kXcd && d->max_blocks <= kBMaxBlocks),
           TC_ERR_ARG, "max_blocks %d outside %d..%d (0: default)", d->max_blocks, kXcd,
           kBMaxBlocks);
  size_t need = 0;
  int rc = tc_batch_workspace_size(d->n, d->lanes, d->dtype, &need);
  if (rc) return rc;
  TC_CHECK(d->ws && d->ws_bytes >= need, TC_ERR_SIZE, "batch workspace too small (%zu < %zu)",
           d->ws_bytes, need);
  tc_batch* S = new tc_batch();
  S->d = *d;
  for (int g = 0; g < d->n_groups; ++g) {
    switch (d->groups[g].model) {
      case TC_MODEL_MS2V: S->gp[g].ms2v = *(const tc_ms2v_params*)d->groups[g].params; break;
      case TC_MODEL_TTP: S->gp[g].ttp = *(const tc_ttp_params*)d->groups[g].params; break;
      case TC_MODEL_CRN: S->gp[g].crn = *(const tc_crn_params*)d->groups[g].params; break;
      default:
        delete S;
        set_error("batch group %d: unknown model %d", g, d->groups[g].model);
        return TC_ERR_ARG;
    }
    S->d.groups[g].params = nullptr;
  }
  S->lg = log2_exact(d->lanes);
  S->G = bat_blocks(d->n, d->lanes, d->max_blocks > 0 ? d->max_blocks : kBMaxBlocks);
  {
    const int64_t gp = ceil_div(d->n * d->lanes, 256 * 4);
    S->Gp = (int)(gp < 1 ? 1 : (gp > 8192 ? 8192 : gp));
  }
  const WsLayout L = ws_layout(d->n, d->lanes, d->dtype);
  char* base = (char*)d->ws;
  S->ctl = (BatCtl*)(base + L.ctl);
  for (int k = 0; k < 3; ++k) S->part[k] = (double*)(base + L.part[k]);
  S->r = base + L.r;
  S->p[0] = base + L.p0;
  S->p[1] = base + L.p1;
  S->Ap = base + L.Ap;
  hipError_t e = hipHostMalloc((void**)&S->host_status, sizeof(StatusRec) * kBRing,
                               hipHostMallocMapped);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&S->dev_view, S->host_status, 0);
  if (e != hipSuccess) {
    set_error("tc_batch_create: %s", hipGetErrorString(e));
    tc_batch_destroy(S);
    return TC_ERR_HIP;
  }
  status_clear(S->host_status, kBRing);
  S->seq = 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_bat_init_ctl, dim3(1), dim3(kLanes), 0, s, S->ctl, d->a_tol, d->r_tol);
  e = hipGetLastError();
  if (e == hipSuccess) {
    rc = d->dtype == TC_F64 ? bat_dinv_nk<double>(S, s) : bat_dinv_nk<float>(S, s);
    if (rc) {
      tc_batch_destroy(S);
      return rc;
    }
    e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    set_error("tc_batch_create: %s", hipGetErrorString(e));
    tc_batch_destroy(S);
    return TC_ERR_HIP;
  }
  *out = S;
  return TC_OK;
}

extern "C" int tc_batch_destroy(tc_batch* S) {
  if (!S) return TC_OK;
  S->graphs.release();
  if (S->host_status) (void)hipHostFree(S->host_status);
  delete S;
  return TC_OK;
}

extern "C" int tc_batch_graph_stats(const tc_batch* S, int64_t* captures, int64_t* replays) {
  TC_CHECK(S && captures && replays, TC_ERR_ARG, "tc_batch_graph_stats: bad arguments");
  *captures = S->graphs.captures;
  *replays = S->graphs.replays;
  return TC_OK;
}

extern "C" int tc_batch_advance(tc_batch* S, int64_t nsteps, int32_t* iters, void* stream) {
  TC_CHECK(S && iters && nsteps >= 0, TC_ERR_ARG, "tc_batch_advance: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (S->d.dtype == TC_F64) return bat_advance_nk<double, double, double>(S, nsteps, iters, s);
  if (S->d.x_f64 && S->d.w_f64) return bat_advance_nk<float, double, double>(S, nsteps, iters, s);
  if (S->d.x_f64) return bat_advance_nk<float, double, float>(S, nsteps, iters, s);
  return bat_advance_nk<float, float, float>(S, nsteps, iters, s);
}

extern "C" int tc_batch_finish(tc_batch* S, int64_t* total_iters, void* stream) {
  TC_CHECK(S && total_iters, TC_ERR_ARG, "tc_batch_finish: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const tc_batch_desc& d = S->d;
  if (d.AT && S->vs.lat_pending) {
    const int64_t m = d.n * d.lanes;
    const dim3 g((int)ceil_div(m, 256)), b(256);
    const void* x = d.x[S->vs.xc];
    if (d.dtype == TC_F64)
      hipLaunchKernelGGL((k_lat_x<double, double>), g, b, 0, s, (const double*)x,
                         (double*)d.AT, (double*)d.RT, m, S->vs.lat_t);
    else if (d.x_f64)
      hipLaunchKernelGGL((k_lat_x<float, double>), g, b, 0, s, (const double*)x, (float*)d.AT,
                         (float*)d.RT, m, S->vs.lat_t);
    else
      hipLaunchKernelGGL((k_lat_x<float, float>), g, b, 0, s, (const float*)x, (float*)d.AT,
                         (float*)d.RT, m, S->vs.lat_t);
    TC_LAUNCH();
    S->vs.lat_pending = false;
  }
  long long tot[kLanes];
  TC_HIP(hipMemcpyAsync(tot, S->ctl->total, sizeof(long long) * kLanes, hipMemcpyDeviceToHost, s));
  TC_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < d.B; ++b) total_iters[b] = tot[b];
  return TC_OK;
}

extern "C" int tc_batch_status(tc_batch* S, double* t, int64_t* steps, int* x_index) {
  TC_CHECK(S && t && steps && x_index, TC_ERR_ARG, "tc_batch_status: bad arguments");
  *t = S->vs.t;
  *steps = S->vs.steps;
  *x_index = S->vs.xc;
  return TC_OK;
}

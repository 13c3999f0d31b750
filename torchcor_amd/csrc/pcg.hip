// Right-hand side and Jacobi-PCG on SELL-64 operators.
//
// One thread owns one row; a wavefront owns one 64-row slice, so the value
// and column loads of a slice are fully coalesced (column-major padding) and
// no cross-lane reduction is needed for a row.  Every product accumulates
// in float64.  Each PCG iteration is two passes:
//   spmv pass:   p' = beta p + Dinv r (first iteration: p' = Dinv r),
//                Ap = A p' with p' formed on the fly for the gathered
//                neighbours, partial sums of p'.Ap
//   update pass: x += alpha p', r -= alpha Ap, partial sums of z.z and r.z
//                with z = Dinv r.
//
// Reductions without atomics: the PCG kernels run on a fixed grid of G
// blocks; each block writes its partial sums to partials[k*G + block], and
// every block of the NEXT pass folds all G partials itself, in block order,
// before using the scalar.  All blocks therefore see bit-identical scalars
// and the result does not depend on scheduling.  The stopping test of
// iteration k is evaluated by the spmv pass of iteration k+1 (or by a one-
// block check pass at the end of a launch chunk), exactly with the reference
// rule, so no host round trip is needed inside an iteration.
//
// XCD-aware work split: block b runs on XCD b % 8; blocks of one XCD get one
// contiguous range of slices so that neighbour gathers stay in that XCD's L2.
#include <vector>

#include "tc_internal.h"

namespace tc {

struct PcgCtl {
  double rz[2];   // r.z of iteration k in rz[k & 1]
  double bnorm;   // ||Dinv b||
  double a_tol, r_tol;
  int converged;  // stop rule met
  int n_iter;     // iterations that did not stop (reference convention)
  int iter;       // iterations executed
  int pad;
};

}  // namespace tc

struct tc_pcg {
  tc_sell A;
  const void* Av;
  const void* Dinv;
  int dtype;
  int64_t n;
  int G;                 // blocks of every PCG pass
  tc::PcgCtl* ctl;       // device
  double* part[3];       // device partial sums: init (2G), spmv (G), update (2G)
  void* r;
  void* p[2];
  void* Ap;
  tc::PcgCtl* host_ctl;  // pinned
  int chunk;             // iterations launched before the first status read
  // Optional per-kernel timing with HIP events on the solve's stream:
  // pairs around each launch; index 0 init, then (spmv, update) per iteration.
  int timing;
  std::vector<hipEvent_t> ev;
  double acc_ms[3];      // init, spmv, update
  int64_t acc_n[3];
};

namespace tc {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxBlocks = 2048;
constexpr int kXcd = 8;

// ------------------------------------------------------------------ SpMV
template <class T>
__global__ void __launch_bounds__(kBlock)
k_spmv(int64_t n, const int32_t* __restrict__ sptr, const int32_t* __restrict__ scol,
       const T* __restrict__ v, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t s = i >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t ns = (n + 63) >> 6;
  if (s >= ns) return;
  const int32_t b = sptr[s], w = (sptr[s + 1] - b) >> 6;
  const int32_t* c = scol + b + lane;
  const T* a = v + b + lane;
  double acc = 0.0;
  for (int j = 0; j < w; ++j) acc += (double)a[j * 64] * (double)x[c[j * 64]];
  if (i < n) y[i] = (T)acc;
}

// b = M w - coef K u
template <class T>
__global__ void __launch_bounds__(kBlock)
k_rhs(int64_t n, const int32_t* __restrict__ sptr, const int32_t* __restrict__ scol,
      const T* __restrict__ Mv, const T* __restrict__ Kv, const T* __restrict__ w,
      const T* __restrict__ u, double coef, T* __restrict__ bout) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t s = i >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t ns = (n + 63) >> 6;
  if (s >= ns) return;
  const int32_t b = sptr[s], wd = (sptr[s + 1] - b) >> 6;
  const int32_t* c = scol + b + lane;
  const T* m = Mv + b + lane;
  const T* k = Kv + b + lane;
  double am = 0.0, ak = 0.0;
  for (int j = 0; j < wd; ++j) {
    const int32_t cj = c[j * 64];
    am += (double)m[j * 64] * (double)w[cj];
    ak += (double)k[j * 64] * (double)u[cj];
  }
  if (i < n) bout[i] = (T)(am - coef * ak);
}

// ------------------------------------------------------- PCG work split
struct Split {
  int64_t s0, s1;  // this block's slice range
};

__device__ inline Split block_slices(int64_t ns) {
  const int G = gridDim.x;
  const int b = blockIdx.x;
  int vb = b;
  if (G % kXcd == 0) vb = (b % kXcd) * (G / kXcd) + b / kXcd;
  const int64_t per = (ns + G - 1) / G;
  Split sp;
  sp.s0 = (int64_t)vb * per;
  sp.s1 = sp.s0 + per < ns ? sp.s0 + per : ns;
  return sp;
}

// Block partial sums -> partials[k*G + blockIdx.x].
template <int NV>
__device__ inline void store_partials(double (&v)[NV], double* partials) {
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
  }
}

// Sum of the previous pass's G partials, in a fixed order, returned to
// every thread of the block.
template <int NV>
__device__ inline void fold_partials(const double* __restrict__ partials, int G,
                                     double (&out)[NV]) {
  __shared__ double bc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) s += partials[k * G + b];
    out[k] = s;
  }
  block_sum<NV>(out);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) bc[k] = out[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = bc[k];
}

// Block-uniform read of the converged flag (one load, shared by all waves).
__device__ inline bool block_converged(const PcgCtl* ctl) {
  __shared__ int conv;
  if (threadIdx.x == 0) conv = ctl->converged;
  __syncthreads();
  return conv != 0;
}

__device__ inline bool stop_rule(double zn, const PcgCtl* ctl) {
  return zn < ctl->a_tol || zn / ctl->bnorm < ctl->r_tol;
}

// PCG start: x0 = 2x - x_prev (linear guess) or x; r = b - A x0;
// partial sums of r.z and (Dinv b)^2.
template <class T, bool LG>
__global__ void __launch_bounds__(kBlock)
k_pcg_init(int64_t n, const int32_t* __restrict__ sptr, const int32_t* __restrict__ scol,
           const T* __restrict__ Av, const T* __restrict__ Dinv, const T* __restrict__ bvec,
           const T* __restrict__ x, const T* __restrict__ xp, T* __restrict__ xnext,
           T* __restrict__ r, PcgCtl* ctl, double a_tol, double r_tol, double* part0) {
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double red[2] = {0.0, 0.0};
  for (int64_t s = sp.s0 + wv; s < sp.s1; s += kWaves) {
    const int64_t i = s * 64 + lane;
    const int32_t b = sptr[s], wd = (sptr[s + 1] - b) >> 6;
    const int32_t* c = scol + b + lane;
    const T* a = Av + b + lane;
    double ax = 0.0;
    for (int j = 0; j < wd; ++j) {
      const int32_t cj = c[j * 64];
      T xj;
      if (LG) xj = (T)(2.0 * (double)x[cj] - (double)xp[cj]);
      else xj = x[cj];
      ax += (double)a[j * 64] * (double)xj;
    }
    if (i < n) {
      if (LG) xnext[i] = (T)(2.0 * (double)x[i] - (double)xp[i]);
      const double d = (double)Dinv[i];
      const double bi = (double)bvec[i];
      const T ri = (T)(bi - ax);
      r[i] = ri;
      red[0] += (double)ri * (d * (double)ri);
      const double db = d * bi;
      red[1] += db * db;
    }
  }
  store_partials<2>(red, part0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->converged = 0;
    ctl->n_iter = 0;
    ctl->iter = 0;
    ctl->a_tol = a_tol;
    ctl->r_tol = r_tol;
  }
}

// spmv pass of iteration k.  pin is the previous p (unused when k == 0),
// pout receives p'.  Iteration k-1's stopping test happens here first.
template <class T>
__global__ void __launch_bounds__(kBlock)
k_pcg_spmv(int64_t n, const int32_t* __restrict__ sptr, const int32_t* __restrict__ scol,
           const T* __restrict__ Av, const T* __restrict__ Dinv, const T* __restrict__ r,
           const T* __restrict__ pin, T* __restrict__ pout, T* __restrict__ Ap, int k,
           PcgCtl* ctl, const double* __restrict__ part0, const double* __restrict__ part2,
           double* part1) {
  if (block_converged(ctl)) return;
  const int G = gridDim.x;
  double beta = 0.0, rz;
  if (k == 0) {
    double v[2];
    fold_partials<2>(part0, G, v);
    rz = v[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->bnorm = sqrt(v[1]);
  } else {
    double v[2];
    fold_partials<2>(part2, G, v);
    if (stop_rule(sqrt(v[0]), ctl)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->converged = 1;
        ctl->n_iter = k - 1;
        ctl->iter = k;
      }
      return;
    }
    rz = v[1];
    beta = rz / ctl->rz[(k - 1) & 1];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->rz[k & 1] = rz;
    ctl->n_iter = k;
    ctl->iter = k;
  }
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double red[1] = {0.0};
  for (int64_t s = sp.s0 + wv; s < sp.s1; s += kWaves) {
    const int64_t i = s * 64 + lane;
    const int32_t b = sptr[s], wd = (sptr[s + 1] - b) >> 6;
    const int32_t* c = scol + b + lane;
    const T* a = Av + b + lane;
    double acc = 0.0;
    if (k == 0) {
      for (int j = 0; j < wd; ++j) {
        const int32_t cj = c[j * 64];
        const T pj = (T)((double)Dinv[cj] * (double)r[cj]);
        acc += (double)a[j * 64] * (double)pj;
      }
    } else {
      for (int j = 0; j < wd; ++j) {
        const int32_t cj = c[j * 64];
        const T pj = (T)(beta * (double)pin[cj] + (double)Dinv[cj] * (double)r[cj]);
        acc += (double)a[j * 64] * (double)pj;
      }
    }
    if (i < n) {
      T pi;
      if (k == 0) pi = (T)((double)Dinv[i] * (double)r[i]);
      else pi = (T)(beta * (double)pin[i] + (double)Dinv[i] * (double)r[i]);
      const T api = (T)acc;
      pout[i] = pi;
      Ap[i] = api;
      red[0] += (double)pi * (double)api;
    }
  }
  store_partials<1>(red, part1);
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_pcg_update(int64_t n, const T* __restrict__ Dinv, const T* __restrict__ p,
             const T* __restrict__ Ap, T* __restrict__ x, T* __restrict__ r, int k, PcgCtl* ctl,
             const double* __restrict__ part1, double* part2) {
  if (block_converged(ctl)) return;
  double v[1];
  fold_partials<1>(part1, gridDim.x, v);
  const double alpha = ctl->rz[k & 1] / v[0];
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double red[2] = {0.0, 0.0};
  for (int64_t s = sp.s0 + wv; s < sp.s1; s += kWaves) {
    const int64_t i = s * 64 + lane;
    if (i < n) {
      const T xi = (T)((double)x[i] + alpha * (double)p[i]);
      const T ri = (T)((double)r[i] - alpha * (double)Ap[i]);
      x[i] = xi;
      r[i] = ri;
      const double z = (double)Dinv[i] * (double)ri;
      red[0] += z * z;
      red[1] += (double)ri * z;
    }
  }
  store_partials<2>(red, part2);
}

// One block: stopping test of the last update of a launch chunk (iteration
// k), so the host can read the state.
__global__ void __launch_bounds__(kBlock)
k_pcg_check(int k, int G, PcgCtl* ctl, const double* __restrict__ part2) {
  if (block_converged(ctl)) return;
  double v[2];
  fold_partials<2>(part2, G, v);
  if (threadIdx.x == 0) {
    const bool stop = stop_rule(sqrt(v[0]), ctl);
    ctl->converged = stop ? 1 : 0;
    ctl->n_iter = stop ? k : k + 1;
    ctl->iter = k + 1;
  }
}

size_t vec_bytes(int64_t n, int dtype) {
  return align_up((size_t)n * (dtype == TC_F64 ? 8 : 4), 256);
}

int pcg_blocks(int64_t n) {
  const int64_t ns = ceil_div(n, 64);
  int64_t g = ceil_div(ns, kWaves);
  if (g > kMaxBlocks) g = kMaxBlocks;
  if (g >= kXcd) g -= g % kXcd;
  return (int)(g < 1 ? 1 : g);
}

// Event pair k brackets launch k of the current solve.
int ev_reserve(tc_pcg* P, size_t pairs) {
  while (P->ev.size() < 2 * pairs) {
    hipEvent_t e;
    TC_HIP(hipEventCreate(&e));
    P->ev.push_back(e);
  }
  return TC_OK;
}

template <class T>
int pcg_solve(tc_pcg* P, const T* b, T* x, const T* xp, T* xn, double a_tol, double r_tol,
              int max_iter, int lg, int* n_iter_host, hipStream_t s) {
  const int64_t n = P->n;
  const dim3 g(P->G), blk(kBlock);
  const int32_t* sp = P->A.slice_ptr;
  const int32_t* sc = P->A.col;
  const T* Av = (const T*)P->Av;
  const T* D = (const T*)P->Dinv;
  T* r = (T*)P->r;
  T* Ap = (T*)P->Ap;
  T* pp[2] = {(T*)P->p[0], (T*)P->p[1]};
  if (P->timing) {
    int rc = ev_reserve(P, 1 + 2 * (size_t)(max_iter > 0 ? max_iter : 0));
    if (rc) return rc;
    TC_HIP(hipEventRecord(P->ev[0], s));
  }
  T* xs;
  if (lg) {
    hipLaunchKernelGGL((k_pcg_init<T, true>), g, blk, 0, s, n, sp, sc, Av, D, b, x, xp, xn, r,
                       P->ctl, a_tol, r_tol, P->part[0]);
    xs = xn;
  } else {
    hipLaunchKernelGGL((k_pcg_init<T, false>), g, blk, 0, s, n, sp, sc, Av, D, b, x, xp,
                       (T*)nullptr, r, P->ctl, a_tol, r_tol, P->part[0]);
    xs = x;
  }
  TC_LAUNCH();
  if (P->timing) TC_HIP(hipEventRecord(P->ev[1], s));
  int it = 0;
  int chunk = P->chunk < 1 ? 1 : P->chunk;
  P->host_ctl->converged = 0;
  P->host_ctl->iter = 0;
  P->host_ctl->n_iter = 0;
  while (it < max_iter) {
    const int c = chunk < max_iter - it ? chunk : max_iter - it;
    for (int j = 0; j < c; ++j) {
      const int k = it + j;
      T* pin = pp[k & 1];
      T* pout = pp[(k + 1) & 1];
      if (P->timing) TC_HIP(hipEventRecord(P->ev[2 + 4 * k], s));
      hipLaunchKernelGGL(k_pcg_spmv<T>, g, blk, 0, s, n, sp, sc, Av, D, r, pin, pout, Ap, k,
                         P->ctl, P->part[0], P->part[2], P->part[1]);
      TC_LAUNCH();
      if (P->timing) {
        TC_HIP(hipEventRecord(P->ev[3 + 4 * k], s));
        TC_HIP(hipEventRecord(P->ev[4 + 4 * k], s));
      }
      hipLaunchKernelGGL(k_pcg_update<T>, g, blk, 0, s, n, D, pout, Ap, xs, r, k, P->ctl,
                         P->part[1], P->part[2]);
      TC_LAUNCH();
      if (P->timing) TC_HIP(hipEventRecord(P->ev[5 + 4 * k], s));
    }
    it += c;
    hipLaunchKernelGGL(k_pcg_check, dim3(1), blk, 0, s, it - 1, P->G, P->ctl, P->part[2]);
    TC_LAUNCH();
    TC_HIP(hipMemcpyAsync(P->host_ctl, P->ctl, sizeof(PcgCtl), hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    if (P->host_ctl->converged) break;
    chunk = chunk * 2 < 64 ? chunk * 2 : 64;
  }
  if (max_iter == 0) TC_HIP(hipStreamSynchronize(s));
  *n_iter_host = P->host_ctl->n_iter;
  P->chunk = P->host_ctl->iter > 0 ? P->host_ctl->iter : 1;
  if (P->timing) {
    // Only iterations that ran count (later launches exit at once).
    float ms = 0.f;
    TC_HIP(hipEventElapsedTime(&ms, P->ev[0], P->ev[1]));
    P->acc_ms[0] += ms;
    P->acc_n[0] += 1;
    for (int k = 0; k < P->host_ctl->iter && k < it; ++k) {
      TC_HIP(hipEventElapsedTime(&ms, P->ev[2 + 4 * k], P->ev[3 + 4 * k]));
      P->acc_ms[1] += ms;
      TC_HIP(hipEventElapsedTime(&ms, P->ev[4 + 4 * k], P->ev[5 + 4 * k]));
      P->acc_ms[2] += ms;
      P->acc_n[1] += 1;
      P->acc_n[2] += 1;
    }
  }
  return TC_OK;
}

}  // namespace
}  // namespace tc

using namespace tc;

extern "C" int tc_sell_spmv(const tc_sell* A, const void* v, const void* x, void* y,
                            int dtype, void* stream) {
  TC_CHECK(A && v && x && y && A->n > 0, TC_ERR_ARG, "tc_sell_spmv: bad arguments");
  dim3 g((int)ceil_div(A->n_slices * 64, kBlock)), b(kBlock);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TC_F32)
    hipLaunchKernelGGL(k_spmv<float>, g, b, 0, s, A->n, A->slice_ptr, A->col, (const float*)v,
                       (const float*)x, (float*)y);
  else if (dtype == TC_F64)
    hipLaunchKernelGGL(k_spmv<double>, g, b, 0, s, A->n, A->slice_ptr, A->col,
                       (const double*)v, (const double*)x, (double*)y);
  else
    TC_CHECK(false, TC_ERR_ARG, "bad dtype %d", dtype);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_rhs(const tc_sell* A, const void* Mv, const void* Kv, const void* w,
                      const void* u, double coef, void* b, int dtype, void* stream) {
  TC_CHECK(A && Mv && Kv && w && u && b && A->n > 0, TC_ERR_ARG, "tc_rhs: bad arguments");
  dim3 g((int)ceil_div(A->n_slices * 64, kBlock)), bl(kBlock);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TC_F32)
    hipLaunchKernelGGL(k_rhs<float>, g, bl, 0, s, A->n, A->slice_ptr, A->col,
                       (const float*)Mv, (const float*)Kv, (const float*)w, (const float*)u,
                       coef, (float*)b);
  else if (dtype == TC_F64)
    hipLaunchKernelGGL(k_rhs<double>, g, bl, 0, s, A->n, A->slice_ptr, A->col,
                       (const double*)Mv, (const double*)Kv, (const double*)w,
                       (const double*)u, coef, (double*)b);
  else
    TC_CHECK(false, TC_ERR_ARG, "bad dtype %d", dtype);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_pcg_workspace_size(int64_t n, int dtype, size_t* bytes) {
  TC_CHECK(bytes && n > 0 && (dtype == TC_F32 || dtype == TC_F64), TC_ERR_ARG,
           "tc_pcg_workspace_size: bad arguments");
  const int G = pcg_blocks(n);
  *bytes = align_up(sizeof(PcgCtl), 256) + 3 * align_up(sizeof(double) * 2 * G, 256) +
           4 * vec_bytes(n, dtype);
  return TC_OK;
}

extern "C" int tc_pcg_create(tc_pcg** plan, const tc_sell* A, const void* Av, const void* Dinv,
                             int dtype, void* ws, size_t ws_bytes) {
  TC_CHECK(plan && A && Av && Dinv && ws && A->n > 0, TC_ERR_ARG, "tc_pcg_create: bad arguments");
  TC_CHECK(dtype == TC_F32 || dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", dtype);
  size_t need = 0;
  int rc = tc_pcg_workspace_size(A->n, dtype, &need);
  if (rc) return rc;
  TC_CHECK(ws_bytes >= need, TC_ERR_SIZE, "pcg workspace too small (%zu < %zu)", ws_bytes, need);
  tc_pcg* P = new tc_pcg();
  P->A = *A;
  P->Av = Av;
  P->Dinv = Dinv;
  P->dtype = dtype;
  P->n = A->n;
  P->G = pcg_blocks(A->n);
  char* b = (char*)ws;
  size_t off = 0;
  P->ctl = (PcgCtl*)(b + off);
  off += align_up(sizeof(PcgCtl), 256);
  for (int k = 0; k < 3; ++k) {
    P->part[k] = (double*)(b + off);
    off += align_up(sizeof(double) * 2 * P->G, 256);
  }
  P->r = b + off; off += vec_bytes(P->n, dtype);
  P->p[0] = b + off; off += vec_bytes(P->n, dtype);
  P->p[1] = b + off; off += vec_bytes(P->n, dtype);
  P->Ap = b + off; off += vec_bytes(P->n, dtype);
  P->chunk = 1;
  P->timing = 0;
  for (int k = 0; k < 3; ++k) {
    P->acc_ms[k] = 0.0;
    P->acc_n[k] = 0;
  }
  hipError_t e = hipHostMalloc((void**)&P->host_ctl, sizeof(PcgCtl), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete P;
    set_error("hipHostMalloc: %s", hipGetErrorString(e));
    return TC_ERR_HIP;
  }
  // Zero the control block; plan creation is a set-up call, never captured.
  e = hipMemset(P->ctl, 0, sizeof(PcgCtl));
  if (e != hipSuccess) {
    (void)hipHostFree(P->host_ctl);
    delete P;
    set_error("hipMemset: %s", hipGetErrorString(e));
    return TC_ERR_HIP;
  }
  *plan = P;
  return TC_OK;
}

extern "C" int tc_pcg_set_timing(tc_pcg* P, int on) {
  TC_CHECK(P, TC_ERR_ARG, "tc_pcg_set_timing: null plan");
  P->timing = on ? 1 : 0;
  for (int k = 0; k < 3; ++k) {
    P->acc_ms[k] = 0.0;
    P->acc_n[k] = 0;
  }
  return TC_OK;
}

extern "C" int tc_pcg_get_timing(tc_pcg* P, double* ms_host, int64_t* count_host) {
  TC_CHECK(P && ms_host && count_host, TC_ERR_ARG, "tc_pcg_get_timing: bad arguments");
  for (int k = 0; k < 3; ++k) {
    ms_host[k] = P->acc_ms[k];
    count_host[k] = P->acc_n[k];
  }
  return TC_OK;
}

extern "C" int tc_pcg_destroy(tc_pcg* P) {
  if (!P) return TC_OK;
  for (hipEvent_t e : P->ev) (void)hipEventDestroy(e);
  if (P->host_ctl) (void)hipHostFree(P->host_ctl);
  delete P;
  return TC_OK;
}

extern "C" int tc_pcg_solve(tc_pcg* P, const void* b, void* x, const void* x_prev,
                            void* x_next, double a_tol, double r_tol, int max_iter,
                            int linear_guess, int* n_iter_host, void* stream) {
  TC_CHECK(P && b && x && n_iter_host && max_iter >= 0, TC_ERR_ARG, "tc_pcg_solve: bad arguments");
  TC_CHECK(!linear_guess || (x_prev && x_next), TC_ERR_ARG,
           "tc_pcg_solve: linear_guess needs x_prev and x_next");
  hipStream_t s = (hipStream_t)stream;
  if (P->dtype == TC_F32)
    return pcg_solve<float>(P, (const float*)b, (float*)x, (const float*)x_prev,
                            (float*)x_next, a_tol, r_tol, max_iter, linear_guess, n_iter_host, s);
  return pcg_solve<double>(P, (const double*)b, (double*)x, (const double*)x_prev,
                           (double*)x_next, a_tol, r_tol, max_iter, linear_guess, n_iter_host, s);
}

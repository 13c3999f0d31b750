// Stand-alone Jacobi-PCG solve and SpMV/RHS entry points (kernels in
// pcg_kernels.h).
#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

#include "pcg_kernels.h"

namespace tc {
namespace {

// Event pair k brackets launch k of the current solve.
int ev_reserve(tc_pcg* P, size_t pairs) {
  while (P->ev.size() < 2 * pairs) {
    hipEvent_t e;
    // timing only: no system-scope fence (no cache write-back per record)
    TC_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    P->ev.push_back(e);
  }
  return TC_OK;
}

// Device buffers and host state of one solve: a plan's (tc_pcg_create) or
// the library's static scratch array (tc_pcg_solve_static).
struct PcgBufs {
  const tc_sell* A;
  const void* Av;
  const void* Dinv;
  int64_t n;
  int G;
  PcgCtl* ctl;
  double* part[3];
  void* r;
  void* p[2];
  void* Ap;
  PcgCtl* host_ctl;  // where the control block is copied back to
  int* chunk;        // iterations launched before the first status read (hint)
  tc_pcg* timed;     // plan whose event timing applies, or null
};

PcgBufs plan_bufs(tc_pcg* P) {
  PcgBufs B;
  B.A = &P->A;
  B.Av = P->Av;
  B.Dinv = P->Dinv;
  B.n = P->n;
  B.G = P->G;
  B.ctl = P->ctl;
  for (int k = 0; k < 3; ++k) B.part[k] = P->part[k];
  B.r = P->r;
  B.p[0] = P->p[0];
  B.p[1] = P->p[1];
  B.Ap = P->Ap;
  B.host_ctl = P->host_ctl;
  B.chunk = &P->chunk;
  B.timed = P->timing ? P : nullptr;
  return B;
}

template <class T, class CI>
int pcg_solve(const PcgBufs& P, const CI* sc, const T* b, T* x, const T* xp, T* xn,
              double a_tol, double r_tol, int max_iter, int lg, int* n_iter_host, hipStream_t s) {
  const int64_t n = P.n;
  const dim3 g(P.G), blk(kPcgBlock);
  const int32_t* sp = hot_slices(*P.A);
  const T* Av = (const T*)P.Av;
  const T* D = (const T*)P.Dinv;
  T* r = (T*)P.r;
  T* Ap = (T*)P.Ap;
  T* pp[2] = {(T*)P.p[0], (T*)P.p[1]};
  tc_pcg* tm = P.timed;
  if (tm) {
    int rc = ev_reserve(tm, 1 + 2 * (size_t)(max_iter > 0 ? max_iter : 0));
    if (rc) return rc;
  }
  auto ev = [&](size_t k) { return tm ? tm->ev[k] : (hipEvent_t) nullptr; };
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, &P.ctl->stall, 0);
  TC_LAUNCH();
  T* xs;
  if (lg) {
    launch(k_pcg_init<T, true, CI>, g, blk, s, ev(0), ev(1), n, sp, sc, P.A->pat, Av, D, b,
           (const T*)x, xp, xn, r, P.ctl, a_tol, r_tol, P.part[0]);
    xs = xn;
  } else {
    launch(k_pcg_init<T, false, CI>, g, blk, s, ev(0), ev(1), n, sp, sc, P.A->pat, Av, D, b,
           (const T*)x, xp, (T*)nullptr, r, P.ctl, a_tol, r_tol, P.part[0]);
    xs = x;
  }
  TC_LAUNCH();
  int it = 0;
  int chunk = *P.chunk < 1 ? 1 : *P.chunk;
  P.host_ctl->converged = 0;
  P.host_ctl->iter = 0;
  P.host_ctl->n_iter = 0;
  while (it < max_iter) {
    const int c = chunk < max_iter - it ? chunk : max_iter - it;
    for (int j = 0; j < c; ++j) {
      const int k = it + j;
      T* pin = pp[k & 1];
      T* pout = pp[(k + 1) & 1];
      launch(k_pcg_spmv<T, CI>, g, blk, s, ev(2 + 4 * k), ev(3 + 4 * k), n, sp, sc, P.A->pat, Av, D,
             (const T*)r, (const T*)pin, pout, Ap, k, P.ctl, (const double*)P.part[0],
             (const double*)P.part[2], P.part[1]);
      TC_LAUNCH();
      launch(k_pcg_update<T, T>, g, blk, s, ev(4 + 4 * k), ev(5 + 4 * k), n, D, (const T*)pout,
             (const T*)Ap, xs, r, k, P.ctl, (const double*)P.part[1], P.part[2],
             (StatusRec*)nullptr, 0, 0, 0LL);
      TC_LAUNCH();
    }
    it += c;
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(kPcgBlock), 0, s, it - 1, P.G, P.ctl, P.part[2],
                       (StatusRec*)nullptr, 0, max_iter, 0LL);
    TC_LAUNCH();
    TC_HIP(hipMemcpyAsync(P.host_ctl, P.ctl, kCtlHostBytes, hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    if (P.host_ctl->converged) break;
    chunk = chunk * 2 < 64 ? chunk * 2 : 64;
  }
  if (max_iter == 0) TC_HIP(hipStreamSynchronize(s));
  *n_iter_host = P.host_ctl->n_iter;
  *P.chunk = P.host_ctl->iter > 0 ? P.host_ctl->iter : 1;
  if (tm) {
    // Only iterations that ran count (later launches exit at once).
    float ms = 0.f;
    TC_HIP(hipEventElapsedTime(&ms, tm->ev[0], tm->ev[1]));
    tm->acc_ms[0] += ms;
    tm->acc_n[0] += 1;
    for (int k = 0; k < P.host_ctl->iter && k < it; ++k) {
      TC_HIP(hipEventElapsedTime(&ms, tm->ev[2 + 4 * k], tm->ev[3 + 4 * k]));
      tm->acc_ms[1] += ms;
      TC_HIP(hipEventElapsedTime(&ms, tm->ev[4 + 4 * k], tm->ev[5 + 4 * k]));
      tm->acc_ms[2] += ms;
      tm->acc_n[1] += 1;
      tm->acc_n[2] += 1;
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

extern "C" int tc_pcg_set_blocks(tc_pcg* P, int blocks) {
  TC_CHECK(P, TC_ERR_ARG, "tc_pcg_set_blocks: null plan");
  const int mx = pcg_blocks(P->n);  // the workspace holds this many partial slots
  TC_CHECK(blocks >= 0 && blocks <= mx, TC_ERR_ARG,
           "tc_pcg_set_blocks: %d blocks outside 1..%d (0: default)", blocks, mx);
  P->G = blocks > 0 ? blocks : mx;
  ++P->version;
  return TC_OK;
}

extern "C" int tc_pcg_set_stage(tc_pcg* P, const tc_stage* st) {
  TC_CHECK(P, TC_ERR_ARG, "tc_pcg_set_stage: null plan");
  if (!st) {
    P->stage = tc_stage{};
    ++P->version;
    return TC_OK;
  }
  TC_CHECK(st->col_ptr && st->own && st->cols && st->lidx && st->blocks >= 1 &&
               st->max_cols >= 1 && st->max_cols <= TC_STAGE_MAX_COLS &&
               (st->chunk == 0 || st->chunk == 8 || st->chunk == 16),
           TC_ERR_ARG, "tc_pcg_set_stage: bad stage (max_cols 1..%d, chunk 0, 8 or 16)",
           TC_STAGE_MAX_COLS);
  P->stage = *st;
  ++P->version;
  return TC_OK;
}

// Host: the distinct columns of each unit's rows (ascending; a unit is a
// block's slice range, or a run of `chunk` slices of it), each unit's
// first-row position and every entry's position in its unit's set.  A
// unit's columns lie in a window [lo, hi] of the row range; a marker array
// over that window finds the set in one pass over the entries and one over
// the window.
extern "C" int tc_sell_stage(int64_t n, int64_t ns, const int32_t* sptr, const int32_t* col,
                             int G, int chunk, int32_t* col_ptr, int32_t* own, int32_t* cols,
                             uint16_t* lidx, int32_t* max_cols) {
  TC_CHECK(n > 0 && ns == (n + 63) / 64 && sptr && col && G >= 1 && col_ptr && own && cols &&
               lidx && max_cols,
           TC_ERR_ARG, "tc_sell_stage: bad arguments");
  TC_CHECK(chunk == 0 || (chunk % 8 == 0 && chunk <= 16), TC_ERR_ARG,
           "tc_sell_stage: chunk %d is not 0, 8 or 16", chunk);
  const int64_t per = (ns + G - 1) / G;
  // units of `chunk` slices inside each block's range (one unit per block
  // without chunking)
  const int64_t upb = chunk ? (per + chunk - 1) / chunk : 1;
  const int64_t ulen = chunk ? chunk : per;
  std::vector<int32_t> pos;
  int64_t total = 0;
  int32_t mx = 0;
  for (int64_t v = 0; v < (int64_t)G * upb; ++v) {
    col_ptr[v] = (int32_t)total;
    const int64_t b0 = v / upb * per, b1 = b0 + per < ns ? b0 + per : ns;
    const int64_t s0 = b0 + v % upb * ulen;
    const int64_t s1 = s0 + ulen < b1 ? s0 + ulen : b1;
    own[v] = 0;
    if (s0 >= s1) continue;
    const int64_t e0 = sptr[s0], e1 = sptr[s1];
    int32_t lo = INT32_MAX, hi = -1;
    for (int64_t e = e0; e < e1; ++e) {
      TC_CHECK(col[e] >= 0 && col[e] < n, TC_ERR_ARG, "tc_sell_stage: column %d out of range",
               col[e]);
      lo = col[e] < lo ? col[e] : lo;
      hi = col[e] > hi ? col[e] : hi;
    }
    if (e1 - e0 > TC_STAGE_MAX_COLS && hi - lo + 1 > 16 * (int64_t)TC_STAGE_MAX_COLS) {
      // a window this wide (a scattered node order) is all but certain to
      // exceed the limit: count exactly only when it could fit
      std::vector<int32_t> c(col + e0, col + e1);
      std::sort(c.begin(), c.end());
      const int64_t distinct = std::unique(c.begin(), c.end()) - c.begin();
      if (distinct > TC_STAGE_MAX_COLS) {
        *max_cols = (int32_t)distinct;
        set_error("tc_sell_stage: a block references %lld columns (at most %d)",
                  (long long)distinct, TC_STAGE_MAX_COLS);
        return TC_ERR_SIZE;
      }
    }
    pos.assign((size_t)(hi - lo + 1), -1);
    for (int64_t e = e0; e < e1; ++e) pos[col[e] - lo] = 0;
    int32_t cnt = 0;
    for (int32_t c = lo; c <= hi; ++c)
      if (pos[c - lo] == 0) {
        pos[c - lo] = cnt++;
        cols[total + cnt - 1] = c;
      }
    const int32_t r0 = (int32_t)(s0 * 64);
    TC_CHECK(r0 >= lo && r0 <= hi && pos[r0 - lo] >= 0, TC_ERR_ARG,
             "tc_sell_stage: unit %lld does not reference its first row", (long long)v);
    own[v] = pos[r0 - lo];
    if (cnt > mx) mx = cnt;
    if (cnt > TC_STAGE_MAX_COLS) break;  // refused below
    if (cnt <= 65536)
      for (int64_t e = e0; e < e1; ++e) lidx[e] = (uint16_t)pos[col[e] - lo];
    total += cnt;
  }
  col_ptr[(int64_t)G * upb] = (int32_t)total;
  *max_cols = mx;
  TC_CHECK(mx <= TC_STAGE_MAX_COLS, TC_ERR_SIZE,
           "tc_sell_stage: a block references %d columns (at most %d)", mx, TC_STAGE_MAX_COLS);
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
  const PcgBufs B = plan_bufs(P);
#define TC_PCG_SOLVE(T, SC)                                                                  \
  pcg_solve<T>(B, SC, (const T*)b, (T*)x, (const T*)x_prev, (T*)x_next, a_tol, r_tol, max_iter, \
               linear_guess, n_iter_host, s)
  if (P->dtype == TC_F32)
    return P->A.dcol ? TC_PCG_SOLVE(float, P->A.dcol) : TC_PCG_SOLVE(float, P->A.col);
  return P->A.dcol ? TC_PCG_SOLVE(double, P->A.dcol) : TC_PCG_SOLVE(double, P->A.col);
#undef TC_PCG_SOLVE
}

// ------------------------------------------------------------------------
// The same solve with no device allocation by the library or the caller:
// the workspace a plan would hold (control block, partial sums, r, p, p',
// Ap) lives in `scratch`, a file-scope device array of fixed size.  Every
// GPU the library runs on holds its own copy of it for the life of the
// process.  Host code never passes `scratch` itself (its host-side name is
// not a device address): the first call on a device resolves the array's
// device address with hipGetSymbolAddress, with that device current, and
// every later call on that device uses that address.
namespace tc {
namespace {

constexpr int64_t kStaticRows = TC_PCG_STATIC_MAX_ROWS;
constexpr size_t kStaticCtlBytes = 8192;
constexpr size_t kStaticPartBytes = 2 * kMaxBlocks * sizeof(double);   // 2G doubles, G <= 1024
constexpr size_t kStaticVecBytes = size_t(kStaticRows) * sizeof(double);
constexpr size_t kStaticBytes = kStaticCtlBytes + 3 * kStaticPartBytes + 4 * kStaticVecBytes;
static_assert(sizeof(PcgCtl) <= kStaticCtlBytes, "control block does not fit");
static_assert(kStaticPartBytes % 256 == 0 && kStaticVecBytes % 256 == 0, "alignment");

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kStaticMaxDevices = 64;
std::mutex g_static_mu;
void* g_static_ptr[kStaticMaxDevices];     // device address of `scratch` per device
int g_static_chunk[kStaticMaxDevices];     // launch-chunk hint per device
PcgCtl g_static_host[kStaticMaxDevices];   // where each device's control block is read back

int static_scratch(int* dev_out, void** ptr_out) {
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  TC_CHECK(dev >= 0 && dev < kStaticMaxDevices, TC_ERR_ARG,
           "tc_pcg_solve_static: device %d beyond the %d devices it serves", dev,
           kStaticMaxDevices);
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void* ptr = nullptr;
    TC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    TC_CHECK(ptr != nullptr, TC_ERR_HIP, "tc_pcg_solve_static: no device address for scratch");
    g_static_ptr[dev] = ptr;
    g_static_chunk[dev] = 1;
  }
  *dev_out = dev;
  *ptr_out = g_static_ptr[dev];
  return TC_OK;
}

}  // namespace
}  // namespace tc

extern "C" int tc_pcg_solve_static(const tc_sell* A, const void* Av, const void* Dinv, int dtype,
                                   const void* b, void* x, const void* x_prev, void* x_next,
                                   double a_tol, double r_tol, int max_iter, int linear_guess,
                                   int* n_iter_host, void* stream) {
  TC_CHECK(A && Av && Dinv && b && x && n_iter_host && max_iter >= 0 && A->n > 0, TC_ERR_ARG,
           "tc_pcg_solve_static: bad arguments");
  TC_CHECK(dtype == TC_F32 || dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", dtype);
  TC_CHECK(!linear_guess || (x_prev && x_next), TC_ERR_ARG,
           "tc_pcg_solve_static: linear_guess needs x_prev and x_next");
  TC_CHECK(A->n <= kStaticRows, TC_ERR_SIZE,
           "tc_pcg_solve_static: %lld rows exceed the static scratch maximum of %lld rows "
           "(use tc_pcg_create/tc_pcg_solve with a caller workspace)",
           (long long)A->n, (long long)kStaticRows);
  int dev = 0;
  void* base = nullptr;
  int rc = static_scratch(&dev, &base);
  if (rc) return rc;
  char* c = (char*)base;
  PcgBufs B;
  B.A = A;
  B.Av = Av;
  B.Dinv = Dinv;
  B.n = A->n;
  B.G = pcg_blocks(A->n);
  B.ctl = (PcgCtl*)c;
  size_t off = kStaticCtlBytes;
  for (int k = 0; k < 3; ++k, off += kStaticPartBytes) B.part[k] = (double*)(c + off);
  const size_t vb = vec_bytes(A->n, dtype);
  B.r = c + off;
  B.p[0] = c + off + vb;
  B.p[1] = c + off + 2 * vb;
  B.Ap = c + off + 3 * vb;
  B.host_ctl = &g_static_host[dev];
  B.chunk = &g_static_chunk[dev];
  B.timed = nullptr;
  hipStream_t s = (hipStream_t)stream;
#define TC_PCG_SOLVE(T, SC)                                                                  \
  pcg_solve<T>(B, SC, (const T*)b, (T*)x, (const T*)x_prev, (T*)x_next, a_tol, r_tol, max_iter, \
               linear_guess, n_iter_host, s)
  if (dtype == TC_F32) return A->dcol ? TC_PCG_SOLVE(float, A->dcol) : TC_PCG_SOLVE(float, A->col);
  return A->dcol ? TC_PCG_SOLVE(double, A->dcol) : TC_PCG_SOLVE(double, A->col);
#undef TC_PCG_SOLVE
}

// Kernels of the right-hand side and the Jacobi-PCG on SELL-64 operators,
// shared by the stand-alone solver (pcg.hip) and the time-step runtime
// (sim.hip).
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
#pragma once
#include <cstddef>
#include <vector>

#include "tc_internal.h"

namespace tc {

constexpr int kDoneGroups = 32;

struct PcgCtl {
  double rz[2];   // r.z of iteration k in rz[k & 1]
  double bnorm;   // ||Dinv b||
  double a_tol, r_tol;
  int converged;  // stop rule met
  int n_iter;     // iterations that did not stop (reference convention)
  int iter;       // iterations executed
  int stall;      // set when a pipelined step ran out of launched iterations;
                  // every later gated kernel then returns at once
  // Completion counts of a checking update pass: block b counts itself in
  // group[b % kDoneGroups] (each on a cache line of its own, so no more than
  // G/32 atomic adds queue on one address: adds on one address serialize at
  // about 10 ns each); the last block of a group resets that count and
  // counts the group in `done`.  The counts follow every field the host
  // reads back (kCtlHostBytes).
  unsigned done;
  struct alignas(128) Count {
    unsigned n;
  } group[kDoneGroups];
};

// Step status in pinned host memory: two 8-byte words, each a 32-bit
// payload beside the low 32 bits of the record's sequence number.  A
// naturally aligned 8-byte store reaches host memory whole and an aligned
// 8-byte host load reads it whole, so a word that carries the awaited
// sequence number carries that record's payload: no fence is needed between
// payload and flag (a system-scope fence would also write back the L2 of
// the writer's XCD, microseconds after a pass that wrote megabytes).
// The leading fields of PcgCtl, which the host reads back.
constexpr size_t kCtlHostBytes = offsetof(PcgCtl, done);

struct StatusRec {
  unsigned long long w[2];
};

// Payload of the time-step runtimes' records: p0 = n_iter, p1 = iter with
// bit 30 = converged and bit 31 = stalled (iteration counts stay below 2^30).
constexpr uint32_t kStConverged = 1u << 30, kStStalled = 1u << 31, kStIterMask = kStConverged - 1;
constexpr int kMaxIterLimit = (int)kStIterMask;

__device__ inline void put_status(StatusRec* r, uint32_t p0, uint32_t p1, long long seq) {
  const unsigned long long q = (unsigned long long)(uint32_t)seq << 32;
  __hip_atomic_store(&r->w[0], q | p0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&r->w[1], q | p1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The status of record `seq` as the host reads it.
struct StepStatus {
  int n_iter, iter, converged, stalled;
};

inline void status_clear(StatusRec* r, int count) {
  for (int k = 0; k < count; ++k) r[k].w[0] = r[k].w[1] = ~0ull;
}

// True (and *out filled) once both words of `r` carry sequence number seq.
inline bool status_take(const StatusRec* r, long long seq, StepStatus* out) {
  const volatile unsigned long long* w = r->w;
  const unsigned long long a = w[0], b = w[1];
  const uint32_t q = (uint32_t)seq;
  if ((uint32_t)(a >> 32) != q || (uint32_t)(b >> 32) != q) return false;
  const uint32_t p1 = (uint32_t)b;
  out->n_iter = (int)(uint32_t)a;
  out->iter = (int)(p1 & kStIterMask);
  out->converged = (p1 & kStConverged) ? 1 : 0;
  out->stalled = (p1 & kStStalled) ? 1 : 0;
  return true;
}

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
  tc_stage stage;        // staged gathers (stage.cols NULL: none)
  int version;           // bumped when the blocks or the stage change
};

namespace tc {
namespace {

constexpr int kBlock = 256;       // stand-alone SpMV / RHS kernels
constexpr int kPcgBlock = 512;    // PCG passes: 8 waves per block
constexpr int kWaves = kPcgBlock / 64;
// At most 4 blocks per CU (32 waves, full occupancy), and few enough partial
// sums (<= 1024) that every block folding all of them stays cheap.
constexpr int kMaxBlocks = 1024;
constexpr int kXcd = 8;

// Column of a SELL entry in row `row`: int32 absolute columns, or int16
// offsets col - row when every offset of the operator fits in 16 bits (half
// the index traffic of the hot passes).
__device__ inline int32_t col_of(int32_t c, int64_t) { return c; }
__device__ inline int32_t col_of(int16_t c, int64_t row) { return (int32_t)row + (int32_t)c; }

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

// ------------------------------------------------- row chunks (PCG passes)
// The hot passes walk a row in chunks of kChunk entries: all column and
// value loads of a chunk are issued before its gathers, and all gathers
// before the (in-order) float64 accumulation, so that each wave keeps many
// loads in flight.  Entries past the row's padded width repeat the last
// entry (a valid address) and are masked out of the sums.
constexpr int kChunk = 4;

// Column source of one slice: the lane's own entries 64 apart (per-row
// columns), or the slice's shared pattern, one offset per entry position
// and the same for every lane (a wave-uniform address, read through the
// scalar cache).  A pattern position that a row lacks holds value 0 and may
// point past the operator's rows at its edge, so its column is clamped to
// [lo, hi], rows whose values are finite.
template <class CI>
struct Cols {
  const CI* p;
  const int32_t* pat;  // the slice's pattern, or NULL for per-row columns
  int32_t lo, hi;
};

template <class CI>
__device__ inline Cols<CI> per_row(const CI* p) {
  return Cols<CI>{p, nullptr, 0, 0};
}

// The slice table of the hot passes: slice_tag when the operator has
// shared slices, else slice_ptr.
__host__ __device__ inline const int32_t* hot_slices(const tc_sell& A) {
  return A.slice_tag ? A.slice_tag : A.slice_ptr;
}

// Slice s of the hot passes' slice table: its entry offset b, width wd and
// whether it is shared.  The table is slice_ptr, or slice_ptr with bit 0 set
// on the shared slices (offsets are multiples of 64), so the flag costs no
// load of its own.
struct SliceHd {
  int32_t b, wd;
  int sh;
};

__device__ inline SliceHd slice_hd(const int32_t* __restrict__ sptr, int64_t s) {
  const int32_t t0 = sptr[s], t1 = sptr[s + 1];
  const int32_t b = t0 & ~1;
  return SliceHd{b, ((t1 & ~1) - b) >> 6, t0 & 1};
}

// Column source of slice s (entries from b): its shared pattern when it has
// one (int16 offsets only), else the lane's per-row columns.
__device__ inline Cols<int32_t> slice_cols(const int32_t* scol, const int32_t*, int, int32_t b,
                                           int lane, int64_t) {
  return per_row(scol + b + lane);
}

__device__ inline Cols<int16_t> slice_cols(const int16_t* dcol, const int32_t* pat, int sh,
                                           int32_t b, int lane, int64_t n) {
  return Cols<int16_t>{dcol + b + lane, sh ? pat + (b >> 6) : nullptr, 0, (int32_t)(n - 1)};
}

template <class CI, class T>
__device__ inline void load_chunk(const Cols<CI>& c, const T* v, int j0, int wd, int64_t i,
                                  int32_t (&cj)[kChunk], T (&vj)[kChunk]) {
  if (c.pat) {
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const int jj = j0 + u < wd ? j0 + u : wd - 1;
      const int32_t x = (int32_t)i + c.pat[jj];
      cj[u] = x < c.lo ? c.lo : (x > c.hi ? c.hi : x);
      vj[u] = v[jj * 64];
    }
  } else {
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const int jj = j0 + u < wd ? j0 + u : wd - 1;
      cj[u] = col_of(c.p[jj * 64], i);
      vj[u] = v[jj * 64];
    }
  }
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

// This block's position in the block order (the index of its slice range).
__device__ inline int block_vindex() {
  const int G = gridDim.x, b = blockIdx.x;
  return G % kXcd == 0 ? (b % kXcd) * (G / kXcd) + b / kXcd : b;
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

// A partial sum read at device scope (past the reading XCD's L2), for sums
// that other blocks of the running kernel stored at device scope.
__device__ inline double load_dev(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the previous pass's G partials, in a fixed order, returned to
// every thread of the block.
// Blocks of kPcgBlock threads and G <= kMaxBlocks: each thread's loads are
// all issued before its (in-order) additions.  DEV reads them at device
// scope (partials stored by other blocks of the same kernel).
template <int NV, bool DEV = false>
__device__ inline void fold_partials(const double* __restrict__ partials, int G,
                                     double (&out)[NV]) {
  __shared__ double bc[NV];
  constexpr int R = (kMaxBlocks + kPcgBlock - 1) / kPcgBlock;
  double v[NV][R];
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      // unconditional loads (a clamped index, then a select) keep the
      // loads of one thread in flight together
      const int b = (int)threadIdx.x + q * kPcgBlock;
      const double* src = &partials[k * G + (b < G ? b : G - 1)];
      const double t = DEV ? load_dev(src) : *src;
      v[k][q] = b < G ? t : 0.0;
    }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) s += v[k][q];
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

// Sums of two products, written with one explicit fused multiply-add so
// that every pass, and every template instance of a pass, rounds them the
// same way (left to contraction, the compiler fuses one product or the
// other depending on register allocation):
// p' = beta p + Dinv r
__device__ inline double pdir(double beta, double p, double d, double r) {
  return __builtin_fma(beta, p, d * r);
}
// Cm (M x0) + theta dt (K x0)
__device__ inline double ax0(double Cm, double mx, double tdt, double kx) {
  return __builtin_fma(Cm, mx, tdt * kx);
}

__device__ inline bool stop_rule(double zn, const PcgCtl* ctl) {
  return zn < ctl->a_tol || zn / ctl->bnorm < ctl->r_tol;
}

// PCG start: x0 = 2x - x_prev (linear guess) or x; r = b - A x0;
// partial sums of r.z and (Dinv b)^2.
template <class T, bool LG, class CI>
__global__ void __launch_bounds__(kPcgBlock)
k_pcg_init(int64_t n, const int32_t* __restrict__ sptr, const CI* __restrict__ scol,
           const int32_t* __restrict__ pat,
           const T* __restrict__ Av, const T* __restrict__ Dinv, const T* __restrict__ bvec,
           const T* __restrict__ x, const T* __restrict__ xp, T* __restrict__ xnext,
           T* __restrict__ r, PcgCtl* ctl, double a_tol, double r_tol, double* part0) {
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int lane = threadIdx.x & 63, wv = wave_id();
  double red[2] = {0.0, 0.0};
  for (int64_t s = sp.s0 + wv; s < sp.s1; s += kWaves) {
    const int64_t i = s * 64 + lane;
    const SliceHd hd = slice_hd(sptr, s);
    const int32_t b = hd.b, wd = hd.wd;
    const Cols<CI> c = slice_cols(scol, pat, hd.sh, b, lane, n);
    const T* a = Av + b + lane;
    double ax = 0.0;
    for (int j0 = 0; j0 < wd; j0 += kChunk) {
      int32_t cj[kChunk];
      T av[kChunk], xj[kChunk];
      load_chunk(c, a, j0, wd, i, cj, av);
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        if (LG) xj[u] = (T)(2.0 * (double)x[cj[u]] - (double)xp[cj[u]]);
        else xj[u] = x[cj[u]];
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const double t = ax + (double)av[u] * (double)xj[u];
        ax = j0 + u < wd ? t : ax;
      }
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

// One chunk of a row of the spmv pass: columns and A values, then the
// gathered Dinv, r (and the previous p) of those columns; add() forms
// p' = beta p + Dinv r for them and accumulates A p' in order.
template <class T, class CI>
struct SpmvChunk {
  int32_t cj[kChunk];
  T av[kChunk], dj[kChunk], rj[kChunk], qj[kChunk];
  __device__ inline void fetch(const Cols<CI>& c, const T* a, int j0, int wd, int64_t i,
                               const T* __restrict__ Dinv, const T* __restrict__ r,
                               const T* __restrict__ pin, int k) {
    load_chunk(c, a, j0, wd, i, cj, av);
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      dj[u] = Dinv[cj[u]];
      rj[u] = r[cj[u]];
      qj[u] = k == 0 ? T(0) : pin[cj[u]];
    }
  }
  __device__ inline double add(double acc, int j0, int wd, int k, double beta) const {
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      T pj;
      if (k == 0) pj = (T)((double)dj[u] * (double)rj[u]);
      else pj = (T)pdir(beta, (double)qj[u], (double)dj[u], (double)rj[u]);
      const double t = acc + (double)av[u] * (double)pj;
      acc = j0 + u < wd ? t : acc;
    }
    return acc;
  }
};

// spmv pass of iteration k.  pin is the previous p (unused when k == 0),
// pout receives p'.  Iteration k-1's stopping test happens here first.  The
// first chunk of each wave's first slice is fetched before the block folds
// the partial sums, so those latencies overlap.  Each wave reads the flags
// beside its slice header and skips that fetch (two dependent loads: the
// columns, then the gathers) when the pass is halted, so a pass launched
// for a step that already stopped costs little more than the launch.  A
// wave and the block can only disagree on the flags if `converged` is set
// during this pass, which block 0 does only when the stopping test holds,
// and then every block returns at that test before using a fetched chunk.
template <class T, class CI>
__global__ void __launch_bounds__(kPcgBlock)
k_pcg_spmv(int64_t n, const int32_t* __restrict__ sptr, const CI* __restrict__ scol,
           const int32_t* __restrict__ pat,
           const T* __restrict__ Av, const T* __restrict__ Dinv, const T* __restrict__ r,
           const T* __restrict__ pin, T* __restrict__ pout, T* __restrict__ Ap, int k,
           PcgCtl* ctl, const double* __restrict__ part0, const double* __restrict__ part2,
           double* part1) {
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int lane = threadIdx.x & 63, wv = wave_id();
  int64_t s = sp.s0 + wv;
  SpmvChunk<T, CI> ch;
  int32_t b = 0, wd = 0;
  int sh = 0;
  const int wave_halted = ctl->converged | ctl->stall;
  bool fetched = false;
  if (s < sp.s1) {
    const SliceHd hd = slice_hd(sptr, s);
    b = hd.b;
    wd = hd.wd;
    sh = hd.sh;
    if (wd > 0 && !wave_halted) {
      ch.fetch(slice_cols(scol, pat, sh, b, lane, n), Av + b + lane, 0, wd, s * 64 + lane, Dinv, r,
               pin, k);
      fetched = true;
    }
  }
  // the flags and the partial sums are read before the first wait
  __shared__ int halted;
  if (threadIdx.x == 0) halted = ctl->converged | ctl->stall;
  const int G = gridDim.x;
  const double rz_prev = ctl->rz[(k - 1) & 1];
  double v[2];
  fold_partials<2>(k == 0 ? part0 : part2, G, v);
  if (halted) return;
  double beta = 0.0, rz;
  if (k == 0) {
    rz = v[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->bnorm = sqrt(v[1]);
  } else {
    if (stop_rule(sqrt(v[0]), ctl)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->converged = 1;
        ctl->n_iter = k - 1;
        ctl->iter = k;
      }
      return;
    }
    rz = v[1];
    beta = rz / rz_prev;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->rz[k & 1] = rz;
    ctl->n_iter = k;
    ctl->iter = k;
  }
  double red[1] = {0.0};
  bool first = true;
  for (; s < sp.s1; s += kWaves) {
    const int64_t i = s * 64 + lane;
    if (!first) {
      const SliceHd hd = slice_hd(sptr, s);
      b = hd.b;
      wd = hd.wd;
      sh = hd.sh;
    }
    const Cols<CI> c = slice_cols(scol, pat, sh, b, lane, n);
    const T* a = Av + b + lane;
    double acc = 0.0;
    for (int j0 = 0; j0 < wd; j0 += kChunk) {
      if (!(first && j0 == 0 && fetched)) ch.fetch(c, a, j0, wd, i, Dinv, r, pin, k);
      acc = ch.add(acc, j0, wd, k, beta);
    }
    first = false;
    if (i < n) {
      T pi;
      if (k == 0) pi = (T)((double)Dinv[i] * (double)r[i]);
      else pi = (T)pdir(beta, (double)pin[i], (double)Dinv[i], (double)r[i]);
      const T api = (T)acc;
      pout[i] = pi;
      Ap[i] = api;
      red[0] += (double)pi * (double)api;
    }
  }
  store_partials<1>(red, part1);
}

// spmv pass of iteration k with staged gathers (a tc_stage; operators
// without shared slices).  The block forms p' = beta p + Dinv r once for
// every distinct column its rows reference (its set, ascending, at most
// kStageR * kPcgBlock) into LDS, then walks its rows as k_pcg_spmv does,
// reading each entry's p' from LDS at the entry's 16-bit position: the same
// products in the same order, so the same bits, with one load of Dinv, r
// and p per distinct column instead of one per entry.  The set's columns
// and the flags load first, then the partial-sum fold and the set's Dinv,
// r, p gathers are in flight together.
constexpr int kStageR = (TC_STAGE_MAX_COLS + kPcgBlock - 1) / kPcgBlock;

template <class T, bool UNITS>
__global__ void __launch_bounds__(kPcgBlock)
k_pcg_spmv_staged(int64_t n, const int32_t* __restrict__ sptr, const uint16_t* __restrict__ lidx,
                  const int32_t* __restrict__ st_ptr, const int32_t* __restrict__ st_own,
                  const int32_t* __restrict__ st_col, int chunk, const T* __restrict__ Av,
                  const T* __restrict__ Dinv, const T* __restrict__ r,
                  const T* __restrict__ pin, T* __restrict__ pout, T* __restrict__ Ap, int k,
                  PcgCtl* ctl, const double* __restrict__ part0, const double* __restrict__ part2,
                  double* part1) {
  extern __shared__ __align__(16) unsigned char stage_lds[];
  T* ps = reinterpret_cast<T*>(stage_lds);
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int vb = block_vindex();
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int G = gridDim.x;
  // the block's units (runs of `chunk` slices, or its whole range), staged
  // one after the other
  // (UNITS false: one unit per block, the chunk argument unused)
  const int64_t per = (ns + G - 1) / G;
  const int upb = UNITS ? (int)((per + chunk - 1) / chunk) : 1;
  const int64_t ulen = UNITS ? chunk : per;
  const int halted_early = ctl->converged | ctl->stall;
  T dv[kStageR], rv[kStageR], qv[kStageR];
  {
    // the set of unit 0: Dinv, r and the previous p of its columns
    // (registers), in flight during the fold
    const int64_t ui = (int64_t)vb * upb;
    const int32_t u0 = st_ptr[ui], nu = st_ptr[ui + 1] - u0;
    if (!halted_early && nu > 0) {
      int32_t cv[kStageR];
#pragma unroll
      for (int q = 0; q < kStageR; ++q) {
        const int t = (int)threadIdx.x + q * kPcgBlock;
        cv[q] = st_col[u0 + (t < nu ? t : nu - 1)];
      }
#pragma unroll
      for (int q = 0; q < kStageR; ++q) {
        dv[q] = Dinv[cv[q]];
        rv[q] = r[cv[q]];
        qv[q] = k == 0 ? T(0) : pin[cv[q]];
      }
    }
  }
  __shared__ int halted;
  if (threadIdx.x == 0) halted = ctl->converged | ctl->stall;
  const double rz_prev = ctl->rz[(k - 1) & 1];
  double v[2];
  fold_partials<2>(k == 0 ? part0 : part2, G, v);
  if (halted) return;
  double beta = 0.0, rz;
  if (k == 0) {
    rz = v[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->bnorm = sqrt(v[1]);
  } else {
    if (stop_rule(sqrt(v[0]), ctl)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->converged = 1;
        ctl->n_iter = k - 1;
        ctl->iter = k;
      }
      return;
    }
    rz = v[1];
    beta = rz / rz_prev;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->rz[k & 1] = rz;
    ctl->n_iter = k;
    ctl->iter = k;
  }
  double red[1] = {0.0};
  // p' of unit c's set into LDS from its Dinv, r, p (d, rr, q)
  auto form_set = [&](int32_t nu, const T (&d)[kStageR], const T (&rr)[kStageR],
                      const T (&q)[kStageR]) {
#pragma unroll
    for (int t0 = 0; t0 < kStageR; ++t0) {
      const int t = (int)threadIdx.x + t0 * kPcgBlock;
      if (t < nu) {
        if (k == 0) ps[t] = (T)((double)d[t0] * (double)rr[t0]);
        else ps[t] = (T)pdir(beta, (double)q[t0], (double)d[t0], (double)rr[t0]);
      }
    }
  };
  // the rows of unit c: wave w takes its slices us0 + w, us0 + w + 8, ...;
  // with units of a multiple of 8 slices each lane's rows come in the order
  // of the unstaged pass (s0 + w, s0 + w + 8, ...)
  auto unit_rows = [&](int c) {
    const int64_t ui = (int64_t)vb * upb + c;
    const int64_t us0 = sp.s0 + c * ulen;
    const int64_t us1 = us0 + ulen < sp.s1 ? us0 + ulen : sp.s1;
    const int64_t r0 = us0 * 64;
    const int32_t own = st_own[ui];
    for (int64_t s = us0 + wv; s < us1; s += kWaves) {
      const int64_t i = s * 64 + lane;
      const int32_t b = sptr[s], wd = (sptr[s + 1] - b) >> 6;
      const uint16_t* li = lidx + b + lane;
      const T* a = Av + b + lane;
      double acc = 0.0;
      for (int j0 = 0; j0 < wd; j0 += kChunk) {
        uint16_t lj[kChunk];
        T aj[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          const int jj = j0 + u < wd ? j0 + u : wd - 1;
          lj[u] = li[jj * 64];
          aj[u] = a[jj * 64];
        }
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          const double t = acc + (double)aj[u] * (double)ps[lj[u]];
          acc = j0 + u < wd ? t : acc;
        }
      }
      if (i < n) {
        const T pi = ps[own + (int32_t)(i - r0)];
        const T api = (T)acc;
        pout[i] = pi;
        Ap[i] = api;
        red[0] += (double)pi * (double)api;
      }
    }
  };
  // unit 0 from the set loaded before the fold, then the later units (each
  // loading its own set once the previous unit's rows are done)
  form_set(st_ptr[(int64_t)vb * upb + 1] - st_ptr[(int64_t)vb * upb], dv, rv, qv);
  __syncthreads();
  unit_rows(0);
  for (int c = 1; c < upb; ++c) {
    __syncthreads();
    T d2[kStageR], r2[kStageR], q2[kStageR];
    const int64_t ui = (int64_t)vb * upb + c;
    const int32_t u0 = st_ptr[ui], nu = st_ptr[ui + 1] - u0;
    if (nu > 0) {
      int32_t cv[kStageR];
#pragma unroll
      for (int q = 0; q < kStageR; ++q) {
        const int t = (int)threadIdx.x + q * kPcgBlock;
        cv[q] = st_col[u0 + (t < nu ? t : nu - 1)];
      }
#pragma unroll
      for (int q = 0; q < kStageR; ++q) {
        d2[q] = Dinv[cv[q]];
        r2[q] = r[cv[q]];
        q2[q] = k == 0 ? T(0) : pin[cv[q]];
      }
    }
    form_set(nu, d2, r2, q2);
    __syncthreads();
    unit_rows(c);
  }
  store_partials<1>(red, part1);
}

// Stopping test of the last update of a launch chunk (iteration k; k < 0
// when no iteration ran), so the host can read the state.  With `status`
// the result is also stored there; with `stall_open` an unconverged result
// below max_iter sets ctl->stall, which gates every later launch of a
// pipelined run until the host has launched more iterations.  Run by one
// block: the one-block check pass, or the last block of a checking update
// pass (DEV: the partial sums come from the running kernel's other blocks).
template <bool DEV>
__device__ inline void check_block(int k, int G, PcgCtl* ctl, const double* __restrict__ part2,
                                   StatusRec* status, int stall_open, int max_iter,
                                   long long seq) {
  __shared__ int flags[2];
  if (threadIdx.x == 0) {
    flags[0] = ctl->stall;
    flags[1] = ctl->converged;
  }
  // the fold reads the partials while the flags arrive (its result is used
  // only when the solve is still running)
  double v[2] = {0.0, 0.0};
  if (k >= 0) fold_partials<2, DEV>(part2, G, v);
  __syncthreads();
  if (flags[0]) return;
  int n_iter, iter, conv, stalled = 0;
  if (flags[1] || k < 0) {
    n_iter = k < 0 ? 0 : ctl->n_iter;
    iter = k < 0 ? 0 : ctl->iter;
    conv = k < 0 ? 0 : 1;
  } else {
    const bool stop = stop_rule(sqrt(v[0]), ctl);
    conv = stop ? 1 : 0;
    n_iter = stop ? k : k + 1;
    iter = k + 1;
    stalled = (!stop && stall_open && k + 1 < max_iter) ? 1 : 0;
    if (threadIdx.x == 0) {
      ctl->converged = conv;
      ctl->n_iter = n_iter;
      ctl->iter = iter;
      if (stalled) ctl->stall = 1;
    }
  }
  if (threadIdx.x == 0 && status)
    put_status(status, (uint32_t)n_iter,
               (uint32_t)iter | (conv ? kStConverged : 0u) | (stalled ? kStStalled : 0u), seq);
}

// XT is the storage type of the iterate x (float64 in float32 runs of the
// time-step runtime, as the reference keeps its CG iterate in float64).
// CHECK: this is the last update of a launch chunk and also runs the
// stopping test (check_block) in the block that finishes last, so no
// separate check pass is launched.  Every block stores its partial sums at
// device scope, waits until the stores are complete and then counts itself
// in its group's count, the last of a group counts the group in ctl->done;
// the block that counts the last group has every partial sum in memory,
// folds them exactly as the check pass does (same order, same bits) and
// resets the counter.  No block waits for another.
template <class T, class XT = T, bool CHECK = false>
__global__ void __launch_bounds__(kPcgBlock)
k_pcg_update(int64_t n, const T* __restrict__ Dinv, const T* __restrict__ p,
             const T* __restrict__ Ap, XT* __restrict__ x, T* __restrict__ r, int k, PcgCtl* ctl,
             const double* __restrict__ part1, double* part2, StatusRec* status = nullptr,
             int stall_open = 0, int max_iter = 0, long long seq = 0) {
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int lane = threadIdx.x & 63, wv = wave_id();
  // the operands of the wave's first two slices (s, s + kWaves) load while
  // the block reads the flags and folds the partial sums
  constexpr int kPre = 2;
  int64_t s = sp.s0 + wv;
  XT xv[kPre];
  T pv[kPre], av[kPre], rv[kPre], dv[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int64_t su = s + u * kWaves, iu = su * 64 + lane;
    xv[u] = XT(0);
    pv[u] = av[u] = rv[u] = dv[u] = T(0);
    if (su < sp.s1 && iu < n) {
      xv[u] = x[iu];
      pv[u] = p[iu];
      av[u] = Ap[iu];
      rv[u] = r[iu];
      dv[u] = Dinv[iu];
    }
  }
  // the flags, r.z and the partial sums are all read before the first wait
  __shared__ int halted;
  if (threadIdx.x == 0) halted = ctl->converged | ctl->stall;
  const double rzk = ctl->rz[k & 1];
  double v[1];
  fold_partials<1>(part1, gridDim.x, v);
  if (halted) {
    // The flags are the same in every block: nothing writes them during
    // this pass before the last count.  A pass halted by convergence
    // publishes the status the check would (from the control block), one
    // halted by a stall publishes nothing; no block counts itself.
    if (CHECK && blockIdx.x == 0 && threadIdx.x == 0 && status && !ctl->stall)
      put_status(status, (uint32_t)ctl->n_iter, (uint32_t)ctl->iter | kStConverged, seq);
    return;
  }
  double red[2] = {0.0, 0.0};
  {
    const double alpha = rzk / v[0];
    auto upd = [&](int64_t i, XT xo, T po, T ao, T ro, T dvo) {
      const XT xi = (XT)((double)xo + alpha * (double)po);
      const T ri = (T)((double)ro - alpha * (double)ao);
      x[i] = xi;
      r[i] = ri;
      const double z = (double)dvo * (double)ri;
      red[0] += z * z;
      red[1] += (double)ri * z;
    };
    // the prefetched slices s and s + kWaves, then any further ones: each
    // lane's sums run over its slices in the same order as one at a time
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
      const int64_t su = s + u * kWaves, iu = su * 64 + lane;
      if (su < sp.s1 && iu < n) upd(iu, xv[u], pv[u], av[u], rv[u], dv[u]);
    }
    for (s += kPre * kWaves; s < sp.s1; s += kWaves) {
      const int64_t i = s * 64 + lane;
      if (i < n) upd(i, x[i], p[i], Ap[i], r[i], Dinv[i]);
    }
  }
  if constexpr (!CHECK) {
    store_partials<2>(red, part2);
  } else {
    __shared__ int last;
    block_sum<2>(red);
    if (threadIdx.x == 0) {
      const int G = gridDim.x;
      __hip_atomic_store(&part2[blockIdx.x], red[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&part2[G + blockIdx.x], red[1], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      // both stores complete before the block counts itself
      __builtin_amdgcn_s_waitcnt(0);
      const int g = blockIdx.x % kDoneGroups;
      const unsigned members = (unsigned)((G - 1 - g) / kDoneGroups + 1);
      const unsigned groups = (unsigned)min(G, kDoneGroups);
      int fin = 0;
      if (__hip_atomic_fetch_add(&ctl->group[g].n, 1u, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
        // every block of the group has stored its sums
        __hip_atomic_store(&ctl->group[g].n, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fin = __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) == groups - 1;
      }
      last = fin;
    }
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0)
      __hip_atomic_store(&ctl->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    check_block<true>(k, gridDim.x, ctl, part2, status, stall_open, max_iter, seq);
  }
}

// One block: the stopping test of the last update of a launch chunk
// (check_block) as a pass of its own.
__global__ void __launch_bounds__(kPcgBlock)
k_pcg_check(int k, int G, PcgCtl* ctl, const double* __restrict__ part2, StatusRec* status,
            int stall_open, int max_iter, long long seq) {
  check_block<false>(k, G, ctl, part2, status, stall_open, max_iter, seq);
}

// Fused right-hand side and PCG start of one time step:
//   u  = x rounded to T            (the solution the ionic pass saw)
//   b  = M w - coef K u
//   x0 = 2x - x_prev (linear guess) or x, in float64
//   r  = b - (Cm M + tdt K) x0,   partial sums of r.z and (Dinv b)^2
// One pass over the shared pattern reading only the M and K values: A x0 is
// formed from the same loads (A = Cm M + theta dt K).  With the linear guess
// the pass also writes x_next = x0 and x_prev_next = x rounded to T
// (solver.py:35-38: the reference keeps x_prev in the dtype of u).
// WT is the type of w; with a float64 w (precise RHS) b is not rounded and
// K u uses the float64 iterate itself.
template <class T, class XT, class WT, bool LG, class CI>
struct RhsChunk {
  int32_t cj[kChunk];
  T mv[kChunk], kv[kChunk];
  XT xj[kChunk];
  WT wj[kChunk];
  T pj[kChunk];
  __device__ inline void fetch(const Cols<CI>& c, const T* m, const T* kk, int j0, int wd,
                               int64_t i, const XT* __restrict__ x, const WT* __restrict__ w,
                               const T* __restrict__ xp) {
    load_chunk(c, m, j0, wd, i, cj, mv);
#pragma unroll
    for (int u = 0; u < kChunk; ++u) kv[u] = kk[(j0 + u < wd ? j0 + u : wd - 1) * 64];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      xj[u] = x[cj[u]];
      wj[u] = w[cj[u]];
      if (LG) pj[u] = xp[cj[u]];
    }
  }
};

template <class T, class XT, class WT, bool LG, class CI>
__global__ void __launch_bounds__(kPcgBlock)
k_rhs_init(int64_t n, const int32_t* __restrict__ sptr, const CI* __restrict__ scol,
           const int32_t* __restrict__ pat,
           const T* __restrict__ Mv, const T* __restrict__ Kv, const T* __restrict__ Dinv,
           const WT* __restrict__ w, const XT* __restrict__ x, const T* __restrict__ xp,
           XT* __restrict__ xnext, T* __restrict__ xpnext, T* __restrict__ r, double coef,
           double Cm, double tdt, PcgCtl* ctl, double a_tol, double r_tol, double* part0) {
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int lane = threadIdx.x & 63, wv = wave_id();
  // the first chunk of each wave's first slice loads while the block reads
  // the stall flag
  int64_t s = sp.s0 + wv;
  RhsChunk<T, XT, WT, LG, CI> ch;
  int32_t b = 0, wd = 0;
  int sh = 0;
  if (s < sp.s1) {
    const SliceHd hd = slice_hd(sptr, s);
    b = hd.b;
    wd = hd.wd;
    sh = hd.sh;
    if (wd > 0)
      ch.fetch(slice_cols(scol, pat, sh, b, lane, n), Mv + b + lane, Kv + b + lane, 0, wd,
               s * 64 + lane, x, w, xp);
  }
  {
    __shared__ int st;
    if (threadIdx.x == 0) st = ctl->stall;
    __syncthreads();
    if (st) return;
  }
  double red[2] = {0.0, 0.0};
  bool first = true;
  for (; s < sp.s1; s += kWaves) {
    const int64_t i = s * 64 + lane;
    if (!first) {
      const SliceHd hd = slice_hd(sptr, s);
      b = hd.b;
      wd = hd.wd;
      sh = hd.sh;
    }
    const Cols<CI> c = slice_cols(scol, pat, sh, b, lane, n);
    const T* m = Mv + b + lane;
    const T* kk = Kv + b + lane;
    double am = 0.0, ak = 0.0, mx = 0.0, kx = 0.0;
    for (int j0 = 0; j0 < wd; j0 += kChunk) {
      if (!(first && j0 == 0)) ch.fetch(c, m, kk, j0, wd, i, x, w, xp);
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        if (j0 + u >= wd) break;
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
    if (i < n) {
      if (LG) {
        xnext[i] = (XT)(2.0 * (double)x[i] - (double)xp[i]);
        xpnext[i] = (T)x[i];
      }
      double bi = am - coef * ak;
      if (sizeof(WT) != 8) bi = (double)(T)bi;
      const double d = (double)Dinv[i];
      const T ri = (T)(bi - ax0(Cm, mx, tdt, kx));
      r[i] = ri;
      red[0] = __builtin_fma((double)ri, d * (double)ri, red[0]);
      const double db = d * bi;
      red[1] = __builtin_fma(db, db, red[1]);
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

// Bytes of LDS the staged RHS + PCG start pass needs for a set of nu
// columns: x (XT), w (WT) and, with the linear guess, x_prev (T) per column.
template <class T, class XT, class WT, bool LG>
__host__ __device__ inline size_t rhs_stage_bytes(int nu) {
  const size_t a = ((size_t)nu * sizeof(XT) + 15) & ~(size_t)15;
  const size_t b = ((size_t)nu * sizeof(WT) + 15) & ~(size_t)15;
  return a + b + (LG ? (size_t)nu * sizeof(T) : 0);
}

// Fused RHS + PCG start with staged gathers (a tc_stage; operators without
// shared slices): the block first loads x, w and x_prev once for every
// distinct column its rows reference into LDS, then walks its rows as
// k_rhs_init does, reading each entry's neighbour values from LDS at the
// entry's 16-bit position; the same terms in the same order, so the same
// bits.  The set's values load while the block reads the stall flag.
//
// NT threads per block: kPcgBlock (8 waves, wave w taking slices s0 + w,
// s0 + w + 8, ...) or 2 kPcgBlock (16 waves, one slice each, for blocks of
// at most 16 slices: the set's LDS allows few blocks per CU, and one slice
// per wave halves each block's chain of dependent loads).  The 16-wave form
// adds wave w + 8's per-lane sums to wave w's and folds the first 8 waves,
// which is the 8-wave form's order of additions: the same partial sums.
template <class T, class XT, class WT, bool LG, int NT, bool UNITS>
__global__ void __launch_bounds__(NT)
k_rhs_init_staged(int64_t n, const int32_t* __restrict__ sptr, const uint16_t* __restrict__ lidx,
                  const int32_t* __restrict__ st_ptr, const int32_t* __restrict__ st_own,
                  const int32_t* __restrict__ st_col, int chunk, const T* __restrict__ Mv,
                  const T* __restrict__ Kv, const T* __restrict__ Dinv, const WT* __restrict__ w,
                  const XT* __restrict__ x, const T* __restrict__ xp, XT* __restrict__ xnext,
                  T* __restrict__ xpnext, T* __restrict__ r, double coef, double Cm,
                  double tdt, PcgCtl* ctl, double a_tol, double r_tol, double* part0) {
  extern __shared__ __align__(16) unsigned char stage_lds[];
  const int64_t ns = (n + 63) >> 6;
  const Split sp = block_slices(ns);
  const int vb = block_vindex();
  const int lane = threadIdx.x & 63, wv = wave_id();
  const int G = gridDim.x;
  // the block's units (runs of `chunk` slices, or its whole range), staged
  // one after the other; with 16 waves a unit holds at most 16 slices
  // (UNITS false: one unit per block, the chunk argument unused)
  const int64_t per = (ns + G - 1) / G;
  const int upb = UNITS ? (int)((per + chunk - 1) / chunk) : 1;
  const int64_t ulen = UNITS ? chunk : per;
  __shared__ int st;
  if (threadIdx.x == 0) st = ctl->stall;
  static_assert(NT == kPcgBlock || NT == 2 * kPcgBlock, "8 or 16 waves");
  constexpr int R = (TC_STAGE_MAX_COLS + NT - 1) / NT;
  constexpr int NW = NT / 64;
  double red[2] = {0.0, 0.0};
  for (int c = 0; c < upb; ++c) {
    const int64_t ui = (int64_t)vb * upb + c;
    const int32_t u0 = st_ptr[ui], nu = st_ptr[ui + 1] - u0;
    XT* xs = reinterpret_cast<XT*>(stage_lds);
    WT* ws = reinterpret_cast<WT*>(stage_lds + (((size_t)nu * sizeof(XT) + 15) & ~(size_t)15));
    T* ps = reinterpret_cast<T*>(reinterpret_cast<unsigned char*>(ws) +
                                 (((size_t)nu * sizeof(WT) + 15) & ~(size_t)15));
    if (c > 0) __syncthreads();  // every wave is done with the previous unit's LDS
    if (nu > 0) {
      int32_t cv[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int t = (int)threadIdx.x + q * NT;
        cv[q] = st_col[u0 + (t < nu ? t : nu - 1)];
      }
      XT xv[R];
      WT wv_[R];
      T pv[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        xv[q] = x[cv[q]];
        wv_[q] = w[cv[q]];
        if (LG) pv[q] = xp[cv[q]];
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int t = (int)threadIdx.x + q * NT;
        if (t < nu) {
          xs[t] = xv[q];
          ws[t] = wv_[q];
          if (LG) ps[t] = pv[q];
        }
      }
    }
    __syncthreads();
    if (st) return;
    const int64_t us0 = sp.s0 + c * ulen;
    const int64_t us1 = us0 + ulen < sp.s1 ? us0 + ulen : sp.s1;
    const int64_t r0 = us0 * 64;
    const int32_t own = st_own[ui];
    // the factors of this lane's last row's terms in the unit (0 when it
    // had none)
    double f_ri = 0.0, f_dri = 0.0, f_db = 0.0;
    for (int64_t s = us0 + wv; s < us1; s += NW) {
      const int64_t i = s * 64 + lane;
      const int32_t b = sptr[s], wd = (sptr[s + 1] - b) >> 6;
      const uint16_t* li = lidx + b + lane;
      const T* m = Mv + b + lane;
      const T* kk = Kv + b + lane;
      double am = 0.0, ak = 0.0, mx = 0.0, kx = 0.0;
      for (int j0 = 0; j0 < wd; j0 += kChunk) {
        uint16_t lj[kChunk];
        T mv[kChunk], kv[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          const int jj = j0 + u < wd ? j0 + u : wd - 1;
          lj[u] = li[jj * 64];
          mv[u] = m[jj * 64];
          kv[u] = kk[jj * 64];
        }
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          if (j0 + u >= wd) break;
          const int q = lj[u];
          const XT xq = xs[q];
          double x0;
          if (LG) x0 = (double)(XT)(2.0 * (double)xq - (double)ps[q]);
          else x0 = (double)xq;
          const double mj = (double)mv[u], kj = (double)kv[u];
          am += mj * (double)ws[q];
          if (sizeof(WT) == 8) ak += kj * (double)xq;
          else ak += kj * (double)(T)xq;
          mx += mj * x0;
          kx += kj * x0;
        }
      }
      if (i < n) {
        const int q = own + (int32_t)(i - r0);
        const XT xi = xs[q];
        if (LG) {
          xnext[i] = (XT)(2.0 * (double)xi - (double)ps[q]);
          xpnext[i] = (T)xi;
        }
        double bi = am - coef * ak;
        if (sizeof(WT) != 8) bi = (double)(T)bi;
        const double d = (double)Dinv[i];
        const T ri = (T)(bi - ax0(Cm, mx, tdt, kx));
        r[i] = ri;
        f_ri = (double)ri;
        f_dri = d * (double)ri;
        f_db = d * bi;
        if (NW == kWaves || wv < kWaves) {
          red[0] = __builtin_fma(f_ri, f_dri, red[0]);
          red[1] = __builtin_fma(f_db, f_db, red[1]);
        }
      }
    }
    if constexpr (NT != kPcgBlock) {
      // waves 8..15 (one row per lane at most in a unit) hand that row's
      // factors to waves 0..7 (through the LDS of the set, free once every
      // wave has read it), which fuse them into their sums as the 8-wave
      // block's second slice of the unit does
      double* comb = reinterpret_cast<double*>(stage_lds);
      __syncthreads();
      if (wv >= kWaves) {
        double* cb = comb + ((wv - kWaves) * 64 + lane) * 3;
        cb[0] = f_ri;
        cb[1] = f_dri;
        cb[2] = f_db;
      }
      __syncthreads();
      if (wv < kWaves) {
        const double* cb = comb + (wv * 64 + lane) * 3;
        red[0] = __builtin_fma(cb[0], cb[1], red[0]);
        red[1] = __builtin_fma(cb[2], cb[2], red[1]);
      }
    }
  }
  if constexpr (NT == kPcgBlock) {
    store_partials<2>(red, part0);
  } else {
    // the first 8 waves fold as the 8-wave block does
    __shared__ double wsum[2][kWaves];
    if (wv < kWaves) {
      red[0] = wave_sum(red[0]);
      red[1] = wave_sum(red[1]);
      if (lane == 0) {
        wsum[0][wv] = red[0];
        wsum[1][wv] = red[1];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double s0 = 0.0, s1 = 0.0;
      for (int q = 0; q < kWaves; ++q) {
        s0 += wsum[0][q];
        s1 += wsum[1][q];
      }
      part0[blockIdx.x] = s0;
      part0[gridDim.x + blockIdx.x] = s1;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->converged = 0;
    ctl->n_iter = 0;
    ctl->iter = 0;
    ctl->a_tol = a_tol;
    ctl->r_tol = r_tol;
  }
}

// dst = src (converted to T) unless the gate is set.
template <class T, class XT>
__global__ void k_copy_gated(const int* gate, const XT* __restrict__ src, T* __restrict__ dst,
                             int64_t n) {
  if (gate && *(volatile const int*)gate) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (T)src[i];
}

// LAT/RT test (monodomain.py:174-176) on u = x rounded to T.
template <class T, class XT>
__global__ void k_lat_x(const XT* __restrict__ x, T* __restrict__ AT, T* __restrict__ RT,
                        int64_t n, double t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T V = (T)x[i];
  T at = AT[i];
  if (V > T(0) && at == T(-1)) {
    at = (T)t;
    AT[i] = at;
  }
  if (at > T(0) && V < T(-70) && RT[i] == T(-1)) RT[i] = (T)t;
}

__global__ void k_set_int(int* p, int v) { *p = v; }

inline size_t vec_bytes(int64_t n, int dtype) {
  return align_up((size_t)n * (dtype == TC_F64 ? 8 : 4), 256);
}

inline int pcg_blocks(int64_t n) {
  const int64_t ns = ceil_div(n, 64);
  int64_t g = ceil_div(ns, kWaves);
  if (g > kMaxBlocks) g = kMaxBlocks;
  if (g >= kXcd) g -= g % kXcd;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace
}  // namespace tc

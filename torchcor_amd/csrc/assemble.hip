// P1 operator assembly on the device: CSR pattern by a stable radix sort of
// the element (row, col) pairs, element matrices in float64 from closed-form
// cofactors, deterministic per-entry gather sums (no float atomics), and
// the SELL-64 layout used by the SpMV kernels.
#include <hipcub/hipcub.hpp>

#include "tc_internal.h"

namespace tc {
namespace {

constexpr int kBlock = 256;

int key_bits(int64_t n_nodes) {
  unsigned __int128 m = (unsigned __int128)n_nodes * (unsigned __int128)n_nodes;
  int b = 1;
  while (b < 64 && ((unsigned __int128)1 << b) < m) ++b;
  return b;
}

struct PatternWs {
  uint64_t* key_in;
  uint64_t* key_out;
  int32_t* id_in;
  int32_t* id_out;
  int32_t* slot;   // inclusive scan of head flags -> CSR slot of each sorted entry
  void* temp;
  size_t temp_bytes;
  size_t total;
};

PatternWs pattern_layout(void* base, int64_t N) {
  PatternWs w{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  size_t o_ki = take(sizeof(uint64_t) * N);
  size_t o_ko = take(sizeof(uint64_t) * N);
  size_t o_ii = take(sizeof(int32_t) * N);
  size_t o_io = take(sizeof(int32_t) * N);
  size_t o_sl = take(sizeof(int32_t) * N);
  size_t sort_bytes = 0, scan_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint64_t*)nullptr,
                                           (uint64_t*)nullptr, (int32_t*)nullptr,
                                           (int32_t*)nullptr, (int)N, 0, 64);
  (void)hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr,
                                         (int32_t*)nullptr, (int)N);
  size_t tb = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
  size_t o_t = take(tb);
  w.total = off;
  char* b = (char*)base;
  w.key_in = (uint64_t*)(b + o_ki);
  w.key_out = (uint64_t*)(b + o_ko);
  w.id_in = (int32_t*)(b + o_ii);
  w.id_out = (int32_t*)(b + o_io);
  w.slot = (int32_t*)(b + o_sl);
  w.temp = b + o_t;
  w.temp_bytes = tb;
  return w;
}

// Entry id = (a*k + b)*E + e: the reference's listing order (assemble.py:225-231).
__global__ void k_pattern_keys(const int32_t* __restrict__ elems, int64_t E, int k,
                               int64_t n, int64_t N, uint64_t* __restrict__ key,
                               int32_t* __restrict__ id) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  int64_t ab = t / E, e = t - ab * E;
  int a = (int)(ab / k), b = (int)(ab - (int64_t)a * k);
  uint64_t r = (uint32_t)elems[e * k + a], c = (uint32_t)elems[e * k + b];
  key[t] = r * (uint64_t)n + c;
  id[t] = (int32_t)t;
}

__global__ void k_head_flags(const uint64_t* __restrict__ key, int64_t N,
                             int32_t* __restrict__ flag) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  flag[t] = (t == 0 || key[t] != key[t - 1]) ? 1 : 0;
}

// One thread per sorted entry t: heads open a CSR slot.
__global__ void k_pattern_finish(const uint64_t* __restrict__ key,
                                 const int32_t* __restrict__ slot, int64_t N, int64_t n,
                                 int64_t nnz, int32_t* __restrict__ row_ptr,
                                 int32_t* __restrict__ col_idx,
                                 int32_t* __restrict__ seg_ptr) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  bool head = (t == 0) || key[t] != key[t - 1];
  if (!head) return;
  int64_t s = slot[t] - 1;
  uint64_t k = key[t];
  int64_t r = (int64_t)(k / (uint64_t)n);
  col_idx[s] = (int32_t)(k - (uint64_t)r * (uint64_t)n);
  seg_ptr[s] = (int32_t)t;
  // Rows (prev_r, r] start at slot s.
  int64_t prev_r = -1;
  if (t > 0) {
    // Previous head's row: the previous entry belongs to the previous slot.
    prev_r = (int64_t)(key[t - 1] / (uint64_t)n);
  }
  for (int64_t rr = prev_r + 1; rr <= r; ++rr) row_ptr[rr] = (int32_t)s;
  if (s == nnz - 1) {
    seg_ptr[nnz] = (int32_t)N;
    for (int64_t rr = r + 1; rr <= n; ++rr) row_ptr[rr] = (int32_t)nnz;
  }
}

// ---------------------------------------------------------------- elements

__device__ inline double det3(const double J[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
         J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

__device__ inline double maybe_f32(double v, int round_f32) {
  return round_f32 ? (double)(float)v : v;
}

// One thread per tetrahedron (assemble.py:163-215): J columns are the edge
// vectors X[c+1]-X[0]; shape gradients G = dN J^-1; Ke = |detJ|/6 G s G^T;
// Me = V/20 (1 + delta) with V = |detJ|/6.
__device__ inline void sigma_of(int64_t e, const double* __restrict__ sig,
                                const double* __restrict__ sl, const double* __restrict__ st,
                                const double* __restrict__ fib, double S[3][3]) {
  if (sig) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) S[r][c] = sig[e * 9 + r * 3 + c];
    return;
  }
  const double f[3] = {fib[e * 3 + 0], fib[e * 3 + 1], fib[e * 3 + 2]};
  const double t = st[e], dlt = sl[e] - t;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) S[r][c] = (r == c ? t : 0.0) + dlt * f[r] * f[c];
}

__global__ void k_local_tet(const double* __restrict__ X, const int32_t* __restrict__ elems,
                            int64_t E, const double* __restrict__ sig,
                            const double* __restrict__ sl,
                            const double* __restrict__ st, const double* __restrict__ fib,
                            int round_f32, double* __restrict__ Kloc,
                            double* __restrict__ Mloc) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double P[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    int64_t v = elems[e * 4 + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) P[a][d] = X[v * 3 + d];
  }
  double J[3][3];  // J[r][c] = P[c+1][r] - P[0][r]
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) J[r][c] = P[c + 1][r] - P[0][r];
  const double det = det3(J);
  // inverse via adjugate: inv[r][c] = cof[c][r] / det
  double inv[3][3];
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  double G[4][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    G[1][d] = inv[0][d];
    G[2][d] = inv[1][d];
    G[3][d] = inv[2][d];
    G[0][d] = -(inv[0][d] + inv[1][d] + inv[2][d]);
  }
  double S[3][3];
  sigma_of(e, sig, sl, st, fib, S);
  const double adet = fabs(det);
  const double scale = adet / 6.0;
  const double vol = adet / 6.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double SG[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) SG[r] = S[r][0] * G[a][0] + S[r][1] * G[a][1] + S[r][2] * G[a][2];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double kab = scale * (G[b][0] * SG[0] + G[b][1] * SG[1] + G[b][2] * SG[2]);
      Kloc[e * 16 + a * 4 + b] = maybe_f32(kab, round_f32);
      Mloc[e * 16 + a * 4 + b] = maybe_f32((vol / 20.0) * (a == b ? 2.0 : 1.0), round_f32);
    }
  }
}

// One thread per surface triangle (assemble.py:99-145): in-plane frame
// R = [alpha; beta]; conductivity R s R^T; J = R [X1-X0, X2-X0];
// Ke = 0.5 detJ G s2 G^T; Me = area/12 (1 + delta).
__global__ void k_local_tri(const double* __restrict__ X, const int32_t* __restrict__ elems,
                            int64_t E, const double* __restrict__ sig,
                            const double* __restrict__ sl,
                            const double* __restrict__ st, const double* __restrict__ fib,
                            int round_f32, double* __restrict__ Kloc,
                            double* __restrict__ Mloc) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double P[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int64_t v = elems[e * 3 + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) P[a][d] = X[v * 3 + d];
  }
  double u1[3], u2[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u1[d] = P[1][d] - P[0][d];
    u2[d] = P[2][d] - P[0][d];
  }
  const double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  double al[3] = {u1[0] / n1, u1[1] / n1, u1[2] / n1};
  const double pr = u2[0] * al[0] + u2[1] * al[1] + u2[2] * al[2];
  double be[3] = {u2[0] - pr * al[0], u2[1] - pr * al[1], u2[2] - pr * al[2]};
  const double n2 = sqrt(be[0] * be[0] + be[1] * be[1] + be[2] * be[2]);
  be[0] /= n2; be[1] /= n2; be[2] /= n2;
  const double cx = u1[1] * u2[2] - u1[2] * u2[1];
  const double cy = u1[2] * u2[0] - u1[0] * u2[2];
  const double cz = u1[0] * u2[1] - u1[1] * u2[0];
  const double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
  const double* R[2] = {al, be};
  // 2x2 Jacobian J[r][c] = R[r] . (P[c+1]-P[0])
  double J[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    J[r][0] = R[r][0] * u1[0] + R[r][1] * u1[1] + R[r][2] * u1[2];
    J[r][1] = R[r][0] * u2[0] + R[r][1] * u2[1] + R[r][2] * u2[2];
  }
  const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  double inv[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
  double G[3][2];
  G[1][0] = inv[0][0]; G[1][1] = inv[0][1];
  G[2][0] = inv[1][0]; G[2][1] = inv[1][1];
  G[0][0] = -(inv[0][0] + inv[1][0]);
  G[0][1] = -(inv[0][1] + inv[1][1]);
  double S[3][3];
  sigma_of(e, sig, sl, st, fib, S);
  double S2[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc += R[r][i] * S[i][j] * R[c][j];
      S2[r][c] = acc;
    }
  const double scale = 0.5 * det;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double SG0 = S2[0][0] * G[a][0] + S2[0][1] * G[a][1];
    double SG1 = S2[1][0] * G[a][0] + S2[1][1] * G[a][1];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double kab = scale * (G[b][0] * SG0 + G[b][1] * SG1);
      Kloc[e * 9 + a * 3 + b] = maybe_f32(kab, round_f32);
      Mloc[e * 9 + a * 3 + b] = maybe_f32((area / 12.0) * (a == b ? 2.0 : 1.0), round_f32);
    }
  }
}

// One thread per CSR entry: sum its element contributions in listing order.
__global__ void k_gather_sum(const int32_t* __restrict__ seg_ptr,
                             const int32_t* __restrict__ contrib, int64_t nnz, int64_t E,
                             int kk, const double* __restrict__ Kloc,
                             const double* __restrict__ Mloc, double* __restrict__ Kv,
                             double* __restrict__ Mv) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  double ks = 0.0, ms = 0.0;
  for (int32_t q = seg_ptr[s]; q < seg_ptr[s + 1]; ++q) {
    int64_t id = contrib[q];
    int64_t ab = id / E, e = id - ab * E;
    ks += Kloc[e * kk + ab];
    ms += Mloc[e * kk + ab];
  }
  Kv[s] = ks;
  Mv[s] = ms;
}

template <class T>
__global__ void k_operator_values(const double* __restrict__ M, const double* __restrict__ K,
                                  const int32_t* __restrict__ row_ptr,
                                  const int32_t* __restrict__ col, int64_t n, double Chi,
                                  double Cm, double tdt, T* __restrict__ Mv,
                                  T* __restrict__ Kv, T* __restrict__ Av,
                                  T* __restrict__ Dinv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T diag = T(0);
  for (int32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
    double kq = K[q] / Chi;
    double aq = Cm * M[q] + tdt * kq;
    Mv[q] = (T)M[q];
    Kv[q] = (T)kq;
    Av[q] = (T)aq;
    if (col[q] == i) diag = (T)aq;
  }
  Dinv[i] = T(1) / diag;
}

// ------------------------------------------------------------------- SELL

__global__ void k_slice_width(const int32_t* __restrict__ row_ptr, int64_t n,
                              int64_t n_slices, int32_t* __restrict__ width) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slices) return;
  int w = 0;
  for (int l = 0; l < kWave; ++l) {
    int64_t i = s * kWave + l;
    if (i < n) {
      int len = row_ptr[i + 1] - row_ptr[i];
      w = len > w ? len : w;
    }
  }
  width[s] = w;
}

// Single-block exclusive scan of 64*width into slice_ptr (n_slices+1).
__global__ void k_slice_scan(const int32_t* width, int64_t n_slices, int32_t* slice_ptr) {
  __shared__ int64_t part[1024];
  const int T_ = blockDim.x;
  const int64_t chunk = ceil_div(n_slices, T_);
  const int64_t b = threadIdx.x * chunk;
  const int64_t e = b + chunk < n_slices ? b + chunk : n_slices;
  int64_t s = 0;
  for (int64_t q = b; q < e; ++q) s += (int64_t)kWave * width[q];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int t = 0; t < T_; ++t) {
      int64_t v = part[t];
      part[t] = run;
      run += v;
    }
    slice_ptr[n_slices] = (int32_t)run;
  }
  __syncthreads();
  int64_t run = part[threadIdx.x];
  for (int64_t q = b; q < e; ++q) {
    const int32_t wq = width[q];  // read before the in-place write below
    slice_ptr[q] = (int32_t)run;
    run += (int64_t)kWave * wq;
  }
}

// Entry e of row i (offset d = col - i) goes to position e of its slice,
// or, in a shared slice, to the position of d in the slice's pattern; the
// other positions are padding (col = i, value 0).  pos_walk visits the w
// positions in order and calls put(j, e) for a real entry, pad(j) else.
template <class Put, class Pad>
__device__ inline void pos_walk(const int32_t* __restrict__ col, int32_t b, int len, int64_t i,
                                int w, const int32_t* pat_s, Put put, Pad pad) {
  if (!pat_s) {
    for (int j = 0; j < w; ++j) {
      if (j < len) put(j, j);
      else pad(j);
    }
    return;
  }
  int e = 0;
  for (int j = 0; j < w; ++j) {
    if (e < len && (int64_t)col[b + e] - i == (int64_t)pat_s[j]) put(j, e++);
    else pad(j);
  }
}

// One wavefront per slice: the slice is shared when every row's offsets
// lie, in order, among those of its first longest row and all fit int16.
// The candidate pattern goes through LDS (the rows' subset walks diverge,
// so they read it there rather than across lanes); every thread reaches
// the block barrier, waves past the last slice included.
__global__ void __launch_bounds__(kBlock)
k_sell_share(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col, int64_t n,
             int64_t n_slices, const int32_t* __restrict__ slice_ptr, int32_t* __restrict__ pat,
             int32_t* __restrict__ slice_tag) {
  __shared__ int32_t U[kBlock / kWave][kWave];
  const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const bool live = s < n_slices;
  const int64_t i = s * kWave + lane;
  const int w = live ? (slice_ptr[s + 1] - slice_ptr[s]) / kWave : 0;
  int len = 0;
  int32_t b = 0;
  if (live && i < n) {
    b = row_ptr[i];
    len = row_ptr[i + 1] - b;
  }
  const unsigned long long longest = __ballot(w > 0 && len == w);
  const bool cand = w > 0 && w <= kWave && longest != 0;
  bool ok = true;
  if (cand && lane < w) {
    const int lead = __ffsll((long long)longest) - 1;
    const int64_t li = s * kWave + lead;
    const int64_t d = (int64_t)col[row_ptr[li] + lane] - li;
    ok = d >= -32767 && d <= 32767;
    U[wv][lane] = (int32_t)d;
  }
  __syncthreads();
  if (cand) {
    // this row's offsets, in order, among the pattern's
    int p = 0;
    for (int e = 0; e < len && ok; ++e) {
      const int64_t d = (int64_t)col[b + e] - i;
      while (p < w && (int64_t)U[wv][p] < d) ++p;
      ok = p < w && (int64_t)U[wv][p] == d;
      ++p;
    }
  }
  const bool all = cand && __ballot(!ok) == 0;
  if (live && lane == 0) slice_tag[s] = slice_ptr[s] | (all ? 1 : 0);
  if (live && lane == 0 && s == n_slices - 1) slice_tag[n_slices] = slice_ptr[n_slices];
  if (all && lane < w) pat[slice_ptr[s] / kWave + lane] = U[wv][lane];
}

__global__ void k_sell_columns(const int32_t* __restrict__ row_ptr,
                               const int32_t* __restrict__ col, int64_t n,
                               int64_t n_slices, const int32_t* __restrict__ slice_ptr,
                               int32_t* __restrict__ scol, const int32_t* __restrict__ pat,
                               const int32_t* __restrict__ slice_tag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_slices * kWave) return;
  int64_t s = i / kWave;
  int l = (int)(i - s * kWave);
  int w = (slice_ptr[s + 1] - slice_ptr[s]) / kWave;
  int64_t base = slice_ptr[s] + l;
  int len = 0;
  int32_t b = 0;
  if (i < n) {
    b = row_ptr[i];
    len = row_ptr[i + 1] - b;
  }
  const int32_t pad = (int32_t)(i < n ? i : n - 1);
  const int32_t* ps = (pat && slice_tag && (slice_tag[s] & 1)) ? pat + slice_ptr[s] / kWave : nullptr;
  pos_walk(col, b, len, i, w, ps,
           [&](int j, int e) { scol[base + (int64_t)j * kWave] = col[b + e]; },
           [&](int j) { scol[base + (int64_t)j * kWave] = pad; });
}

__global__ void k_zero_int(int32_t* p) { *p = 0; }

// dcol = scol - row (int16, saturated); *max_abs = the largest |scol - row|.
__global__ void k_sell_deltas(const int32_t* __restrict__ scol, int64_t n_slices,
                              const int32_t* __restrict__ slice_ptr, int16_t* __restrict__ dcol,
                              int32_t* max_abs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t m = 0;
  if (i < n_slices * kWave) {
    const int64_t s = i / kWave;
    const int l = (int)(i - s * kWave);
    const int w = (slice_ptr[s + 1] - slice_ptr[s]) / kWave;
    const int64_t base = slice_ptr[s] + l;
    for (int j = 0; j < w; ++j) {
      const int64_t e = base + (int64_t)j * kWave;
      int64_t d = (int64_t)scol[e] - i;
      const int64_t a = d < 0 ? -d : d;
      if (a > m) m = a > INT32_MAX ? INT32_MAX : (int32_t)a;
      d = d > 32767 ? 32767 : (d < -32767 ? -32767 : d);
      dcol[e] = (int16_t)d;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t v = __shfl_xor(m, o);
    m = v > m ? v : m;
  }
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(max_abs, m);
}

template <class T>
__global__ void k_sell_values(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                              int64_t n, int64_t n_slices, const int32_t* __restrict__ slice_ptr,
                              const T* __restrict__ v, T* __restrict__ sv,
                              const int32_t* __restrict__ pat,
                              const int32_t* __restrict__ slice_tag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_slices * kWave) return;
  int64_t s = i / kWave;
  int l = (int)(i - s * kWave);
  int w = (slice_ptr[s + 1] - slice_ptr[s]) / kWave;
  int64_t base = slice_ptr[s] + l;
  int len = 0;
  int32_t b = 0;
  if (i < n) {
    b = row_ptr[i];
    len = row_ptr[i + 1] - b;
  }
  const int32_t* ps = (pat && slice_tag && (slice_tag[s] & 1)) ? pat + slice_ptr[s] / kWave : nullptr;
  pos_walk(col, b, len, i, w, ps,
           [&](int j, int e) { sv[base + (int64_t)j * kWave] = v[b + e]; },
           [&](int j) { sv[base + (int64_t)j * kWave] = T(0); });
}

}  // namespace
}  // namespace tc

using namespace tc;

extern "C" int tc_pattern_workspace_size(int64_t n_elems, int k, size_t* bytes) {
  TC_CHECK(bytes && n_elems > 0 && (k == 3 || k == 4), TC_ERR_ARG,
           "tc_pattern_workspace_size: bad arguments");
  int64_t N = n_elems * k * k;
  TC_CHECK(N < (int64_t)INT32_MAX, TC_ERR_SIZE, "too many element entries (%lld)", (long long)N);
  *bytes = pattern_layout(nullptr, N).total;
  return TC_OK;
}

extern "C" int tc_pattern_sort(const int32_t* elems, int64_t E, int k, int64_t n, void* ws,
                               size_t ws_bytes, int64_t* nnz_host, void* stream) {
  TC_CHECK(elems && ws && nnz_host && E > 0 && n > 0 && (k == 3 || k == 4), TC_ERR_ARG,
           "tc_pattern_sort: bad arguments");
  TC_CHECK(n < (int64_t)INT32_MAX, TC_ERR_SIZE, "too many nodes");
  int64_t N = E * k * k;
  PatternWs w = pattern_layout(ws, N);
  TC_CHECK(ws_bytes >= w.total, TC_ERR_SIZE, "pattern workspace too small (%zu < %zu)",
           ws_bytes, w.total);
  hipStream_t s = (hipStream_t)stream;
  int grid = (int)ceil_div(N, kBlock);
  hipLaunchKernelGGL(k_pattern_keys, dim3(grid), dim3(kBlock), 0, s, elems, E, k, n, N,
                     w.key_in, w.id_in);
  TC_LAUNCH();
  size_t tb = w.temp_bytes;
  TC_HIP(hipcub::DeviceRadixSort::SortPairs(w.temp, tb, w.key_in, w.key_out, w.id_in,
                                            w.id_out, (int)N, 0, key_bits(n), s));
  hipLaunchKernelGGL(k_head_flags, dim3(grid), dim3(kBlock), 0, s, w.key_out, N, w.id_in);
  TC_LAUNCH();
  tb = w.temp_bytes;
  TC_HIP(hipcub::DeviceScan::InclusiveSum(w.temp, tb, w.id_in, w.slot, (int)N, s));
  int32_t last = 0;
  TC_HIP(hipMemcpyAsync(&last, w.slot + (N - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  TC_HIP(hipStreamSynchronize(s));
  *nnz_host = last;
  return TC_OK;
}

extern "C" int tc_pattern_finish(int64_t E, int k, int64_t n, int64_t nnz, const void* ws,
                                 int32_t* row_ptr, int32_t* col_idx, int32_t* seg_ptr,
                                 int32_t* contrib, void* stream) {
  TC_CHECK(ws && row_ptr && col_idx && seg_ptr && contrib && nnz > 0, TC_ERR_ARG,
           "tc_pattern_finish: bad arguments");
  int64_t N = E * k * k;
  PatternWs w = pattern_layout(const_cast<void*>(ws), N);
  hipStream_t s = (hipStream_t)stream;
  int grid = (int)ceil_div(N, kBlock);
  hipLaunchKernelGGL(k_pattern_finish, dim3(grid), dim3(kBlock), 0, s, w.key_out, w.slot, N,
                     n, nnz, row_ptr, col_idx, seg_ptr);
  TC_LAUNCH();
  TC_HIP(hipMemcpyAsync(contrib, w.id_out, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, s));
  return TC_OK;
}

extern "C" int tc_assemble_p1(const double* nodes, const int32_t* elems, int64_t E, int k,
                              const double* sig, const double* sl, const double* st,
                              const double* fib, const int32_t* seg_ptr,
                              const int32_t* contrib, int64_t nnz, double* local_ws,
                              int round_f32, double* M_vals, double* K_vals, void* stream) {
  TC_CHECK(nodes && elems && (sig || (sl && st && fib)) && seg_ptr && contrib && local_ws &&
               M_vals && K_vals && E > 0 && nnz > 0 && (k == 3 || k == 4),
           TC_ERR_ARG, "tc_assemble_p1: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int kk = k * k;
  double* Kloc = local_ws;
  double* Mloc = local_ws + (size_t)kk * E;
  int grid = (int)ceil_div(E, kBlock);
  if (k == 4)
    hipLaunchKernelGGL(k_local_tet, dim3(grid), dim3(kBlock), 0, s, nodes, elems, E, sig, sl,
                       st, fib, round_f32, Kloc, Mloc);
  else
    hipLaunchKernelGGL(k_local_tri, dim3(grid), dim3(kBlock), 0, s, nodes, elems, E, sig, sl,
                       st, fib, round_f32, Kloc, Mloc);
  TC_LAUNCH();
  hipLaunchKernelGGL(k_gather_sum, dim3((int)ceil_div(nnz, kBlock)), dim3(kBlock), 0, s,
                     seg_ptr, contrib, nnz, E, kk, Kloc, Mloc, K_vals, M_vals);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_operator_values(const double* M, const double* K, const int32_t* row_ptr,
                                  const int32_t* col, int64_t n, int64_t nnz, double Chi,
                                  double Cm, double theta, double dt, int dtype, void* Mv,
                                  void* Kv, void* Av, void* Dinv, void* stream) {
  TC_CHECK(M && K && row_ptr && col && Mv && Kv && Av && Dinv && n > 0 && nnz > 0,
           TC_ERR_ARG, "tc_operator_values: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  int grid = (int)ceil_div(n, kBlock);
  const double tdt = theta * dt;
  if (dtype == TC_F32)
    hipLaunchKernelGGL(k_operator_values<float>, dim3(grid), dim3(kBlock), 0, s, M, K, row_ptr,
                       col, n, Chi, Cm, tdt, (float*)Mv, (float*)Kv, (float*)Av, (float*)Dinv);
  else if (dtype == TC_F64)
    hipLaunchKernelGGL(k_operator_values<double>, dim3(grid), dim3(kBlock), 0, s, M, K,
                       row_ptr, col, n, Chi, Cm, tdt, (double*)Mv, (double*)Kv, (double*)Av,
                       (double*)Dinv);
  else
    TC_CHECK(false, TC_ERR_ARG, "bad dtype %d", dtype);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_sell_layout(const int32_t* row_ptr, int64_t n, int32_t* slice_ptr,
                              int64_t* padded_host, void* stream) {
  TC_CHECK(row_ptr && slice_ptr && padded_host && n > 0, TC_ERR_ARG,
           "tc_sell_layout: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  int64_t ns = ceil_div(n, kWave);
  // The widths are staged in slice_ptr[0..ns) and scanned in place by one
  // block: each thread owns one contiguous chunk, sums it before a barrier
  // and then rewrites it element by element, reading each width first.
  hipLaunchKernelGGL(k_slice_width, dim3((int)ceil_div(ns, kBlock)), dim3(kBlock), 0, s,
                     row_ptr, n, ns, slice_ptr);
  TC_LAUNCH();
  hipLaunchKernelGGL(k_slice_scan, dim3(1), dim3(1024), 0, s, slice_ptr, ns, slice_ptr);
  TC_LAUNCH();
  int32_t total = 0;
  TC_HIP(hipMemcpyAsync(&total, slice_ptr + ns, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  TC_HIP(hipStreamSynchronize(s));
  *padded_host = total;
  return TC_OK;
}

extern "C" int tc_sell_share(const int32_t* row_ptr, const int32_t* col, int64_t n,
                             const int32_t* slice_ptr, int32_t* pat, int32_t* slice_tag,
                             void* stream) {
  TC_CHECK(row_ptr && col && slice_ptr && pat && slice_tag && n > 0, TC_ERR_ARG,
           "tc_sell_share: bad arguments");
  const int64_t ns = ceil_div(n, kWave);
  hipLaunchKernelGGL(k_sell_share, dim3((int)ceil_div(ns * kWave, kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, row_ptr, col, n, ns, slice_ptr, pat, slice_tag);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_sell_columns(const int32_t* row_ptr, const int32_t* col, int64_t n,
                               const int32_t* slice_ptr, int32_t* scol, const int32_t* pat,
                               const int32_t* slice_tag, void* stream) {
  TC_CHECK(row_ptr && col && slice_ptr && scol && n > 0, TC_ERR_ARG,
           "tc_sell_columns: bad arguments");
  int64_t ns = ceil_div(n, kWave);
  hipLaunchKernelGGL(k_sell_columns, dim3((int)ceil_div(ns * kWave, kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, row_ptr, col, n, ns, slice_ptr, scol, pat, slice_tag);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_sell_deltas(const tc_sell* A, int16_t* dcol, int32_t* max_abs, void* stream) {
  TC_CHECK(A && A->slice_ptr && A->col && dcol && max_abs && A->n > 0, TC_ERR_ARG,
           "tc_sell_deltas: bad arguments");
  const int64_t ns = ceil_div(A->n, kWave);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, max_abs);
  TC_LAUNCH();
  hipLaunchKernelGGL(k_sell_deltas, dim3((int)ceil_div(ns * kWave, kBlock)), dim3(kBlock), 0, s,
                     A->col, ns, A->slice_ptr, dcol, max_abs);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_sell_values(const int32_t* row_ptr, const int32_t* col_for_values, int64_t n,
                              const int32_t* slice_ptr, const void* v, void* sv, int dtype,
                              const int32_t* pat, const int32_t* slice_tag, void* stream) {
  TC_CHECK(row_ptr && slice_ptr && v && sv && n > 0, TC_ERR_ARG, "tc_sell_values: bad arguments");
  TC_CHECK(!pat || (col_for_values && slice_tag), TC_ERR_ARG,
           "tc_sell_values: a pattern needs the CSR columns and the slice tags");
  int64_t ns = ceil_div(n, kWave);
  dim3 g((int)ceil_div(ns * kWave, kBlock)), b(kBlock);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TC_F32)
    hipLaunchKernelGGL(k_sell_values<float>, g, b, 0, s, row_ptr, col_for_values, n, ns,
                       slice_ptr, (const float*)v, (float*)sv, pat, slice_tag);
  else if (dtype == TC_F64)
    hipLaunchKernelGGL(k_sell_values<double>, g, b, 0, s, row_ptr, col_for_values, n, ns,
                       slice_ptr, (const double*)v, (double*)sv, pat, slice_tag);
  else
    TC_CHECK(false, TC_ERR_ARG, "bad dtype %d", dtype);
  TC_LAUNCH();
  return TC_OK;
}

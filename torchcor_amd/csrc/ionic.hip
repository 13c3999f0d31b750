// Fused node-parallel ionic updates: one thread per node reads V and its
// states once, interpolates the lookup-table rows, evaluates the currents,
// advances the states (forward Euler for concentrations, Rush-Larsen for
// gates, all from the old state) and writes the states, -Iion and the
// right-hand-side input w = Cm V + dt (dV/100 + Istim/100) in one pass.
// The same pass applies the activation/repolarisation-time test to V, which
// is the previous step's solution.
//
// Lookup tables are computed on the device in float64 on the reference's
// float32 grids and stored in the working dtype.  The interpolation index
// and lower grid position are computed with explicitly rounded float32
// operations so that a float32 run picks the same table rows as the
// reference's float32 run.
#include "tc_internal.h"

namespace tc {
namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------ table grids

struct TabParam {
  double mn, mx, res, step;
  int64_t mx_idx;
  int64_t rows;
};

// (mn, mx, res) as in the reference; step = 1/res and
// mx_idx = int((mx-mn)*step) - 1 in double, rows = ceil((mx-mn)/res).
TabParam make_tab(double mn, double mx, double res) {
  TabParam t;
  t.mn = mn;
  t.mx = mx;
  t.res = res;
  t.step = 1 / res;
  t.mx_idx = (int64_t)((mx - mn) * t.step) - 1;
  t.rows = (int64_t)ceil((mx - mn) / res);
  return t;
}

// Grid point i of torch's float32 CPU arange(start, end, res): the base of
// each 8-wide vector is formed in double, lanes add k*float(res) with a
// single rounding.
__device__ inline double grid_f32(const TabParam& t, int64_t i) {
#pragma clang fp contract(off)
  const int64_t q = i & ~(int64_t)7;
  const float base = (float)(t.mn + t.res * (double)q);
  const double k = (double)(i - q);
  return (double)(float)((double)base + k * (double)(float)t.res);
}

template <class T>
struct Lerp {
  int64_t lo;
  T w;
};

// Reference interpolate() (TenTusscherPanfilov.py:188-195) in float32.
// Every operation rounds separately, as in the reference's elementwise tensor
// ops: contraction into fused multiply-adds is switched off here (the HIP
// *_rn helpers are plain operators and would still be fused).
__device__ inline Lerp<float> lerp_at(float X, const TabParam& t) {
#pragma clang fp contract(off)
  const float mn = (float)t.mn, mx = (float)t.mx, res = (float)t.res, st = (float)t.step;
  X = X < mn ? mn : (X > mx ? mx : X);
  const float s = (X - mn) * st;
  int64_t idx = (int64_t)s;
  idx = idx < 0 ? 0 : (idx > t.mx_idx - 1 ? t.mx_idx - 1 : idx);
  const float lp = (float)idx * res + mn;
  Lerp<float> L;
  L.lo = idx;
  L.w = __fdiv_rn(X - lp, res);
  return L;
}

// Same in float64; lower_pos stays float32 as the reference's int64 x float
// promotion makes it.
__device__ inline Lerp<double> lerp_at(double X, const TabParam& t) {
#pragma clang fp contract(off)
  X = X < t.mn ? t.mn : (X > t.mx ? t.mx : X);
  const double s = (X - t.mn) * t.step;
  int64_t idx = (int64_t)s;
  idx = idx < 0 ? 0 : (idx > t.mx_idx - 1 ? t.mx_idx - 1 : idx);
  const double lp = (double)((float)idx * (float)t.res + (float)t.mn);
  Lerp<double> L;
  L.lo = idx;
  L.w = (X - lp) / t.res;
  return L;
}

template <class T>
struct Row {
  const T* r0;
  T w, w1;
  __device__ inline T operator()(int c, int ncol) const {
    return w1 * r0[c] + w * r0[c + ncol];
  }
};

// Both interpolation rows of a table with NS (a multiple of 16 bytes) values
// per row, loaded into registers with 16-byte vector loads.
template <class T> struct Vec16;
template <> struct Vec16<float> { typedef float4 type; static constexpr int n = 4; };
template <> struct Vec16<double> { typedef double2 type; static constexpr int n = 2; };

template <class T, int NS>
struct RowV {
  T lo[NS], hi[NS];
  T w, w1;
  __device__ inline T operator()(int c, int) const { return w1 * lo[c] + w * hi[c]; }
};

// Position of the interpolation rows without loading them; load_group(g)
// then fetches the 16-byte group g (columns g*E .. g*E+E-1) of both rows.
template <class T, int NS>
struct RowG : RowV<T, NS> {
  const T* base;
  __device__ inline void at(const T* tab, T X, const TabParam& t) {
    const Lerp<T> L = lerp_at(X, t);
    base = tab + L.lo * NS;
    this->w = L.w;
    this->w1 = T(1) - L.w;
  }
  __device__ inline void load_group(int g) {
    typedef typename Vec16<T>::type V;
    constexpr int E = Vec16<T>::n;
    const V* p = reinterpret_cast<const V*>(base);
    const V a = p[g], b = p[g + NS / E];
    const T* x = reinterpret_cast<const T*>(&a);
    const T* y = reinterpret_cast<const T*>(&b);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      this->lo[g * E + e] = x[e];
      this->hi[g * E + e] = y[e];
    }
  }
};

template <class T>
__device__ inline Row<T> row_at(const T* tab, int ncol, T X, const TabParam& t) {
  Lerp<T> L = lerp_at(X, t);
  Row<T> R;
  R.r0 = tab + L.lo * ncol;
  R.w = L.w;
  R.w1 = T(1) - L.w;
  return R;
}

__device__ inline float t_exp(float x) { return expf(x); }
__device__ inline double t_exp(double x) { return exp(x); }
__device__ inline float t_log(float x) { return logf(x); }
__device__ inline double t_log(double x) { return log(x); }
__device__ inline float t_expm1(float x) { return expm1f(x); }
__device__ inline double t_expm1(double x) { return expm1(x); }

// Where entry i of a launch lives and what its epilogue applies.
// Element idx of an array with a wave-uniform base, addressed by a 32-bit
// byte offset: loads and stores then use a scalar base register and one
// offset register per lane.  The host checks that every per-node index of
// an ionic step stays below 2^29, so idx * 8 fits 32 bits.
template <class T>
__device__ inline T& at32(T* base, uint32_t idx) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (uint32_t)sizeof(T));
}
template <class T>
__device__ inline const T& at32(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     idx * (uint32_t)sizeof(T));
}

// Pool slice list and per-system fields (tc_ionic_io's pool_* members),
// carried separately where they come from read-only kernel arguments.
struct PoolArgs {
  const int32_t* slices;
  const int32_t* slice_sys;
  const int64_t* row0;
  const int64_t* n;
  const double* t;
  const long long* steps;
  const uint32_t* bits;
};

struct NodeCtx {
  int64_t vi;          // entry in the per-node arrays (V, w, dV, stim_mask, AT, RT)
  int64_t si, ss;      // state index and state row stride
  int64_t sp0;         // state block offset (blocks of 64 entries; 0 in rows)
  uint32_t bits;       // stimuli on at this step for this entry
  const double* amp;   // intensity of stimulus k at amp[k * amp_stride]
  int amp_stride;
  int lat;             // LAT/RT test on, at lat_t
  double lat_t;
  int ok;              // a real node (pool padding rows are skipped)
};

// Plain: entry i.  Batch group (lanes > 0): lane lane0 + i % lanes of node
// i / lanes, rows of lane_stride entries.  Pool: row slice*64 + i%64 of the
// listed slice, with its system's row range, stimulus clock and LAT time.
template <bool PA = false>
__device__ inline NodeCtx node_ctx(const tc_ionic_io& io, int64_t i,
                                   const PoolArgs& pa = PoolArgs{}) {
  NodeCtx c;
  if (io.pool_slices) {
    // one slice per wavefront: a scalar index, so its descriptors load as scalars
    const PoolArgs q = PA ? pa
                          : PoolArgs{io.pool_slices, io.pool_slice_sys, io.pool_sys_row0,
                                     io.pool_sys_n, io.pool_sys_t, io.pool_sys_steps,
                                     io.pool_sys_bits};
    const int64_t sl = q.slices[__builtin_amdgcn_readfirstlane((int)(i >> 6))];
    const int64_t row = sl * 64 + (i & 63);
    const int sys = q.slice_sys[sl];
    c.vi = row;
    c.si = row;
    c.ss = io.state_stride;
    c.sp0 = 0;
    c.ok = row - q.row0[sys] < q.n[sys];
    const uint32_t on = q.bits[sys];
    c.bits = (on && io.stim_mask) ? io.stim_mask[row] & on : 0u;
    c.amp = &io.pool_sys_stim[(int64_t)sys * TC_POOL_STIM].amp;
    c.amp_stride = (int)(sizeof(tc_stimulus) / sizeof(double));
    c.lat = io.AT != nullptr && q.steps[sys] > 0;
    c.lat_t = q.t[sys];
    return c;
  }
  if (io.lanes > 0) {
    const int64_t node = i / io.lanes;
    c.vi = node * io.lane_stride + io.lane0 + (i - node * io.lanes);
  } else {
    c.vi = i;
  }
  c.si = i;
  c.ss = io.state_stride > 0 ? io.state_stride : io.n;
  c.sp0 = 0;
  c.ok = 1;
  c.bits = (io.stim_active && io.stim_mask) ? at32(io.stim_mask, (uint32_t)c.vi) & io.stim_active
                                             : 0u;
  c.amp = io.stim_amp;
  c.amp_stride = 1;
  c.lat = io.AT != nullptr;
  c.lat_t = io.lat_t;
  return c;
}

// States in blocks of 64 entries (io.state_block): entry i's state k at
// (i / 64) * S * 64 + k * 64 + i % 64, so the S loads and stores of a
// wavefront all fall in one contiguous S * 256-byte (float32) span instead
// of S rows n entries apart.  S is the model's state count.
template <int S>
__device__ inline void state_blocks(NodeCtx& c, const tc_ionic_io& io) {
  if (io.state_block) {
    c.sp0 = (c.si >> 6) * (int64_t)(S * 64);
    c.si &= 63;
    c.ss = 64;
  }
}

// State k of this node: row k of the (S, stride) block, or its place in
// the node's block of 64.
template <class T>
__device__ inline T& state(T* st, const NodeCtx& c, int k) {
  return at32(st + c.sp0 + (int64_t)k * c.ss, (uint32_t)c.si);
}

// State k of this node, with non-temporal loads and stores when NT (the
// io's stream_states): the states stream through once per step (76 MB each
// way for TTP at 1 M nodes) and so no longer displace the operator values
// and columns, which the RHS and PCG passes then find in the last-level
// cache step after step.
template <bool NT, class T>
__device__ inline T ld_state(const T* st, const NodeCtx& c, int k) {
  const T* q = &at32(st + c.sp0 + (int64_t)k * c.ss, (uint32_t)c.si);
  if constexpr (NT) return __builtin_nontemporal_load(q);
  return *q;
}

template <bool NT, class T>
__device__ inline void st_state(T* st, const NodeCtx& c, int k, T v) {
  T* q = &at32(st + c.sp0 + (int64_t)k * c.ss, (uint32_t)c.si);
  if constexpr (NT) __builtin_nontemporal_store(v, q);
  else *q = v;
}

// The membrane potential: float64 (VF64, the runtime's float64 iterate) or T.
template <class T, bool VF64>
__device__ inline T load_v(const tc_ionic_io& io, int64_t vi) {
  if (VF64) return (T)at32((const double*)io.V, (uint32_t)vi);
  return at32((const T*)io.V, (uint32_t)vi);
}

// The pipeline's stall flag (0 without one), read as a plain wave-uniform
// load at kernel start; kernels test it before their first store, so the
// flag's latency overlaps the node loads.
__device__ inline int gate_value(const tc_ionic_io& io) { return io.gate ? *io.gate : 0; }

// AT of this node, loaded with the other node inputs (used by finish_node).
template <class T>
__device__ inline T load_at(const tc_ionic_io& io, const NodeCtx& c) {
  return c.lat ? at32((const T*)io.AT, (uint32_t)c.vi) : T(0);
}

// Stimulus, RHS input and LAT/RT epilogue shared by all models.
template <class T>
__device__ inline void finish_node(const tc_ionic_io& io, const NodeCtx& c, T V, T dV, T at) {
  const uint32_t i = (uint32_t)c.vi;
  if (io.dV) at32((T*)io.dV, i) = dV;
  if (io.w) {
    const uint32_t bits = c.bits;
    if (io.w_f64) {
      // float64 right-hand-side input from the model's dV
      double b = (double)V * io.Cm + io.dt_rhs * ((double)dV / 100.0);
      if (bits) {
        double ist = 0.0;
        for (int k = 0; k < TC_MAX_STIMULI; ++k)
          if (bits & (1u << k)) ist += c.amp[k * c.amp_stride];
        b = b + io.dt_rhs * (ist / 100.0);
      }
      at32((double*)io.w, i) = b;
    } else {
      // the reference's order of operations in the working dtype
      const T du = dV / T(100);
      T b = V * (T)io.Cm + (T)io.dt_rhs * du;
      if (bits) {
        T ist = T(0);
        for (int k = 0; k < TC_MAX_STIMULI; ++k)
          if (bits & (1u << k)) ist += (T)c.amp[k * c.amp_stride];
        b = b + (T)io.dt_rhs * (ist / T(100));
      }
      at32((T*)io.w, i) = b;
    }
  }
  if (c.lat) {
    T* AT = (T*)io.AT;
    T* RT = (T*)io.RT;
    if (V > T(0) && at == T(-1)) {
      at = (T)c.lat_t;
      at32(AT, i) = at;
    }
    if (at > T(0) && V < T(-70) && at32(RT, i) == T(-1)) at32(RT, i) = (T)c.lat_t;
  }
}

// ---------------------------------------------------------------- MS2v

template <class T, bool VF64, bool NT, bool PA = false>
__device__ __forceinline__ void ms2v_node(const tc_ms2v_params& p, T* __restrict__ H,
                                          const tc_ionic_io& io, int64_t i,
                                          const PoolArgs& pa = PoolArgs{}) {
  if (i >= io.n) return;
  const int gated = gate_value(io);
  NodeCtx c = node_ctx<PA>(io, i, pa);
  state_blocks<1>(c, io);
  if (!c.ok) return;
  const T U = load_v<T, VF64>(io, c.vi);
  const T h = ld_state<NT>(H, c, 0);
  const T at = load_at<T>(io, c);
  const T DV = (T)(p.vmax - p.vmin);
  const T Uad = (U - (T)p.vmin) / DV;
  const T J_in = T(-1) * h * Uad * (Uad - (T)p.u_crit) * (T(1) - Uad) / (T)p.tau_in;
  const T J_out = (T(1) - h) * Uad / (T)p.tau_out;
  const T dU = -(DV * (J_in + J_out));
  const T dH = Uad > (T)p.u_gate ? -h / (T)p.tau_close : (T(1) - h) / (T)p.tau_open;
  if (gated) return;
  st_state<NT>(H, c, 0, T(h + (T)io.dt * dH));
  finish_node<T>(io, c, U, dU, at);
}

template <class T, bool VF64, bool NT>
__global__ void __launch_bounds__(kBlock)
k_ms2v(tc_ms2v_params p, T* __restrict__ H, tc_ionic_io io) {
  ms2v_node<T, VF64, NT>(p, H, io, (int64_t)blockIdx.x * kBlock + threadIdx.x);
}

// ---------------------------------------------------------------- TTP

enum {
  TTP_D_A, TTP_D_B, TTP_F2_A, TTP_F2_B, TTP_F_A, TTP_F_B, TTP_H_A, TTP_H_B, TTP_INaCa_A,
  TTP_INaCa_B, TTP_J_A, TTP_J_B, TTP_M_A, TTP_M_B, TTP_R_A, TTP_R_B, TTP_S_A, TTP_S_B,
  TTP_Xr1_A, TTP_Xr1_B, TTP_Xr2_A, TTP_Xr2_B, TTP_Xs_A, TTP_Xs_B, TTP_a2, TTP_rec_iNaK,
  TTP_rec_ipK, TTP_VCOLS
};
// row stride of the V table: the 27 columns and one zero, 16-byte rows
constexpr int TTP_VSTRIDE = 28;
enum { TS_CaSR, TS_CaSS, TS_Cai, TS_D, TS_F, TS_F2, TS_FCaSS, TS_H, TS_J, TS_Ki, TS_M, TS_Nai,
       TS_R, TS_R_, TS_S, TS_Xr1, TS_Xr2, TS_Xs, TS_N };

struct TtpTabs {
  TabParam V, CaSS, VEk;
};
TtpTabs ttp_tabs() {
  return {make_tab(-800., 800., 0.05), make_tab(0.00001, 10.0, 0.00001),
          make_tab(-800., 800., 0.05)};
}

// Table columns (TenTusscherPanfilov.py:197-318), float64.
__global__ void k_ttp_vtab(tc_ttp_params p, double dt, TabParam t, void* out, int f64) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.rows) return;
  const double V = grid_f32(t, i);
  const double RTONF = (p.Rconst * p.T) / p.Fconst;
  const double F_RT = 1. / RTONF;
  const double invKmCa_Cao = 1. / (p.KmCa + p.Cao);
  const double KmNai3 = p.KmNai * p.KmNai * p.KmNai;
  const double Nao3 = p.Nao * p.Nao * p.Nao;
  const double pmf_INaCa = p.knaca * (1. / (KmNai3 + Nao3)) * invKmCa_Cao;
  const double D_inf = 1. / (1. + exp(((-8. + p.D_CaL_off) - V) / 7.5));
  const double F2_inf = 0.67 / (1. + exp((V + 35.) / 7.)) + 0.33;
  const double F_inf = 1. / (1. + exp((V + 20.) / 7.));
  const double eh = 1. + exp((V + 71.55) / 7.43);
  const double H_inf = 1. / (eh * eh);
  const double em = 1. + exp((-56.86 - V) / 9.03);
  const double M_inf = 1. / (em * em);
  const double R_inf = 1. / (1. + exp((20. - V) / 6.));
  const double S_inf = 1. / (1. + exp((V + 20.) / 5.));
  const double Xr1_inf = 1. / (1. + exp((-26. - V) / 7.));
  const double Xr2_inf = 1. / (1. + exp((V - (-88. + p.xr2_off)) / 24.));
  const double Xs_inf = 1. / (1. + exp((-p.vHalfXs - V) / 14.));
  const double a2 = 0.25 * exp((2. * (V - 15.)) * F_RT);
  const double aa_D = 1.4 / (1. + exp((-35. - V) / 13.)) + 0.25;
  const double aa_F = 1102.5 * exp((-(V + 27.) * (V + 27.)) / 225.);
  const double aa_F2 = 562. * exp((-(V + 27.) * (V + 27.)) / 240.);
  const double aa_H = V >= -40. ? 0. : 0.057 * exp(-(V + 80.) / 6.8);
  const double aa_J = V >= -40. ? 0.
      : ((-2.5428e4 * exp(0.2444 * V) - 6.948e-6 * exp(-0.04391 * V)) * (V + 37.78)) /
            (1. + exp(0.311 * (V + 79.23)));
  const double aa_M = 1. / (1. + exp((-60. - V) / 5.));
  const double aa_Xr1 = 450. / (1. + exp((-45. - V) / 10.));
  const double aa_Xr2 = 3. / (1. + exp((-60. - V) / 20.));
  const double aa_Xs = 1400. / sqrt(1. + exp((5. - V) / 6.));
  const double bb_D = 1.4 / (1. + exp((V + 5.) / 5.));
  const double bb_F = 200. / (1. + exp((13. - V) / 10.));
  const double bb_F2 = 31. / (1. + exp((25. - V) / 10.));
  const double bb_H = V >= -40. ? 0.77 / (0.13 * (1. + exp(-(V + 10.66) / 11.1)))
                                : 2.7 * exp(0.079 * V) + 3.1e5 * exp(0.3485 * V);
  const double bb_J = V >= -40. ? (0.6 * exp(0.057 * V)) / (1. + exp(-0.1 * (V + 32.)))
                                : (0.02424 * exp(-0.01052 * V)) / (1. + exp(-0.1378 * (V + 40.14)));
  const double bb_M = 0.1 / (1. + exp((V + 35.) / 5.)) + 0.10 / (1. + exp((V - 50.) / 200.));
  const double bb_Xr1 = 6. / (1. + exp((V - (-30.)) / 11.5));
  const double bb_Xr2 = 1.12 / (1. + exp((V - 60.) / 20.));
  const double bb_Xs = 1. / (1. + exp((V - 35.) / 15.));
  const double cc_D = 1. / (1. + exp((50. - V) / 20.));
  const double cc_F = 180. / (1. + exp((V + 30.) / 10.)) + 20.;
  const double cc_F2 = 80. / (1. + exp((V + 30.) / 10.));
  const double den = pmf_INaCa / (1. + p.ksat * exp(((p.n - 1.) * V) * F_RT));
  const double tau_R = 9.5 * exp(-(V + 40.) * (V + 40.) / 1800.) + 0.8;
  const double tau_S = p.endo ? 1000. * exp(-(V + 67.) * (V + 67.) / 1000.) + 8.
                              : 85. * exp(-(V + 45.) * (V + 45.) / 320.) +
                                    5. / (1. + exp((V - 20.) / 5.)) + 3.;
  const double tau_D = aa_D * bb_D + cc_D;
  const double tau_F2 = aa_F2 + bb_F2 + cc_F2;
  const double tau_Ff = aa_F + bb_F + cc_F;
  const double tau_H = 1.0 / (aa_H + bb_H);
  const double tau_J = 1.0 / (aa_J + bb_J);
  const double tau_M = aa_M * bb_M;
  const double tau_Xr1 = aa_Xr1 * bb_Xr1;
  const double tau_Xr2 = aa_Xr2 * bb_Xr2;
  const double tau_Xs = aa_Xs * bb_Xs + 80.;
  const double tau_F = V > 0. ? tau_Ff * p.scl_tau_f : tau_Ff;
  double c[TTP_VCOLS];
#define RL(G, inf, tau)                    \
  c[TTP_##G##_A] = -(inf) * expm1(-dt / (tau)); \
  c[TTP_##G##_B] = exp(-dt / (tau));
  RL(D, D_inf, tau_D)
  RL(F2, F2_inf, tau_F2)
  RL(F, F_inf, tau_F)
  RL(H, H_inf, tau_H)
  RL(J, H_inf, tau_J)
  RL(M, M_inf, tau_M)
  RL(R, R_inf, tau_R)
  RL(S, S_inf, tau_S)
  RL(Xr1, Xr1_inf, tau_Xr1)
  RL(Xr2, Xr2_inf, tau_Xr2)
  RL(Xs, Xs_inf, tau_Xs)
#undef RL
  c[TTP_INaCa_A] = den * p.Cao * exp((p.n * V) * F_RT);
  c[TTP_INaCa_B] = den * exp(((p.n - 1.) * V) * F_RT) * Nao3 * 2.5;
  c[TTP_a2] = a2;
  c[TTP_rec_iNaK] = 1. / ((1. + 0.1245 * exp((-0.1 * V) * F_RT)) + 0.0353 * exp(-V * F_RT));
  c[TTP_rec_ipK] = 1. / (1. + exp((25. - V) / 5.98));
  for (int k = 0; k < TTP_VSTRIDE; ++k) {
    const double v = k < TTP_VCOLS ? c[k] : 0.0;
    if (f64) ((double*)out)[i * TTP_VSTRIDE + k] = v;
    else ((float*)out)[i * TTP_VSTRIDE + k] = (float)v;
  }
}

__global__ void k_ttp_casstab(double dt, TabParam t, void* out, int f64) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.rows) return;
  const double C = grid_f32(t, i);
  const double q = (C / 0.05) * (C / 0.05);
  const double inf = 0.6 / (1. + q) + 0.4;
  const double tau = 80. / (1. + q) + 2.;
  // column 0 (CaSS2) is never filled by the reference and stays zero
  const double c[3] = {0.0, -inf * expm1(-dt / tau), exp(-dt / tau)};
  for (int k = 0; k < 3; ++k) {
    if (f64) ((double*)out)[i * 3 + k] = c[k];
    else ((float*)out)[i * 3 + k] = (float)c[k];
  }
}

__global__ void k_ttp_vektab(TabParam t, void* out, int f64) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.rows) return;
  const double V = grid_f32(t, i);
  const double a = 0.1 / (1. + exp(0.06 * (V - 200.)));
  const double b = (3. * exp(0.0002 * (V + 100.)) + exp(0.1 * (V - 10.))) / (1. + exp(-0.5 * V));
  const double v = a / (a + b);
  if (f64) ((double*)out)[i] = v;
  else ((float*)out)[i] = (float)v;
}

// The constants of a TTP step in the working dtype.  Each is formed on the
// host in float64 from the parameters, then rounded once, exactly as the
// step would form it, so passing them precomputed changes no result; it
// keeps the kernel's scalar registers free for addresses.
template <class T>
struct TtpK {
  T RTONF, F_RT, hRTONF, inverseVcF2, inverseVssF2, invVcF_Cm, pmf_INaK, GKr_sqrtKo, a1c;
  T Cao, Ko, Nao, KoNa, pKNa, GpCa, KpCa, GKs, GNa, KmNa, GbCa, GbNa, GpK, Gto, GK1;
  T Vleak, Vmaxup, Kup2, Vxfer, EC, maxsr, dsr, Kbufc, CAPACITANCE, VsrVc, BufcK;
  T k4, k2_, k1_, k3, Vrel, Kbufsr, BufsrK, Kbufss, VcVss, VsrVss, BufssK;
};

template <class T>
TtpK<T> ttp_constants(const tc_ttp_params& p) {
  TtpK<T> K;
  const double RTONFd = (p.Rconst * p.T) / p.Fconst;
  K.RTONF = (T)RTONFd;
  K.F_RT = (T)(1. / RTONFd);
  K.hRTONF = (T)(0.5 * RTONFd);
  K.inverseVcF2 = (T)(1. / ((2. * p.Vc) * p.Fconst));
  K.inverseVssF2 = (T)(1. / ((2. * p.Vss) * p.Fconst));
  K.invVcF_Cm = (T)((1. / (p.Vc * p.Fconst)) * p.CAPACITANCE);
  K.pmf_INaK = (T)(p.knak * (p.Ko / (p.Ko + p.KmK)));
  K.GKr_sqrtKo = (T)(p.GKr * sqrt(p.Ko / 5.4));
  K.a1c = (T)(p.GCaL * p.Fconst * (1. / RTONFd) * 4.);
  K.Cao = (T)p.Cao;
  K.Ko = (T)p.Ko;
  K.Nao = (T)p.Nao;
  K.KoNa = (T)(p.Ko + p.pKNa * p.Nao);
  K.pKNa = (T)p.pKNa;
  K.GpCa = (T)p.GpCa;
  K.KpCa = (T)p.KpCa;
  K.GKs = (T)p.GKs;
  K.GNa = (T)p.GNa;
  K.KmNa = (T)p.KmNa;
  K.GbCa = (T)p.GbCa;
  K.GbNa = (T)p.GbNa;
  K.GpK = (T)p.GpK;
  K.Gto = (T)p.Gto;
  K.GK1 = (T)p.GK1;
  K.Vleak = (T)p.Vleak;
  K.Vmaxup = (T)p.Vmaxup;
  K.Kup2 = (T)(p.Kup * p.Kup);
  K.Vxfer = (T)p.Vxfer;
  K.EC = (T)p.EC;
  K.maxsr = (T)p.maxsr;
  K.dsr = (T)(p.maxsr - p.minsr);
  K.Kbufc = (T)p.Kbufc;
  K.CAPACITANCE = (T)p.CAPACITANCE;
  K.VsrVc = (T)(p.Vsr / p.Vc);
  K.BufcK = (T)(p.Bufc * p.Kbufc);
  K.k4 = (T)p.k4;
  K.k2_ = (T)p.k2_;
  K.k1_ = (T)p.k1_;
  K.k3 = (T)p.k3;
  K.Vrel = (T)p.Vrel;
  K.Kbufsr = (T)p.Kbufsr;
  K.BufsrK = (T)(p.Bufsr * p.Kbufsr);
  K.Kbufss = (T)p.Kbufss;
  K.VcVss = (T)(p.Vc / p.Vss);
  K.VsrVss = (T)(p.Vsr / p.Vss);
  K.BufssK = (T)(p.Bufss * p.Kbufss);
  return K;
}

// One TTP step per node (TenTusscherPanfilov.py:353-454).
template <class T, bool VF64, bool NT, bool PA = false>
__device__ __forceinline__ void ttp_node(const TtpK<T>& K, T* __restrict__ st,
                                         const T* __restrict__ Vtab, const T* __restrict__ Ctab,
                                         const T* __restrict__ Ktab, const TtpTabs& tb,
                                         const tc_ionic_io& io, int64_t i,
                                         const PoolArgs& pa = PoolArgs{}) {
  const int64_t n = io.n;
  if (i >= n) return;
  const int gated = gate_value(io);
  NodeCtx c = node_ctx<PA>(io, i, pa);
  state_blocks<TS_N>(c, io);
  if (!c.ok) return;
  const T V = load_v<T, VF64>(io, c.vi);
  T s[TS_N];
#pragma unroll
  for (int k = 0; k < TS_N; ++k) s[k] = ld_state<NT>(st, c, k);
  const T at = load_at<T>(io, c);
  const T RTONF = K.RTONF, F_RT = K.F_RT;

  // Phase 1: currents and concentrations, with the V-table columns they use
  // (INaCa_A/B, a2, rec_iNaK, rec_ipK) and the CaSS and VEk tables; the
  // concentration states, FCaSS and the epilogue are written here.
  RowG<T, TTP_VSTRIDE> vr;
  vr.at(Vtab, V, tb.V);
  constexpr int E = Vec16<T>::n;
  // 16-byte groups holding the phase-1 columns
  constexpr unsigned ph1 = (1u << (TTP_INaCa_A / E)) | (1u << (TTP_INaCa_B / E)) |
                           (1u << (TTP_a2 / E)) | (1u << (TTP_rec_iNaK / E)) |
                           (1u << (TTP_rec_ipK / E));
#pragma unroll
  for (int g = 0; g < TTP_VSTRIDE / E; ++g)
    if (ph1 >> g & 1u) vr.load_group(g);
  const Row<T> cr = row_at<T>(Ctab, 3, s[TS_CaSS], tb.CaSS);
  const T Cai = s[TS_Cai], CaSS = s[TS_CaSS], CaSR = s[TS_CaSR], Ki = s[TS_Ki], Nai = s[TS_Nai];
  const T Eca = K.hRTONF * t_log(K.Cao / Cai);
  const T Ek = RTONF * t_log(K.Ko / Ki);
  const T Eks = RTONF * t_log(K.KoNa / (Ki + K.pKNa * Nai));
  const T Ena = RTONF * t_log(K.Nao / Nai);
  const T IpCa = (K.GpCa * Cai) / (K.KpCa + Cai);
  const T V15 = V - T(15);
  const T ghk = V == T(15) ? T(0.5) * F_RT : V15 / t_expm1((T(2) * V15) * F_RT);
  const T a1 = K.a1c * ghk;
  const T ICaL_A = a1 * vr(TTP_a2, TTP_VCOLS);
  const T ICaL_B = a1 * K.Cao;
  const T IKr = K.GKr_sqrtKo * s[TS_Xr1] * s[TS_Xr2] * (V - Ek);
  const T IKs = K.GKs * s[TS_Xs] * s[TS_Xs] * (V - Eks);
  const T m = s[TS_M];
  const T INa = K.GNa * m * m * m * s[TS_H] * s[TS_J] * (V - Ena);
  const T INaK = K.pmf_INaK * (Nai / (Nai + K.KmNa)) * vr(TTP_rec_iNaK, TTP_VCOLS);
  const T IbCa = K.GbCa * (V - Eca);
  const T IbNa = K.GbNa * (V - Ena);
  const T IpK = K.GpK * vr(TTP_rec_ipK, TTP_VCOLS) * (V - Ek);
  const T Ito = K.Gto * s[TS_R] * s[TS_S] * (V - Ek);
  const Row<T> kr = row_at<T>(Ktab, 1, V - Ek, tb.VEk);
  const T ICaL = s[TS_D] * s[TS_F] * s[TS_F2] * s[TS_FCaSS] * (ICaL_A * CaSS - ICaL_B);
  const T INaCa = vr(TTP_INaCa_A, TTP_VCOLS) * Nai * Nai * Nai - vr(TTP_INaCa_B, TTP_VCOLS) * Cai;
  const T IK1 = K.GK1 * kr(0, 1) * (V - Ek);
  const T Iion = IKr + IKs + IK1 + Ito + INa + IbNa + ICaL + IbCa + INaK + INaCa + IpCa + IpK;

  const T Ileak = K.Vleak * (CaSR - Cai);
  const T Iup = K.Vmaxup / (T(1) + K.Kup2 / (Cai * Cai));
  const T Ixfer = K.Vxfer * (CaSS - Cai);
  const T diff_Ki = -((IK1 + Ito + IKr + IKs - T(2) * INaK) + IpK) * K.invVcF_Cm;
  const T diff_Nai = -((INa + IbNa + T(3) * INaK) + T(3) * INaCa) * K.invVcF_Cm;
  const T ecs = K.EC / CaSR;
  const T kCaSR = K.maxsr - K.dsr / (T(1) + ecs * ecs);
  const T kbc = K.Kbufc + Cai;
  const T diff_Cai = ((Ixfer - ((IbCa + IpCa - T(2) * INaCa) * K.inverseVcF2) * K.CAPACITANCE) -
                      (Iup - Ileak) * K.VsrVc) /
                     (T(1) + (K.BufcK / kbc) / kbc);
  const T R_ = s[TS_R_];
  const T diff_R_ = K.k4 * (T(1) - R_) - K.k2_ * kCaSR * CaSS * R_;
  const T k1 = K.k1_ / kCaSR;
  const T c2 = cr(0, 3);
  const T O = (k1 * c2 * R_) / (K.k3 + k1 * c2);
  const T Irel = K.Vrel * O * (CaSR - CaSS);
  const T ksr = CaSR + K.Kbufsr;
  const T diff_CaSR = (Iup - Irel - Ileak) / (T(1) + (K.BufsrK / ksr) / ksr);
  const T kss = CaSS + K.Kbufss;
  const T diff_CaSS = ((-Ixfer * K.VcVss + Irel * K.VsrVss) +
                       (-ICaL * K.inverseVssF2) * K.CAPACITANCE) /
                      (T(1) + (K.BufssK / kss) / kss);
  const T dt = (T)io.dt;
  if (gated) return;
  st_state<NT>(st, c, TS_CaSR, T(CaSR + diff_CaSR * dt));
  st_state<NT>(st, c, TS_CaSS, T(CaSS + diff_CaSS * dt));
  st_state<NT>(st, c, TS_Cai, T(Cai + diff_Cai * dt));
  st_state<NT>(st, c, TS_Ki, T(Ki + diff_Ki * dt));
  st_state<NT>(st, c, TS_Nai, T(Nai + diff_Nai * dt));
  st_state<NT>(st, c, TS_R_, T(R_ + diff_R_ * dt));
  st_state<NT>(st, c, TS_FCaSS, T(cr(1, 3) + cr(2, 3) * s[TS_FCaSS]));
  finish_node<T>(io, c, V, -Iion, at);
  // Phase 2 (no memory access moves across this point): the remaining
  // V-table columns and the Rush-Larsen gate updates from the old gates.
  asm volatile("" ::: "memory");
#pragma unroll
  for (int g = 0; g < TTP_VSTRIDE / E; ++g)
    if (!(ph1 >> g & 1u)) vr.load_group(g);
  T o[TS_N];
  o[TS_D] = vr(TTP_D_A, TTP_VCOLS) + vr(TTP_D_B, TTP_VCOLS) * s[TS_D];
  o[TS_F] = vr(TTP_F_A, TTP_VCOLS) + vr(TTP_F_B, TTP_VCOLS) * s[TS_F];
  o[TS_F2] = vr(TTP_F2_A, TTP_VCOLS) + vr(TTP_F2_B, TTP_VCOLS) * s[TS_F2];
  o[TS_H] = vr(TTP_H_A, TTP_VCOLS) + vr(TTP_H_B, TTP_VCOLS) * s[TS_H];
  o[TS_J] = vr(TTP_J_A, TTP_VCOLS) + vr(TTP_J_B, TTP_VCOLS) * s[TS_J];
  o[TS_M] = vr(TTP_M_A, TTP_VCOLS) + vr(TTP_M_B, TTP_VCOLS) * s[TS_M];
  o[TS_R] = vr(TTP_R_A, TTP_VCOLS) + vr(TTP_R_B, TTP_VCOLS) * s[TS_R];
  o[TS_S] = vr(TTP_S_A, TTP_VCOLS) + vr(TTP_S_B, TTP_VCOLS) * s[TS_S];
  o[TS_Xr1] = vr(TTP_Xr1_A, TTP_VCOLS) + vr(TTP_Xr1_B, TTP_VCOLS) * s[TS_Xr1];
  o[TS_Xr2] = vr(TTP_Xr2_A, TTP_VCOLS) + vr(TTP_Xr2_B, TTP_VCOLS) * s[TS_Xr2];
  o[TS_Xs] = vr(TTP_Xs_A, TTP_VCOLS) + vr(TTP_Xs_B, TTP_VCOLS) * s[TS_Xs];
  const int gates[11] = {TS_D, TS_F, TS_F2, TS_H, TS_J, TS_M, TS_R, TS_S, TS_Xr1, TS_Xr2, TS_Xs};
#pragma unroll
  for (int k = 0; k < 11; ++k) st_state<NT>(st, c, gates[k], T(o[gates[k]]));
}

template <class T, bool VF64, bool NT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_ttp(TtpK<T> K, T* __restrict__ st, const T* __restrict__ Vtab,
      const T* __restrict__ Ctab, const T* __restrict__ Ktab, TtpTabs tb, tc_ionic_io io) {
  ttp_node<T, VF64, NT>(K, st, Vtab, Ctab, Ktab, tb, io, (int64_t)blockIdx.x * kBlock + threadIdx.x);
}

// ---------------------------------------------------------------- CRN

enum { CA_carow_1, CA_carow_2, CA_carow_3, CA_conCa, CA_f_Ca_A, CA_NCOLS };
enum {
  CV_GKur, CV_INaK, CV_IbNa, CV_d_A, CV_d_B, CV_f_A, CV_f_B, CV_h_A, CV_h_B, CV_j_A, CV_j_B,
  CV_m_A, CV_m_B, CV_oa_A, CV_oa_B, CV_oi_A, CV_oi_B, CV_ua_A, CV_ua_B, CV_ui_A, CV_ui_B,
  CV_vrow_29, CV_vrow_31, CV_vrow_32, CV_vrow_36, CV_vrow_7, CV_w_A, CV_w_B, CV_xr_A, CV_xr_B,
  CV_xs_A, CV_xs_B, CV_NCOLS
};
enum { CF_u_A, CF_v_A, CF_v_B, CF_NCOLS };
enum { CS_Ca_rel, CS_Ca_up, CS_Cai, CS_Ki, CS_d, CS_f, CS_f_Ca, CS_h, CS_j, CS_m, CS_oa, CS_oi,
       CS_u, CS_ua, CS_ui, CS_v, CS_w, CS_xr, CS_xs, CS_N };

struct CrnTabs {
  TabParam Cai, V, fn;
};
CrnTabs crn_tabs() {
  return {make_tab(3e-4, 30.0, 3e-4), make_tab(-200., 200., 0.1),
          make_tab(-2e-11, 10.0e-11, 2e-15)};
}

__device__ inline void put(void* out, int64_t idx, double v, int f64) {
  if (f64) ((double*)out)[idx] = v;
  else ((float*)out)[idx] = (float)v;
}

// Tables of CourtemancheRamirezNattel.py:212-328.
__global__ void k_crn_caitab(tc_crn_params p, double dt, TabParam t, void* out, int f64) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.rows) return;
  const double Ca = grid_f32(t, i);
  const double conCa = Ca / 1000;
  const double f_Ca_C = expm1(-dt / p.tau_f_Ca);
  double c[CA_NCOLS];
  c[CA_conCa] = conCa;
  c[CA_carow_1] = (p.factorGup * p.maxIup) / (1. + p.K_up / conCa);
  c[CA_carow_2] = ((((p.maxTrpn * p.KmTrpn) / (conCa + p.KmTrpn)) / (conCa + p.KmTrpn) +
                    ((p.maxCmdn * p.KmCmdn) / (conCa + p.KmCmdn)) / (conCa + p.KmCmdn) + 1.) /
                   p.C_B1c) / 1000.;
  c[CA_carow_3] = (p.maxIpCa * conCa) / (0.0005 + conCa) -
                  ((((p.GbCa * p.R) * p.T) / 2.) / p.F) * log(p.Cao / conCa);
  c[CA_f_Ca_A] = -(1. / (1. + conCa / 0.00035)) * f_Ca_C;
  for (int k = 0; k < CA_NCOLS; ++k) put(out, i * CA_NCOLS + k, c[k], f64);
}

__global__ void k_crn_vtab(tc_crn_params p, double dt, TabParam t, void* out, int f64) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.rows) return;
  const double V = grid_f32(t, i);
  const double R = p.R, T_ = p.T, F = p.F;
  const double E_Na = ((R * T_) / F) * log(p.Nao / p.Nai);
  const double sigma = (exp(p.Nao / 67.3) - 1.) / 7.;
  const double a_h = V >= -40. ? 0.0 : 0.135 * exp(((V + 80.) - p.factorhGate) / -6.8);
  const double a_j = V < -40. ? ((-127140. * exp(0.2444 * V) - 3.474e-5 * exp(-0.04391 * V)) *
                                 (V + 37.78)) / (1. + exp(0.311 * (V + 79.23)))
                              : 0.;
  const double a_m = V == -47.13 ? 3.2 : (0.32 * (V + 47.13)) / (1. - exp(-0.1 * (V + 47.13)));
  const double aa_oa = 0.65 / (exp((V + 10.) / -8.5) + exp((30. - V) / 59.0));
  const double aa_oi = 1. / (18.53 + exp((V + 113.7) / 10.95));
  const double aa_ua = 0.65 / (exp((V + 10.) / -8.5) + exp((V - 30.) / -59.0));
  const double aa_ui = 1. / (21. + exp((V - 185.) / -28.));
  const double aa_xr = p.factorxrGate * ((0.0003 * (V + 14.1)) / (1. - exp((V + 14.1) / -5.)));
  const double aa_xs = (4.e-5 * (V - 19.9)) / (1. - exp((19.9 - V) / 17.));
  const double b_h = V >= -40. ? (1. / 0.13) / (1. + exp(-(V + 10.66) / 11.1))
                               : 3.56 * exp(0.079 * V) + 3.1e5 * exp(0.35 * V);
  const double b_j = V >= -40. ? (0.3 * exp(-2.535e-7 * V)) / (1. + exp(-0.1 * (V + 32.)))
                               : (0.1212 * exp(-0.01052 * V)) / (1. + exp(-0.1378 * (V + 40.14)));
  const double b_m = 0.08 * exp(-(V - p.factormGate) / 11.);
  const double bb_oa = 0.65 / (2.5 + exp((V + 82.) / 17.));
  const double bb_oi = 1. / (35.56 + exp(-(V + 1.26) / 7.44));
  const double bb_ua = 0.65 / (2.5 + exp((V + 82.) / 17.));
  const double bb_ui = exp((V - 158.) / 16.);
  const double bb_xr = (1.0 / p.factorxrGate) *
                       ((7.3898e-5 * (V - 3.3328)) / (exp((V - 3.3328) / 5.1237) - 1.));
  const double bb_xs = (3.5e-5 * (V - 19.9)) / (exp((V - 19.9) / 9.) - 1.);
  const double d_inf = 1. / (1. + exp(-(V + 10.) / 8.));
  const double f_NaK = 1. / ((1. + 0.1245 * exp((((-0.1 * F) * V) / R) / T_)) +
                             (0.0365 * sigma) * exp(((-F * V) / R) / T_));
  const double f_inf = 1. / (1. + exp((V + 28.) / 6.9));
  const double oa_inf = 1. / (1. + exp(-((V + 20.47) - p.factoroaGate) / 17.54));
  const double oi_inf = 1. / (1. + exp((V + 43.1) / 5.3));
  const double tau_d = V == -10. ? ((1. / 6.24) / 0.035) / 2.
      : ((1. - exp(-(V + 10.) / 6.24)) / 0.035) / (V + 10.) / (1. + exp(-(V + 10.) / 6.24));
  const double tau_f = 9. / (0.0197 * exp(((-0.0337 * 0.0337) * (V + 10.)) * (V + 10.)) + 0.02);
  const double tau_w = ((6. * (1. - exp((7.9 - V) / 5.))) / (1. + 0.3 * exp((7.9 - V) / 5.))) /
                       (V - 7.9);
  const double ua_inf = 1. / (1. + exp((V + 30.3) / -9.6));
  const double ui_inf = 1. / (1. + exp((V - 99.45) / 27.48));
  const double w_inf = 1. - 1. / (1. + exp((40. - V) / 17.));
  const double xr_inf = 1. / (1. + exp((V + 14.1) / -6.5));
  const double xs_inf = 1. / sqrt(1. + exp((V - 19.9) / -12.7));
  const double Nao3 = p.Nao * p.Nao * p.Nao;
  const double sat = 1. + p.k_sat * exp(((((p.gamma - 1.) * F) * V) / R) / T_);
  double c[CV_NCOLS];
  c[CV_GKur] = 0.005 + 0.05 / (1. + exp((15. - V) / 13.));
  c[CV_IbNa] = p.GbNa * (V - E_Na);
  c[CV_vrow_29] = p.GCaL * (V - 65.);
  c[CV_vrow_31] = ((((p.maxINaCa * exp((((p.gamma * F) * V) / R) / T_)) * p.Nai * p.Nai * p.Nai *
                     p.Cao) / (p.KmNa3 + Nao3)) / (p.KmCa + p.Cao)) / sat;
  c[CV_vrow_32] = (((((p.maxINaCa * exp(((((p.gamma - 1.) * F) * V) / R) / T_)) * Nao3) /
                     (p.KmNa3 + Nao3)) / (p.KmCa + p.Cao)) / sat) / 1000.;
  c[CV_vrow_36] = V * p.GbCa;
  c[CV_vrow_7] = p.GNa * (V - E_Na);
  c[CV_INaK] = (((p.maxINaK * f_NaK) / (1. + pow(p.KmNai / p.Nai, 1.5))) * p.Ko) / (p.Ko + p.KmKo);
#define RL(G, inf, tau)                         \
  c[CV_##G##_A] = -(inf) * expm1(-dt / (tau));  \
  c[CV_##G##_B] = exp(-dt / (tau));
  RL(d, d_inf, tau_d)
  RL(f, f_inf, tau_f)
  RL(w, w_inf, tau_w)
  RL(oa, oa_inf, (1. / (aa_oa + bb_oa)) / p.K_Q10)
  RL(oi, oi_inf, (1. / (aa_oi + bb_oi)) / p.K_Q10)
  RL(ua, ua_inf, (1. / (aa_ua + bb_ua)) / p.K_Q10)
  RL(ui, ui_inf, (1. / (aa_ui + bb_ui)) / p.K_Q10)
  RL(xr, xr_inf, 1. / (aa_xr + bb_xr))
  RL(xs, xs_inf, 0.5 / (aa_xs + bb_xs))
#undef RL
  c[CV_h_A] = (-a_h / (a_h + b_h)) * expm1(-dt * (a_h + b_h));
  c[CV_h_B] = exp(-dt * (a_h + b_h));
  c[CV_j_A] = (-a_j / (a_j + b_j)) * expm1(-dt * (a_j + b_j));
  c[CV_j_B] = exp(-dt * (a_j + b_j));
  c[CV_m_A] = (-a_m / (a_m + b_m)) * expm1(-dt * (a_m + b_m));
  c[CV_m_B] = exp(-dt * (a_m + b_m));
  for (int k = 0; k < CV_NCOLS; ++k) put(out, i * CV_NCOLS + k, c[k], f64);
}

__global__ void k_crn_fntab(tc_crn_params p, double dt, TabParam t, void* out, int f64) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.rows) return;
  const double fn = grid_f32(t, i);
  const double u_C = expm1(-dt / p.tau_u);
  const double tau_v = 1.91 + 2.09 / (1. + exp((3.4175e-13 - fn) / 13.67e-16));
  const double u_inf = 1. / (1. + exp((3.4175e-13 - fn) / 13.67e-16));
  const double v_inf = 1. - 1. / (1. + exp((6.835e-14 - fn) / 13.67e-16));
  put(out, i * 3 + CF_u_A, -u_inf * u_C, f64);
  put(out, i * 3 + CF_v_A, -v_inf * expm1(-dt / tau_v), f64);
  put(out, i * 3 + CF_v_B, exp(-dt / tau_v), f64);
}

// The constants of a CRN step in the working dtype, formed on the host in
// float64 from the parameters and the step length, then rounded once (as the
// step would form them).
template <class T>
struct CrnK {
  T RTF, Ko, gach, Gto, factorGKur, GKr, GKs, GK1, factorGtr, tau_tr, factorGrel, k_rel,
      maxIup_Caup, KmCsqn, C_dCa_rel, FVoli, C_dCaup, C_B1d, C_B1e, C_Fn1, C_Fn2, e_fCa, e_u;
};

template <class T>
CrnK<T> crn_constants(const tc_crn_params& p, double dt) {
  CrnK<T> K;
  K.RTF = (T)((p.R * p.T) / p.F);
  K.Ko = (T)p.Ko;
  K.gach = (T)(p.GACh * (10.0 / (1.0 + 9.13652 / pow(p.ACh, 0.477811))));
  K.Gto = (T)p.Gto;
  K.factorGKur = (T)p.factorGKur;
  K.GKr = (T)p.GKr;
  K.GKs = (T)p.GKs;
  K.GK1 = (T)p.GK1;
  K.factorGtr = (T)p.factorGtr;
  K.tau_tr = (T)p.tau_tr;
  K.factorGrel = (T)p.factorGrel;
  K.k_rel = (T)p.k_rel;
  K.maxIup_Caup = (T)(p.maxIup / p.maxCaup);
  K.KmCsqn = (T)p.KmCsqn;
  K.C_dCa_rel = (T)p.C_dCa_rel;
  K.FVoli = (T)(p.F * p.Voli);
  K.C_dCaup = (T)p.C_dCaup;
  K.C_B1d = (T)p.C_B1d;
  K.C_B1e = (T)p.C_B1e;
  K.C_Fn1 = (T)p.C_Fn1;
  K.C_Fn2 = (T)p.C_Fn2;
  K.e_fCa = (T)exp(-dt / p.tau_f_Ca);
  K.e_u = (T)exp(-dt / p.tau_u);
  return K;
}

// One CRN step per node (CourtemancheRamirezNattel.py:358-444), in two
// phases like the TTP step: currents and concentrations first (with the
// V-table groups they read), then the Rush-Larsen gates.
template <class T, bool VF64, bool NT, bool PA = false>
__device__ __forceinline__ void crn_node(const CrnK<T>& K, T* __restrict__ st,
                                         const T* __restrict__ Ctab, const T* __restrict__ Vtab,
                                         const T* __restrict__ Ftab, const CrnTabs& tb,
                                         const tc_ionic_io& io, int64_t i,
                                         const PoolArgs& pa = PoolArgs{}) {
  const int64_t n = io.n;
  if (i >= n) return;
  const int gated = gate_value(io);
  NodeCtx c = node_ctx<PA>(io, i, pa);
  state_blocks<CS_N>(c, io);
  if (!c.ok) return;
  const T V = load_v<T, VF64>(io, c.vi);
  T s[CS_N];
#pragma unroll
  for (int k = 0; k < CS_N; ++k) s[k] = ld_state<NT>(st, c, k);
  const T at = load_at<T>(io, c);
  RowG<T, CV_NCOLS> vr;
  vr.at(Vtab, V, tb.V);
  constexpr int E = Vec16<T>::n;
  // 16-byte groups holding the phase-1 columns
  constexpr unsigned ph1 = (1u << (CV_GKur / E)) | (1u << (CV_INaK / E)) | (1u << (CV_IbNa / E)) |
                           (1u << (CV_vrow_29 / E)) | (1u << (CV_vrow_31 / E)) |
                           (1u << (CV_vrow_32 / E)) | (1u << (CV_vrow_36 / E)) |
                           (1u << (CV_vrow_7 / E));
#pragma unroll
  for (int g = 0; g < CV_NCOLS / E; ++g)
    if (ph1 >> g & 1u) vr.load_group(g);
  const Row<T> car = row_at<T>(Ctab, CA_NCOLS, s[CS_Cai], tb.Cai);
  const T E_K = K.RTF * t_log(K.Ko / s[CS_Ki]);
  const T VmEK = V - E_K;
  const T ICaL = vr(CV_vrow_29, CV_NCOLS) * s[CS_d] * s[CS_f] * s[CS_f_Ca];
  const T IKACh = K.gach * (T(0.0517) + T(0.4516) / (T(1) + t_exp((V + T(59.53)) / T(17.18)))) * VmEK;
  const T m = s[CS_m];
  const T INa = vr(CV_vrow_7, CV_NCOLS) * m * m * m * s[CS_h] * s[CS_j];
  const T INaCa = vr(CV_vrow_31, CV_NCOLS) - s[CS_Cai] * vr(CV_vrow_32, CV_NCOLS);
  const T vrow_13 = K.Gto * VmEK;
  const T vrow_18 = (K.factorGKur * vr(CV_GKur, CV_NCOLS)) * VmEK;
  const T vrow_21 = (K.GKr * VmEK) / (T(1) + t_exp((V + T(15)) / T(22.4)));
  const T vrow_24 = K.GKs * VmEK;
  const T IK1 = (K.GK1 * VmEK) / (T(1) + t_exp(T(0.07) * (V + T(80))));
  const T IKr = vrow_21 * s[CS_xr];
  const T IKs = vrow_24 * s[CS_xs] * s[CS_xs];
  const T ua = s[CS_ua];
  const T IKur = vrow_18 * ua * ua * ua * s[CS_ui];
  const T IpCa = car(CA_carow_3, CA_NCOLS) + vr(CV_vrow_36, CV_NCOLS);
  const T oa = s[CS_oa];
  const T Ito = vrow_13 * oa * oa * oa * s[CS_oi];
  const T INaK = vr(CV_INaK, CV_NCOLS);
  const T Iion = INa + IK1 + Ito + IKur + IKr + IKs + ICaL + IpCa + INaCa +
                 vr(CV_IbNa, CV_NCOLS) + INaK + IKACh;

  const T Ca_rel = s[CS_Ca_rel], Ca_up = s[CS_Ca_up];
  const T Itr = (K.factorGtr * (Ca_up - Ca_rel)) / K.tau_tr;
  const T u = s[CS_u];
  const T Irel = K.factorGrel * u * u * s[CS_v] * K.k_rel * s[CS_w] *
                 (Ca_rel - car(CA_conCa, CA_NCOLS));
  const T dIups = car(CA_carow_1, CA_NCOLS) - K.maxIup_Caup * Ca_up;
  const T kc = Ca_rel + K.KmCsqn;
  const T diff_Ca_rel = (Itr - Irel) / (T(1) + (K.C_dCa_rel / kc) / kc);
  const T diff_Ki = -((Ito + IKr + IKur + IKs + IK1 + IKACh) - T(2) * INaK) / K.FVoli;
  const T diff_Ca_up = dIups - Itr * K.C_dCaup;
  const T diff_Cai = ((K.C_B1d * ((INaCa + INaCa - IpCa) - ICaL) - K.C_B1e * dIups) + Irel) /
                     car(CA_carow_2, CA_NCOLS);
  const T dt = (T)io.dt;
  const T fn = K.C_Fn1 * Irel - K.C_Fn2 * (ICaL - T(0.4) * INaCa);
  const Row<T> fr = row_at<T>(Ftab, CF_NCOLS, fn, tb.fn);
  if (gated) return;
  st_state<NT>(st, c, CS_Ca_rel, T(Ca_rel + diff_Ca_rel * dt));
  st_state<NT>(st, c, CS_Ca_up, T(Ca_up + diff_Ca_up * dt));
  st_state<NT>(st, c, CS_Cai, T(s[CS_Cai] + diff_Cai * dt));
  st_state<NT>(st, c, CS_Ki, T(s[CS_Ki] + diff_Ki * dt));
  st_state<NT>(st, c, CS_f_Ca, T(car(CA_f_Ca_A, CA_NCOLS) + K.e_fCa * s[CS_f_Ca]));
  st_state<NT>(st, c, CS_u, T(fr(CF_u_A, CF_NCOLS) + K.e_u * u));
  st_state<NT>(st, c, CS_v, T(fr(CF_v_A, CF_NCOLS) + fr(CF_v_B, CF_NCOLS) * s[CS_v]));
  finish_node<T>(io, c, V, -Iion, at);
  // Phase 2 (no memory access moves across this point): the gate columns
  // of the V table and the Rush-Larsen updates from the old gates.
  asm volatile("" ::: "memory");
#pragma unroll
  for (int g = 0; g < CV_NCOLS / E; ++g) vr.load_group(g);
#define RLU(G) \
  st_state<NT>(st, c, CS_##G, T(vr(CV_##G##_A, CV_NCOLS) + vr(CV_##G##_B, CV_NCOLS) * s[CS_##G]));
  RLU(d) RLU(f) RLU(h) RLU(j) RLU(m) RLU(oa) RLU(oi) RLU(ua) RLU(ui) RLU(w) RLU(xr) RLU(xs)
#undef RLU
}

template <class T, bool VF64, bool NT>
__global__ void __launch_bounds__(kBlock)
k_crn(CrnK<T> K, T* __restrict__ st, const T* __restrict__ Ctab,
      const T* __restrict__ Vtab, const T* __restrict__ Ftab, CrnTabs tb, tc_ionic_io io) {
  crn_node<T, VF64, NT>(K, st, Ctab, Vtab, Ftab, tb, io, (int64_t)blockIdx.x * kBlock + threadIdx.x);
}

// The three models of a pool in one launch: blocks [0, b1) step the MS2v
// entries, [b1, b2) the TTP entries, the rest the CRN entries (each model's
// io lists its own slices).  One launch instead of three removes two kernel
// boundaries from every pooled time step.  PoolIonic gathers the host side
// of the arguments.
template <class T>
struct PoolIonic {
  tc_ms2v_params mp;
  TtpK<T> tk;
  CrnK<T> ck;
  T* st[3];
  const T* ttab[3];
  const T* ctab[3];
  TtpTabs ttb;
  CrnTabs ctb;
  tc_ionic_io io[3];
  tc_ionic_io sh;  // an io of a model present in the launch: the per-system fields
  int b1, b2;
};

// (separate by-value arguments: the kernel argument block is read in place)
template <class T, bool VF64, bool NT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_pool_ionic(tc_ms2v_params mp, TtpK<T> tk, CrnK<T> ck, T* __restrict__ st0,
             T* __restrict__ st1, T* __restrict__ st2, const T* __restrict__ tt0,
             const T* __restrict__ tt1, const T* __restrict__ tt2, const T* __restrict__ ct0,
             const T* __restrict__ ct1, const T* __restrict__ ct2, TtpTabs ttb, CrnTabs ctb,
             tc_ionic_io io0, tc_ionic_io io1, tc_ionic_io io2, int b1, int b2,
             const int32_t* __restrict__ sl0, const int32_t* __restrict__ sl1,
             const int32_t* __restrict__ sl2, const int32_t* __restrict__ slice_sys,
             const int64_t* __restrict__ sys_row0, const int64_t* __restrict__ sys_n,
             const double* __restrict__ sys_t, const long long* __restrict__ sys_steps,
             const uint32_t* __restrict__ sys_bits) {
  // The slice lists and per-system fields also come as read-only kernel
  // arguments (the io copies of the same pointers are not used): loads
  // through them, uniform per wavefront, then compile to scalar loads, so a
  // wavefront's chain slice -> system -> row range, stimulus bits and clock
  // does not wait on its vector memory counter.
  const int b = blockIdx.x;
  if (b < b1) {
    const PoolArgs pa{sl0, slice_sys, sys_row0, sys_n, sys_t, sys_steps, sys_bits};
    ms2v_node<T, VF64, NT, true>(mp, st0, io0, (int64_t)b * kBlock + threadIdx.x, pa);
  } else if (b < b2) {
    const PoolArgs pa{sl1, slice_sys, sys_row0, sys_n, sys_t, sys_steps, sys_bits};
    ttp_node<T, VF64, NT, true>(tk, st1, tt0, tt1, tt2, ttb, io1,
                                (int64_t)(b - b1) * kBlock + threadIdx.x, pa);
  } else {
    const PoolArgs pa{sl2, slice_sys, sys_row0, sys_n, sys_t, sys_steps, sys_bits};
    crn_node<T, VF64, NT, true>(ck, st2, ct0, ct1, ct2, ctb, io2,
                                (int64_t)(b - b2) * kBlock + threadIdx.x, pa);
  }
}

template <class T>
__global__ void k_lat(const T* __restrict__ u, T* __restrict__ AT, T* __restrict__ RT,
                      int64_t n, double t) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const T V = u[i];
  T at = AT[i];
  if (V > T(0) && at == T(-1)) {
    at = (T)t;
    AT[i] = at;
  }
  if (at > T(0) && V < T(-70) && RT[i] == T(-1)) RT[i] = (T)t;
}

int check_io(const tc_ionic_io* io, int dtype) {
  TC_CHECK(io && io->V && io->n > 0, TC_ERR_ARG, "ionic step: bad io");
  TC_CHECK(dtype == TC_F32 || dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", dtype);
  TC_CHECK(!io->AT || io->RT, TC_ERR_ARG, "ionic step: AT without RT");
  // per-node arrays are addressed with 32-bit byte offsets (at32)
  int64_t top = io->n;
  if (io->pool_slices) top = io->state_stride;
  else if (io->lanes > 0) top = (io->n - 1) / io->lanes * io->lane_stride + io->lane0 + io->lanes;
  TC_CHECK(io->pool_slices || io->state_block || io->state_stride == 0 ||
               io->state_stride >= io->n,
           TC_ERR_ARG, "ionic step: state stride %lld below n %lld",
           (long long)io->state_stride, (long long)io->n);
  TC_CHECK(!io->state_block || io->lanes == 0, TC_ERR_ARG,
           "ionic step: state blocks are not supported with lanes");
  TC_CHECK(top < ((int64_t)1 << 29), TC_ERR_SIZE,
           "ionic step: %lld per-node entries exceed the 2^29 limit", (long long)top);
  return TC_OK;
}


int check_tab(const TabParam& t) {
  TC_CHECK(t.mx_idx >= 1 && t.mx_idx <= t.rows - 1, TC_ERR_ARG,
           "table bounds: mx_idx %lld rows %lld", (long long)t.mx_idx, (long long)t.rows);
  return TC_OK;
}

}  // namespace
}  // namespace tc

using namespace tc;

int tc::ms2v_step(const tc_ms2v_params* p, void* states, const tc_ionic_io* io, int dtype,
                  hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  int rc = check_io(io, dtype);
  if (rc) return rc;
  TC_CHECK(p && states, TC_ERR_ARG, "tc_ms2v_step: bad arguments");
  dim3 g((int)ceil_div(io->n, kBlock)), b(kBlock);
  // float32 kernels read V as float64 when io->v_f64 is set (the runtime's
  // float64 iterate); float64 kernels read double either way
  const bool nt = io->stream_states != 0;
  if (dtype == TC_F32 && io->v_f64)
    launch(nt ? k_ms2v<float, true, true> : k_ms2v<float, true, false>, g, b, s, e0, e1, *p,
           (float*)states, *io);
  else if (dtype == TC_F32)
    launch(nt ? k_ms2v<float, false, true> : k_ms2v<float, false, false>, g, b, s, e0, e1, *p,
           (float*)states, *io);
  else
    launch(nt ? k_ms2v<double, false, true> : k_ms2v<double, false, false>, g, b, s, e0, e1, *p,
           (double*)states, *io);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_ms2v_step(const tc_ms2v_params* p, void* states, const tc_ionic_io* io,
                            int dtype, void* stream) {
  return ms2v_step(p, states, io, dtype, (hipStream_t)stream, nullptr, nullptr);
}

extern "C" int tc_ttp_tables(const tc_ttp_params* p, double dt, void* V_tab, void* CaSS_tab,
                             void* VEk_tab, int dtype, void* stream) {
  TC_CHECK(p && V_tab && CaSS_tab && VEk_tab, TC_ERR_ARG, "tc_ttp_tables: bad arguments");
  TC_CHECK(dtype == TC_F32 || dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", dtype);
  TtpTabs tb = ttp_tabs();
  hipStream_t s = (hipStream_t)stream;
  int f64 = dtype == TC_F64;
  hipLaunchKernelGGL(k_ttp_vtab, dim3((int)ceil_div(tb.V.rows, kBlock)), dim3(kBlock), 0, s, *p,
                     dt, tb.V, V_tab, f64);
  TC_LAUNCH();
  hipLaunchKernelGGL(k_ttp_casstab, dim3((int)ceil_div(tb.CaSS.rows, kBlock)), dim3(kBlock), 0,
                     s, dt, tb.CaSS, CaSS_tab, f64);
  TC_LAUNCH();
  hipLaunchKernelGGL(k_ttp_vektab, dim3((int)ceil_div(tb.VEk.rows, kBlock)), dim3(kBlock), 0, s,
                     tb.VEk, VEk_tab, f64);
  TC_LAUNCH();
  return TC_OK;
}

int tc::ttp_step(const tc_ttp_params* p, void* states, const void* V_tab, const void* CaSS_tab,
                 const void* VEk_tab, const tc_ionic_io* io, int dtype, hipStream_t s,
                 hipEvent_t e0, hipEvent_t e1) {
  int rc = check_io(io, dtype);
  if (rc) return rc;
  TC_CHECK(p && states && V_tab && CaSS_tab && VEk_tab, TC_ERR_ARG, "tc_ttp_step: bad arguments");
  TtpTabs tb = ttp_tabs();
  if ((rc = check_tab(tb.V)) || (rc = check_tab(tb.CaSS)) || (rc = check_tab(tb.VEk))) return rc;
  dim3 g((int)ceil_div(io->n, kBlock)), b(kBlock);
  const bool nt = io->stream_states != 0;
  if (dtype == TC_F32 && io->v_f64)
    launch(nt ? k_ttp<float, true, true> : k_ttp<float, true, false>, g, b, s, e0, e1,
           ttp_constants<float>(*p), (float*)states, (const float*)V_tab,
           (const float*)CaSS_tab, (const float*)VEk_tab, tb, *io);
  else if (dtype == TC_F32)
    launch(nt ? k_ttp<float, false, true> : k_ttp<float, false, false>, g, b, s, e0, e1,
           ttp_constants<float>(*p), (float*)states, (const float*)V_tab,
           (const float*)CaSS_tab, (const float*)VEk_tab, tb, *io);
  else
    launch(nt ? k_ttp<double, false, true> : k_ttp<double, false, false>, g, b, s, e0, e1,
           ttp_constants<double>(*p), (double*)states, (const double*)V_tab,
           (const double*)CaSS_tab, (const double*)VEk_tab, tb, *io);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_ttp_step(const tc_ttp_params* p, void* states, const void* V_tab,
                           const void* CaSS_tab, const void* VEk_tab, const tc_ionic_io* io,
                           int dtype, void* stream) {
  return ttp_step(p, states, V_tab, CaSS_tab, VEk_tab, io, dtype, (hipStream_t)stream, nullptr,
                  nullptr);
}

extern "C" int tc_crn_tables(const tc_crn_params* p, double dt, void* Cai_tab, void* V_tab,
                             void* fn_tab, int dtype, void* stream) {
  TC_CHECK(p && Cai_tab && V_tab && fn_tab, TC_ERR_ARG, "tc_crn_tables: bad arguments");
  TC_CHECK(dtype == TC_F32 || dtype == TC_F64, TC_ERR_ARG, "bad dtype %d", dtype);
  CrnTabs tb = crn_tabs();
  hipStream_t s = (hipStream_t)stream;
  int f64 = dtype == TC_F64;
  hipLaunchKernelGGL(k_crn_caitab, dim3((int)ceil_div(tb.Cai.rows, kBlock)), dim3(kBlock), 0, s,
                     *p, dt, tb.Cai, Cai_tab, f64);
  TC_LAUNCH();
  hipLaunchKernelGGL(k_crn_vtab, dim3((int)ceil_div(tb.V.rows, kBlock)), dim3(kBlock), 0, s, *p,
                     dt, tb.V, V_tab, f64);
  TC_LAUNCH();
  hipLaunchKernelGGL(k_crn_fntab, dim3((int)ceil_div(tb.fn.rows, kBlock)), dim3(kBlock), 0, s,
                     *p, dt, tb.fn, fn_tab, f64);
  TC_LAUNCH();
  return TC_OK;
}

int tc::crn_step(const tc_crn_params* p, void* states, const void* Cai_tab, const void* V_tab,
                 const void* fn_tab, const tc_ionic_io* io, int dtype, hipStream_t s,
                 hipEvent_t e0, hipEvent_t e1) {
  int rc = check_io(io, dtype);
  if (rc) return rc;
  TC_CHECK(p && states && Cai_tab && V_tab && fn_tab, TC_ERR_ARG, "tc_crn_step: bad arguments");
  CrnTabs tb = crn_tabs();
  if ((rc = check_tab(tb.Cai)) || (rc = check_tab(tb.V)) || (rc = check_tab(tb.fn))) return rc;
  dim3 g((int)ceil_div(io->n, kBlock)), b(kBlock);
  const bool nt = io->stream_states != 0;
  if (dtype == TC_F32 && io->v_f64)
    launch(nt ? k_crn<float, true, true> : k_crn<float, true, false>, g, b, s, e0, e1,
           crn_constants<float>(*p, io->dt), (float*)states, (const float*)Cai_tab,
           (const float*)V_tab, (const float*)fn_tab, tb, *io);
  else if (dtype == TC_F32)
    launch(nt ? k_crn<float, false, true> : k_crn<float, false, false>, g, b, s, e0, e1,
           crn_constants<float>(*p, io->dt), (float*)states, (const float*)Cai_tab,
           (const float*)V_tab, (const float*)fn_tab, tb, *io);
  else
    launch(nt ? k_crn<double, false, true> : k_crn<double, false, false>, g, b, s, e0, e1,
           crn_constants<double>(*p, io->dt), (double*)states, (const double*)Cai_tab,
           (const double*)V_tab, (const double*)fn_tab, tb, *io);
  TC_LAUNCH();
  return TC_OK;
}

extern "C" int tc_crn_step(const tc_crn_params* p, void* states, const void* Cai_tab,
                           const void* V_tab, const void* fn_tab, const tc_ionic_io* io,
                           int dtype, void* stream) {
  return crn_step(p, states, Cai_tab, V_tab, fn_tab, io, dtype, (hipStream_t)stream, nullptr,
                  nullptr);
}

template <class T>
int pool_launch(const PoolIonic<T>& a, bool vf64, bool nt, int blocks, hipStream_t s) {
  auto k = vf64 ? (nt ? k_pool_ionic<T, true, true> : k_pool_ionic<T, true, false>)
                : (nt ? k_pool_ionic<T, false, true> : k_pool_ionic<T, false, false>);
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kBlock), 0, s, a.mp, a.tk, a.ck, a.st[0], a.st[1],
                     a.st[2], a.ttab[0], a.ttab[1], a.ttab[2], a.ctab[0], a.ctab[1], a.ctab[2],
                     a.ttb, a.ctb, a.io[0], a.io[1], a.io[2], a.b1, a.b2, a.io[0].pool_slices,
                     a.io[1].pool_slices, a.io[2].pool_slices, a.sh.pool_slice_sys,
                     a.sh.pool_sys_row0, a.sh.pool_sys_n, a.sh.pool_sys_t, a.sh.pool_sys_steps,
                     a.sh.pool_sys_bits);
  TC_LAUNCH();
  return TC_OK;
}

int tc::pool_ionic_step(const tc_ms2v_params* mp, const tc_ttp_params* tp,
                        const tc_crn_params* cp, void* const states[3],
                        const void* const ttp_tab[3], const void* const crn_tab[3],
                        const tc_ionic_io io[3], int dtype, hipStream_t s) {
  int64_t nb[3];
  for (int m = 0; m < 3; ++m) {
    nb[m] = io[m].n > 0 ? ceil_div(io[m].n, kBlock) : 0;
    if (nb[m] == 0) continue;
    int rc = check_io(&io[m], dtype);
    if (rc) return rc;
    TC_CHECK(states[m] && (m != 0 || mp) && (m != 1 || (tp && ttp_tab[0] && ttp_tab[1] && ttp_tab[2])) &&
                 (m != 2 || (cp && crn_tab[0] && crn_tab[1] && crn_tab[2])),
             TC_ERR_ARG, "pool ionic step: model %d lacks parameters, states or tables", m);
  }
  int vf = -1, ns = -1;
  for (int m = 0; m < 3; ++m) {
    if (nb[m] == 0) continue;
    TC_CHECK(vf < 0 || vf == io[m].v_f64, TC_ERR_ARG, "pool ionic step: mixed V types");
    TC_CHECK(ns < 0 || ns == (io[m].stream_states != 0), TC_ERR_ARG,
             "pool ionic step: mixed stream_states");
    vf = io[m].v_f64;
    ns = io[m].stream_states != 0;
  }
  // the launch reads one set of per-system fields for all models
  int m0 = -1;
  for (int m = 0; m < 3; ++m) {
    if (nb[m] == 0) continue;
    if (m0 < 0) {
      m0 = m;
      continue;
    }
    const tc_ionic_io &a = io[m0], &b = io[m];
    TC_CHECK(a.pool_slice_sys == b.pool_slice_sys && a.pool_sys_row0 == b.pool_sys_row0 &&
                 a.pool_sys_n == b.pool_sys_n && a.pool_sys_t == b.pool_sys_t &&
                 a.pool_sys_steps == b.pool_sys_steps && a.pool_sys_bits == b.pool_sys_bits,
             TC_ERR_ARG, "pool ionic step: models with different per-system fields");
  }
  const int64_t blocks = nb[0] + nb[1] + nb[2];
  if (blocks == 0) return TC_OK;
  TC_CHECK(blocks < (1LL << 31), TC_ERR_SIZE, "pool ionic step: too many entries");
  const bool vf64 = dtype == TC_F32 && vf == 1;
  TtpTabs ttb = ttp_tabs();
  CrnTabs ctb = crn_tabs();
  int rc;
  if ((rc = check_tab(ttb.V)) || (rc = check_tab(ttb.CaSS)) || (rc = check_tab(ttb.VEk)) ||
      (rc = check_tab(ctb.Cai)) || (rc = check_tab(ctb.V)) || (rc = check_tab(ctb.fn)))
    return rc;
  auto fill = [&](auto& a, auto tag) -> void {
    using T = decltype(tag);
    a.mp = mp ? *mp : tc_ms2v_params{};
    if (tp) a.tk = ttp_constants<T>(*tp);
    if (cp) a.ck = crn_constants<T>(*cp, io[2].dt);
    for (int m = 0; m < 3; ++m) {
      a.st[m] = (T*)states[m];
      a.io[m] = io[m];
      if (nb[m] == 0) a.io[m].n = 0;
      a.ttab[m] = (const T*)ttp_tab[m];
      a.ctab[m] = (const T*)crn_tab[m];
    }
    a.sh = io[m0];
    a.ttb = ttb;
    a.ctb = ctb;
    a.b1 = (int)nb[0];
    a.b2 = (int)(nb[0] + nb[1]);
  };
  if (dtype == TC_F32) {
    PoolIonic<float> a{};
    fill(a, 0.0f);
    return pool_launch(a, vf64, ns == 1, (int)blocks, s);
  }
  PoolIonic<double> a{};
  fill(a, 0.0);
  return pool_launch(a, false, ns == 1, (int)blocks, s);
}

extern "C" int tc_lat_update(const void* u, void* AT, void* RT, int64_t n, double t, int dtype,
                             void* stream) {
  TC_CHECK(u && AT && RT && n > 0, TC_ERR_ARG, "tc_lat_update: bad arguments");
  dim3 g((int)ceil_div(n, kBlock)), b(kBlock);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TC_F32)
    hipLaunchKernelGGL(k_lat<float>, g, b, 0, s, (const float*)u, (float*)AT, (float*)RT, n, t);
  else if (dtype == TC_F64)
    hipLaunchKernelGGL(k_lat<double>, g, b, 0, s, (const double*)u, (double*)AT, (double*)RT, n, t);
  else
    TC_CHECK(false, TC_ERR_ARG, "bad dtype %d", dtype);
  TC_LAUNCH();
  return TC_OK;
}

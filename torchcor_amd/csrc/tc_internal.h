// Shared helpers for the torchcor_amd HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>

#include "../../include/torchcor_amd.h"

namespace tc {

// Thread-local last error text returned by tc_last_error().
void set_error(const char* fmt, ...);

#define TC_HIP(expr)                                                           \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::tc::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,             \
                      hipGetErrorString(_e));                                  \
      return TC_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

#define TC_CHECK(cond, code, ...)                                              \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::tc::set_error(__VA_ARGS__);                                            \
      return (code);                                                           \
    }                                                                          \
  } while (0)

// Launch-error check after a kernel launch.
#define TC_LAUNCH()                                                            \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) {                                                    \
      ::tc::set_error("%s:%d kernel launch -> %s", __FILE__, __LINE__,         \
                      hipGetErrorString(_e));                                  \
      return TC_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

constexpr int kWave = 64;

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Wavefront index in the block as a wave-uniform (scalar) value, so that
// per-wavefront indices derived from it (slice, row, list entry) and the
// loads they address stay scalar.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Wave64 sum of a double with cross-lane shuffles.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum of NV doubles (blockDim.x a multiple of 64, <= 1024).
// Result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double red[NV][16];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += red[k][w];
      v[k] = s;
    }
  }
}

// Bit k set when stimulus k is on at time t (stimulation.py:51-63): on iff
// t/period <= count and start <= t % period <= start + duration, with
// Python's % on positive doubles equal to fmod.
__host__ __device__ inline uint32_t stimulus_bits(const tc_stimulus* st, int n_stim, double t) {
  uint32_t bits = 0;
  for (int k = 0; k < n_stim; ++k) {
    const tc_stimulus& s = st[k];
    const double p_t = fmod(t, s.period);
    const double t_t = t / s.period;
    if (t_t <= s.count && p_t >= s.start && p_t <= s.start + s.duration) bits |= 1u << k;
  }
  return bits;
}

// Kernel launch.  With a start/stop event pair the events take the
// kernel's own begin and end timestamps (hipExtLaunchKernel), so per-kernel
// timing adds no packet to the stream; null events launch plainly.
template <class F, class... Args>
inline void launch(F kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                   Args... args) {
  if (e0 || e1) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

// Kernel launch with dynamic LDS bytes (as launch).
template <class F, class... Args>
inline void launch_lds(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, hipEvent_t e0,
                       hipEvent_t e1, Args... args) {
  if (e0 || e1) hipExtLaunchKernelGGL(kernel, grid, block, lds, s, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

// Instantiated HIP graphs of an engine's fixed launch sequences, keyed by
// what a sequence depends on (buffer parity, iteration count).  A sequence
// is recorded once by stream capture on a private stream and then replayed
// with one hipGraphLaunch on the caller's stream.  Recorded sequences hold
// kernel launches only (no memset or copy nodes).  An engine clears its
// cache whenever an argument of a recorded launch changes.
struct GraphCache {
  static constexpr size_t kMax = 64;
  struct Ent {
    long long key, key2;
    hipGraphExec_t exec;
  };
  std::vector<Ent> ents;
  hipStream_t cap = nullptr;
  long long captures = 0, replays = 0;
  void clear() {
    for (Ent& e : ents) (void)hipGraphExecDestroy(e.exec);
    ents.clear();
  }
  void release() {
    clear();
    if (cap) (void)hipStreamDestroy(cap);
    cap = nullptr;
  }
};

// Replays the graph of (key, key2), recording it first with rec(stream)
// (which returns a TC_* code) when the cache has none.
template <class F>
int graph_replay(GraphCache& c, long long key, long long key2, hipStream_t s, F&& rec) {
  hipGraphExec_t exec = nullptr;
  for (const GraphCache::Ent& e : c.ents)
    if (e.key == key && e.key2 == key2) exec = e.exec;
  if (!exec) {
    if (c.ents.size() >= GraphCache::kMax) c.clear();
    if (!c.cap) TC_HIP(hipStreamCreateWithFlags(&c.cap, hipStreamNonBlocking));
    TC_HIP(hipStreamBeginCapture(c.cap, hipStreamCaptureModeRelaxed));
    const int rc = rec(c.cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c.cap, &g);
    if (rc != TC_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc != TC_OK) return rc;
      set_error("graph capture -> %s", hipGetErrorString(e));
      return TC_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      set_error("graph instantiate -> %s", hipGetErrorString(ei));
      return TC_ERR_HIP;
    }
    c.ents.push_back({key, key2, exec});
    ++c.captures;
  }
  TC_HIP(hipGraphLaunch(exec, s));
  ++c.replays;
  return TC_OK;
}

// Fused ionic steps (the tc_*_step entry points), the kernel optionally
// bracketed by timing events.
int ms2v_step(const tc_ms2v_params* p, void* states, const tc_ionic_io* io, int dtype,
              hipStream_t s, hipEvent_t e0, hipEvent_t e1);
int ttp_step(const tc_ttp_params* p, void* states, const void* V_tab, const void* CaSS_tab,
             const void* VEk_tab, const tc_ionic_io* io, int dtype, hipStream_t s, hipEvent_t e0,
             hipEvent_t e1);
int crn_step(const tc_crn_params* p, void* states, const void* Cai_tab, const void* V_tab,
             const void* fn_tab, const tc_ionic_io* io, int dtype, hipStream_t s, hipEvent_t e0,
             hipEvent_t e1);
// The three models of a pool (MS2v, TTP, CRN; io[m].n == 0 skips a model) in
// one launch.  Every io must read V the same way (v_f64).
int pool_ionic_step(const tc_ms2v_params* mp, const tc_ttp_params* tp, const tc_crn_params* cp,
                    void* const states[3], const void* const ttp_tab[3],
                    const void* const crn_tab[3], const tc_ionic_io io[3], int dtype,
                    hipStream_t s);

}  // namespace tc

// Shared helpers for the torchcor_amd HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>

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

// Deterministic grid reduction: every block stores its NV partial sums in
// partials[k*nblocks + blockIdx.x]; the last block to arrive (atomic ticket)
// adds them in block order and returns true in thread 0 with the totals in v.
// The ticket is reset by that block, so the counter can be reused by the
// next launch on the same stream.
template <int NV>
__device__ inline bool grid_sum_last(double (&v)[NV], double* partials,
                                     unsigned* ticket) {
  block_sum<NV>(v);
  __shared__ bool last;
  const unsigned nb = gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * nb + blockIdx.x] = v[k];
    __threadfence();
    unsigned t = atomicAdd(ticket, 1u);
    last = (t == nb - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  // Sum partials in fixed block order: each thread takes a strided subset,
  // then a fixed-shape block reduction; the order depends only on nb.
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (unsigned b = threadIdx.x; b < nb; b += blockDim.x)
      s += __hip_atomic_load(&partials[k * nb + b], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    acc[k] = s;
  }
  block_sum<NV>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = acc[k];
    *ticket = 0u;
  }
  return threadIdx.x == 0;
}

}  // namespace tc

// Error text and version of the C ABI.
#include "tc_internal.h"

namespace tc {
namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

}  // namespace tc

extern "C" const char* tc_last_error(void) { return tc::g_last_error.c_str(); }

extern "C" int tc_abi_version(void) { return 3; }

// ------------------------------------------------------------------------
// Plain device copy with one 16-byte load and store per thread: the
// bandwidth reference that the benchmark sets the bandwidth-bound passes
// against.
namespace tc {
namespace {
__global__ void __launch_bounds__(256)
k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < m) dst[i] = src[i];
}
}  // namespace
}  // namespace tc

extern "C" int tc_stream_copy(const void* src, void* dst, int64_t nbytes, void* stream) {
  TC_CHECK(src && dst && nbytes > 0 && nbytes % 16 == 0 && ((uintptr_t)src % 16) == 0 &&
               ((uintptr_t)dst % 16) == 0,
           TC_ERR_ARG, "tc_stream_copy: bad arguments (16-byte aligned, size a multiple of 16)");
  const int64_t m = nbytes / 16;
  hipLaunchKernelGGL(tc::k_stream_copy, dim3((unsigned)tc::ceil_div(m, 256)), dim3(256), 0,
                     (hipStream_t)stream, (const uint4*)src, (uint4*)dst, m);
  TC_LAUNCH();
  return TC_OK;
}

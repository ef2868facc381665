// util.hip -- a plain streaming copy, used to calibrate HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) against a kernel of known traffic.
#include <hip/hip_runtime.h>
#include "internal.h"

namespace {

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, long n) {
  const long stride = (long)gridDim.x * 256;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int qbg_stream_copy(qbg_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst || !src || bytes % 16) return QBG_E_INVAL;
  const long n = (long)(bytes / 16);
  const long blocks = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::max<long>(blocks, 1)), dim3(256), 0, c->stream,
                     (const uint4*)src, (uint4*)dst, n);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return c->fail(QBG_E_HIP, hipGetErrorString(e));
  return QBG_OK;
}

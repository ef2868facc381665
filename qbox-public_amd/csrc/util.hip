// util.hip -- a plain streaming copy, used to calibrate HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) against a kernel of known traffic.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "internal.h"

namespace {

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, long n) {
  const long stride = (long)gridDim.x * 256;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int qbg_stream_copy(qbg_ctx* c, void* dst, const void* src, size_t bytes) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !dst || !src || bytes % 16) return QBG_E_INVAL;
    const long n = (long)(bytes / 16);
    const long blocks = std::min<long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::max<long>(blocks, 1)), dim3(256), 0, c->stream,
                       (const uint4*)src, (uint4*)dst, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return c->fail(QBG_E_HIP, hipGetErrorString(e));
    return QBG_OK;
  });
}

namespace qbg {

// fill (used instead of memset nodes, which graphs must not record)
__global__ __launch_bounds__(256) void k_fill(double* __restrict__ p, size_t n, double v) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = v;
}
void fill_value(double* p, size_t n, double v, hipStream_t s) {
  if (!n) return;
  const size_t nb = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, s, p, n, v);
}
void fill_zero(double* p, size_t n, hipStream_t s) { fill_value(p, n, 0.0, s); }

// gather scattered device scalars into one (host-mapped) buffer:
// dst[i] = src segment by segment, segments (ptr, count) given in seg
__global__ __launch_bounds__(256) void k_gather(const double* const* __restrict__ seg_ptr,
                                                const int* __restrict__ seg_n, int nseg,
                                                double* __restrict__ dst) {
  int off = 0;
  for (int k = 0; k < nseg; k++) {
    const double* p = seg_ptr[k];
    for (int i = threadIdx.x; i < seg_n[k]; i += 256) dst[off + i] = p[i];
    off += seg_n[k];
  }
}
void gather_segments(const double* const* seg_ptr, const int* seg_n, int nseg, double* dst,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(1), dim3(256), 0, s, seg_ptr, seg_n, nseg, dst);
}

// out[i] = a[i] + b[i]
__global__ void k_add2(const double* __restrict__ a, const double* __restrict__ b,
                       double* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}
void add2(const double* a, const double* b, double* out, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_add2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, out, n);
}

}  // namespace qbg

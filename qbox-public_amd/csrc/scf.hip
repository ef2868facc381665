// scf.hip -- kernels of one SCF iteration after H*psi: the wavefunction
// steppers (src/SDWavefunctionStepper.cpp, src/PSDWavefunctionStepper.cpp,
// src/PSDAWavefunctionStepper.cpp) with the diagonal preconditioner
// (src/Preconditioner.cpp), the Higham / Bjorck-Bowie polar iteration of
// SlaterDet::ortho_align (src/Matrix.cpp:2446-2517), and the charge mixing of
// BOSampleStepper (src/BOSampleStepper.cpp:810-879) with the AndersonMixer
// (src/AndersonMixer.cpp).
//
// All of it is bandwidth-bound elementwise work over the coefficients
// (2 ngw x nst doubles) or the density (2 ngv doubles), plus single-block
// kernels for the small dense steps.  Decisions that depend on data (the
// polar iteration's convergence and branch, PSDA's extrapolation, the
// mixer's ring of past vectors) are flags in device memory that the kernels
// read, so the host enqueues a fixed sequence and the sequence can be
// captured into a graph and replayed.
#include <hip/hip_runtime.h>
#include <cmath>
#include "scf.h"

namespace qbg {

namespace {

// sum of NV values over the block (up to 1024 threads); result in thread 0
template <int NV>
__device__ __forceinline__ void bsum(double (&v)[NV]) {
  __shared__ double sh[NV][16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double t = 0.0;
      for (int i = 0; i < nw; i++) t += sh[k][i];
      v[k] = t;
    }
  __syncthreads();
}

// Preconditioner::diag (src/Preconditioner.cpp:48-85), basic adaptive form
// or fixed ecutprec, no confinement term
__device__ __forceinline__ double pdiag(double ekin_n, double g2, double ecutprec) {
  const double e = 0.5 * g2;
  if (ecutprec == 0.0) {
    if (ekin_n == 0.0) ekin_n = 1.0;
    return e < ekin_n ? 0.5 / ekin_n : 0.5 / e;
  }
  return e < ecutprec ? 0.5 / ecutprec : 0.5 / e;
}

}  // namespace

// ------------------------------------------------------------ band sums
// Preconditioner::update (src/Preconditioner.cpp:88-143) and the terms of
// SlaterDet::dot (src/SlaterDet.cpp:1122-1148): BS_CHUNKS blocks per band
// write partial sums, one thread per band combines them in a fixed order
constexpr int BS_CHUNKS = 8;

__global__ __launch_bounds__(256) void k_band_partials(const double* __restrict__ c, long ldc2,
                                                       const double* __restrict__ d, long ldd2,
                                                       const double* __restrict__ g2, int ngw,
                                                       double* __restrict__ part) {
  const int n = blockIdx.y;
  const double2* cn = reinterpret_cast<const double2*>(c + ldc2 * n);
  const double2* dn = reinterpret_cast<const double2*>(d + ldd2 * n);
  const int per = (ngw + BS_CHUNKS - 1) / BS_CHUNKS;
  const int g0 = blockIdx.x * per, g1 = min(ngw, g0 + per);
  double v[3] = {0.0, 0.0, 0.0};
  for (int g = g0 + threadIdx.x; g < g1; g += 256) {
    const double2 cc = cn[g];
    const double2 dd = dn[g];
    const double p2 = cc.x * cc.x + cc.y * cc.y;
    v[0] += p2;
    v[1] += p2 * g2[g];
    v[2] += cc.x * dd.x + cc.y * dd.y;
  }
  bsum<3>(v);
  if (threadIdx.x == 0) {
    double* o = part + 3 * ((long)n * BS_CHUNKS + blockIdx.x);
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
  }
}

__global__ __launch_bounds__(256) void k_band_finish(const double* __restrict__ part,
                                                     const double* __restrict__ c, long ldc2,
                                                     const double* __restrict__ d, long ldd2,
                                                     int nst, double* __restrict__ ekin,
                                                     double* __restrict__ dotp) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= nst) return;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  for (int k = 0; k < BS_CHUNKS; k++) {
    const double* o = part + 3 * ((long)n * BS_CHUNKS + k);
    v0 += o[0];
    v1 += o[1];
    v2 += o[2];
  }
  const double* cn = c + ldc2 * n;
  const double c0r = cn[0], c0i = cn[1];
  const double bn = 2.0 * (v0 - 0.5 * (c0r * c0r + c0i * c0i));
  const double be = 2.0 * v1;
  ekin[n] = 0.5 * bn > 0.0 ? 0.5 * be / bn : 0.0;
  dotp[n] = 2.0 * v2 - c0r * d[ldd2 * n];
}

int band_partials(int nst) { return 3 * BS_CHUNKS * nst; }

void launch_band_sums(const double* c, long ldc2, const double* d, long ldd2, const double* g2,
                      int ngw, int nst, double* part, double* ekin, double* dotp,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_band_partials, dim3(BS_CHUNKS, nst), dim3(256), 0, s, c, ldc2, d, ldd2,
                     g2, ngw, part);
  hipLaunchKernelGGL(k_band_finish, dim3((nst + 255) / 256), dim3(256), 0, s, part, c, ldc2, d,
                     ldd2, nst, ekin, dotp);
}

__global__ __launch_bounds__(256) void k_sum(const double* __restrict__ v, int n,
                                             double* __restrict__ out) {
  double t[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += 256) t[0] += v[i];
  bsum<1>(t);
  if (threadIdx.x == 0) out[0] = t[0];
}

void launch_sum(const double* v, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, v, n, out);
}

// ------------------------------------------------------------ SD and PSD
__global__ __launch_bounds__(256) void k_psd_update(double* __restrict__ c, long ldc2,
                                                    const double* __restrict__ d, long ldd2,
                                                    const double* __restrict__ g2,
                                                    const double* __restrict__ ekin, int ngw,
                                                    double ecutprec) {
  const int g = blockIdx.x * 256 + threadIdx.x, n = blockIdx.y;
  if (g >= ngw) return;
  const double f = pdiag(ekin[n], g2[g], ecutprec);
  double2* cp = reinterpret_cast<double2*>(c + ldc2 * n) + g;
  const double2 dv = reinterpret_cast<const double2*>(d + ldd2 * n)[g];
  double2 cv = *cp;
  // c -= f d as two rounded operations (src/PSDWavefunctionStepper.cpp:99-105)
  cv.x = __dsub_rn(cv.x, __dmul_rn(f, dv.x));
  cv.y = __dsub_rn(cv.y, __dmul_rn(f, dv.y));
  *cp = cv;
}

void launch_psd_update(double* c, long ldc2, const double* d, long ldd2, const double* g2,
                       const double* ekin, int ngw, int nst, double ecutprec, hipStream_t s) {
  hipLaunchKernelGGL(k_psd_update, dim3((ngw + 255) / 256, nst), dim3(256), 0, s, c, ldc2, d,
                     ldd2, g2, ekin, ngw, ecutprec);
}

__global__ __launch_bounds__(256) void k_sd_update(double* __restrict__ c, long ldc2,
                                                   const double* __restrict__ d, long ldd2,
                                                   int ngw, double alpha) {
  const int g = blockIdx.x * 256 + threadIdx.x, n = blockIdx.y;
  if (g >= ngw) return;
  double2* cp = reinterpret_cast<double2*>(c + ldc2 * n) + g;
  const double2 dv = reinterpret_cast<const double2*>(d + ldd2 * n)[g];
  double2 cv = *cp;
  cv.x = cv.x + (-alpha) * dv.x;
  cv.y = cv.y + (-alpha) * dv.y;
  *cp = cv;
}

void launch_sd_update(double* c, long ldc2, const double* d, long ldd2, int ngw, int nst,
                      double alpha, hipStream_t s) {
  hipLaunchKernelGGL(k_sd_update, dim3((ngw + 255) / 256, nst), dim3(256), 0, s, c, ldc2, d,
                     ldd2, ngw, alpha);
}

// ------------------------------------------------------------------ PSDA
// one block per 1024 coefficients of one band
int psda_partials(int ngw, int nst) { return ((ngw + 1023) / 1024) * nst; }

__global__ __launch_bounds__(256) void k_psda_prec(double* __restrict__ d, long ldd2,
                                                   const double* __restrict__ dl,
                                                   const double* __restrict__ g2,
                                                   const double* __restrict__ ekin, int ngw,
                                                   double ecutprec, double* __restrict__ partial) {
  const int n = blockIdx.y;
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int g = blockIdx.x * 1024 + j * 256 + threadIdx.x;
    if (g < ngw) {
      const double f = pdiag(ekin[n], g2[g], ecutprec);
      double2* dp = reinterpret_cast<double2*>(d + ldd2 * n) + g;
      const double2 lv = reinterpret_cast<const double2*>(dl + 2L * ngw * n)[g];
      const double2 dv = *dp;
      const double f0 = -f * dv.x, f1 = -f * dv.y;
      *dp = double2{f0, f1};
      const double d0 = f0 - lv.x, d1 = f1 - lv.y;
      // G and -G count twice, G = 0 once (src/PSDAWavefunctionStepper.cpp:140-163)
      const double w = g == 0 ? 1.0 : 2.0;
      v[0] += w * (f0 * d0 + f1 * d1);
      v[1] += w * (d0 * d0 + d1 * d1);
    }
  }
  bsum<2>(v);
  if (threadIdx.x == 0) {
    const long b = (long)n * gridDim.x + blockIdx.x;
    partial[2 * b] = v[0];
    partial[2 * b + 1] = v[1];
  }
}

void launch_psda_prec(double* d, long ldd2, const double* dl, const double* g2,
                      const double* ekin, int ngw, int nst, double ecutprec, double* partial,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_psda_prec, dim3((ngw + 1023) / 1024, nst), dim3(256), 0, s, d, ldd2, dl,
                     g2, ekin, ngw, ecutprec, partial);
}

// theta = -a/b limited to [-1 -> 0, 2] (src/PSDAWavefunctionStepper.cpp:172-183);
// 0 on the first update after preprocess, which makes the update c = x + f
__global__ __launch_bounds__(256) void k_psda_theta(const double* __restrict__ partial, int nblk,
                                                    int* __restrict__ extrapolate,
                                                    double* __restrict__ theta) {
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nblk; i += 256) {
    v[0] += partial[2 * i];
    v[1] += partial[2 * i + 1];
  }
  bsum<2>(v);
  if (threadIdx.x == 0) {
    double th = 0.0;
    if (*extrapolate) {
      if (v[1] != 0.0) th = -v[0] / v[1];
      if (th < -1.0) th = 0.0;
      th = fmin(2.0, th);
    }
    *theta = th;
    *extrapolate = 1;
  }
}

void launch_psda_theta(const double* partial, int nblk, int* extrapolate, double* theta,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_psda_theta, dim3(1), dim3(256), 0, s, partial, nblk, extrapolate, theta);
}

__global__ __launch_bounds__(256) void k_psda_update(double* __restrict__ c, long ldc2,
                                                     const double* __restrict__ d, long ldd2,
                                                     double* __restrict__ cl,
                                                     double* __restrict__ dl,
                                                     const double* __restrict__ theta, int ngw) {
  const int g = blockIdx.x * 256 + threadIdx.x, n = blockIdx.y;
  if (g >= ngw) return;
  const double th = *theta;
  double2* cp = reinterpret_cast<double2*>(c + ldc2 * n) + g;
  double2* lp = reinterpret_cast<double2*>(cl + 2L * ngw * n) + g;
  double2* fp = reinterpret_cast<double2*>(dl + 2L * ngw * n) + g;
  const double2 x = *cp, xl = *lp, f = reinterpret_cast<const double2*>(d + ldd2 * n)[g],
                fl = *fp;
  double2 out;
  out.x = (x.x + th * (x.x - xl.x)) + (f.x + th * (f.x - fl.x));
  out.y = (x.y + th * (x.y - xl.y)) + (f.y + th * (f.y - fl.y));
  *cp = out;
  *lp = x;
  *fp = f;
}

void launch_psda_update(double* c, long ldc2, const double* d, long ldd2, double* cl, double* dl,
                        const double* theta, int ngw, int nst, hipStream_t s) {
  hipLaunchKernelGGL(k_psda_update, dim3((ngw + 255) / 256, nst), dim3(256), 0, s, c, ldc2, d,
                     ldd2, cl, dl, theta, ngw);
}

// rows, lddst, ldsrc even: columns of complex values copied as double2
__global__ __launch_bounds__(256) void k_copy_cols_if(double* __restrict__ dst, long lddst,
                                                      const double* __restrict__ src, long ldsrc,
                                                      long rows, const int* __restrict__ info) {
  if (*info) return;
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (2 * r < rows)
    reinterpret_cast<double2*>(dst + lddst * j)[r] =
        reinterpret_cast<const double2*>(src + ldsrc * j)[r];
}

void launch_copy_cols_if(double* dst, long lddst, const double* src, long ldsrc, long rows,
                         int cols, const int* info, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_cols_if, dim3((unsigned)((rows / 2 + 255) / 256), cols), dim3(256), 0,
                     s, dst, lddst, src, ldsrc, rows, info);
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

void launch_set_int(int* p, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, p, v);
}

// ------------------------------------------------------------------ polar
__global__ void k_polar_init(int* flags) {
  if (threadIdx.x < 8) flags[threadIdx.x] = 0;
}

void launch_polar_init(int* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_polar_init, dim3(1), dim3(64), 0, s, flags);
}

// q <- q + I (q held -x^T x), qnrm2 = ||q||_F; the step taken next:
// Bjorck-Bowie if qnrm2 < 1, Higham otherwise; done once qnrm2 <= tol
// (the loop test of src/Matrix.cpp:2458 for the following iteration)
__global__ __launch_bounds__(1024) void k_polar_norm(double* __restrict__ q, int n, double tol,
                                                     int* __restrict__ flags,
                                                     double* __restrict__ qnrm) {
  if (flags[0]) {
    if (threadIdx.x == 0) { flags[1] = 1; flags[2] = 1; }
    return;
  }
  double v[1] = {0.0};
  const long nn = (long)n * n;
  for (long i = threadIdx.x; i < nn; i += 1024) {
    const long col = i / n, row = i - col * n;
    const double x = q[i] + (row == col ? 1.0 : 0.0);
    q[i] = x;
    v[0] += x * x;
  }
  bsum<1>(v);
  if (threadIdx.x == 0) {
    const double qn = sqrt(v[0]);
    *qnrm = qn;
    flags[1] = qn < 1.0 ? 0 : 1;
    flags[2] = qn < 1.0 ? 1 : 0;
    flags[3] += 1;
    flags[0] = qn <= tol ? 1 : 0;
  }
}

void launch_polar_norm(double* q, int n, double tol, int* flags, double* qnrm, hipStream_t s) {
  hipLaunchKernelGGL(k_polar_norm, dim3(1), dim3(1024), 0, s, q, n, tol, flags, qnrm);
}

__global__ __launch_bounds__(256) void k_ident_axpy(double* __restrict__ t,
                                                    const double* __restrict__ q, int n,
                                                    double a, const int* __restrict__ skip) {
  if (*skip) return;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)n * n) return;
  const long col = i / n, row = i - col * n;
  t[i] = (row == col ? 1.0 : 0.0) + a * q[i];
}

void launch_ident_axpy(double* t, const double* q, int n, double a, const int* skip,
                       hipStream_t s) {
  const long nn = (long)n * n;
  hipLaunchKernelGGL(k_ident_axpy, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, t, q, n,
                     a, skip);
}

__global__ __launch_bounds__(256) void k_copy_if_not(double* __restrict__ dst,
                                                     const double* __restrict__ src, long count,
                                                     const int* __restrict__ skip) {
  if (*skip) return;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < count) dst[i] = src[i];
}

void launch_copy_if_not(double* dst, const double* src, long count, const int* skip,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_copy_if_not, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, dst,
                     src, count, skip);
}

// pieces of the Higham step x <- (x + x^-T) / 2 (src/Matrix.cpp:2500-2513),
// with x^-T = x (x^T x)^-1 from the device-wide Cholesky of x^T x:
// g = transpose of a row-major lower inverse (so g^T g = (x^T x)^-1),
// x = (x + z) / 2, and the singular flag from the factorization
__global__ __launch_bounds__(256) void k_transpose_if(double* __restrict__ dst,
                                                      const double* __restrict__ src, int n,
                                                      const int* __restrict__ skip) {
  if (*skip) return;
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int j = ty; j < 32; j += 8)
    if (by + j < n && bx + tx < n) tile[j][tx] = src[(long)(by + j) * n + bx + tx];
  __syncthreads();
  for (int j = ty; j < 32; j += 8)
    if (bx + j < n && by + tx < n) dst[(long)(bx + j) * n + by + tx] = tile[tx][j];
}

void launch_transpose_if(double* dst, const double* src, int n, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(k_transpose_if, dim3((n + 31) / 32, (n + 31) / 32), dim3(256), 0, s, dst, src,
                     n, skip);
}

__global__ __launch_bounds__(256) void k_higham_update(double* __restrict__ x,
                                                       const double* __restrict__ z, long count,
                                                       const int* __restrict__ skip) {
  if (*skip) return;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < count) x[i] = 0.5 * (x[i] + z[i]);
}

void launch_higham_update(double* x, const double* z, int n, const int* skip, hipStream_t s) {
  const long nn = (long)n * n;
  hipLaunchKernelGGL(k_higham_update, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, x, z,
                     nn, skip);
}

__global__ void k_polar_singular(const int* __restrict__ info, int* __restrict__ flags) {
  if (threadIdx.x == 0 && *info > 0) flags[4] = 1;
}

void launch_polar_singular(const int* info, int* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_polar_singular, dim3(1), dim3(64), 0, s, info, flags);
}

// out[0..3] = eigenvalue sum, theta, polar iterations, orthogonalization
// status (Cholesky pivot + 1, or -1 for a singular polar step); out may be
// pinned host memory
__global__ void k_wf_results(const double* __restrict__ scal, const int* __restrict__ flags,
                             const int* __restrict__ info, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  out[0] = scal[0];
  out[1] = scal[1];
  out[2] = flags ? (double)flags[3] : 0.0;
  out[3] = *info ? (double)*info : ((flags && flags[4]) ? -1.0 : 0.0);
}

void launch_wf_results(const double* scal, const int* flags, const int* info, double* out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_wf_results, dim3(1), dim3(64), 0, s, scal, flags, info, out);
}

// ----------------------------------------------------------- charge mixing
__global__ void k_mix_restart(int* st) {
  if (threadIdx.x == 0) { st[0] = 0; st[1] = -1; st[2] = -1; st[3] = 0; }
}

void launch_mix_restart(int* st, hipStream_t s) {
  hipLaunchKernelGGL(k_mix_restart, dim3(1), dim3(64), 0, s, st);
}

__global__ void k_mix_begin(int* st, int nmax, int anderson) {
  if (threadIdx.x != 0) return;
  const int mixed = st[0] > 0 ? 1 : 0;
  st[3] = mixed;
  if (mixed && anderson) {
    st[2] = (st[2] + 1) % (nmax + 1);
    if (st[1] < nmax) st[1]++;
  }
  if (st[0] < (1 << 30)) st[0]++;
}

void launch_mix_begin(int* st, int nmax, int anderson, hipStream_t s) {
  hipLaunchKernelGGL(k_mix_begin, dim3(1), dim3(64), 0, s, st, nmax, anderson);
}

// first call: rhog_in = rhog.  Later: drhog = rhog - rhog_in; Anderson
// stores x = rhog_in, f = drhog / wls in ring slot k; simple mixing
// rhog_in += alpha drhog wk and rhog = rhog_in.  ng complex entries: the
// densities of the spins one after the other, each of nw entries (the
// weights wls, wk, of length nw, repeat per spin)
__global__ __launch_bounds__(256) void k_mix_stage(double* __restrict__ rhog,
                                                   double* __restrict__ rhog_in,
                                                   const double* __restrict__ wls,
                                                   const double* __restrict__ wk,
                                                   double* __restrict__ xs,
                                                   double* __restrict__ fs,
                                                   const int* __restrict__ st, int ng,
                                                   int nw, double alpha, int anderson) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ng) return;
  const int iw = i % nw;
  const double2 r = reinterpret_cast<const double2*>(rhog)[i];
  const double2 ri = reinterpret_cast<const double2*>(rhog_in)[i];
  if (!st[3]) {
    reinterpret_cast<double2*>(rhog_in)[i] = r;
    return;
  }
  const double dx = r.x - ri.x, dy = r.y - ri.y;
  if (anderson) {
    const long m = 2L * ng, k = st[2];
    xs[k * m + 2 * i] = ri.x;
    xs[k * m + 2 * i + 1] = ri.y;
    fs[k * m + 2 * i] = dx / wls[iw];
    fs[k * m + 2 * i + 1] = dy / wls[iw];
  } else {
    const double w = wk[iw];
    const double2 rin = {ri.x + (alpha * dx) * w, ri.y + (alpha * dy) * w};
    reinterpret_cast<double2*>(rhog_in)[i] = rin;
    reinterpret_cast<double2*>(rhog)[i] = rin;
  }
}

void launch_mix_stage(double* rhog, double* rhog_in, const double* wls, const double* wk,
                      double* xs, double* fs, const int* st, int ng, int nw, double alpha,
                      int anderson, hipStream_t s) {
  hipLaunchKernelGGL(k_mix_stage, dim3((ng + 255) / 256), dim3(256), 0, s, rhog, rhog_in, wls,
                     wk, xs, fs, st, ng, nw, alpha, anderson);
}

// Anderson sums a_ij = sum (f_k - f_i)(f_k - f_j) (j <= i), b_i = sum (f_k - f_i) f_k
// (src/AndersonMixer.cpp:76-104): per-block partials of NSUM values
constexpr int MIX_NA = MIX_NMAX * (MIX_NMAX + 1) / 2;
constexpr int MIX_NSUM = MIX_NA + MIX_NMAX;
constexpr int MIX_BLOCKS = 256;
int mix_partials() { return MIX_BLOCKS * MIX_NSUM; }

__global__ __launch_bounds__(256) void k_mix_dots(const double* __restrict__ fs,
                                                  const int* __restrict__ st, int nmax, long m,
                                                  double* __restrict__ partial) {
  if (!st[3]) return;
  const int nv = st[1], k = st[2];
  if (nv <= 0) return;
  const double* fk = fs + (long)k * m;
  const double* fi[MIX_NMAX];
#pragma unroll
  for (int i = 0; i < MIX_NMAX; i++)
    fi[i] = i < nv ? fs + (long)((k - i + nmax) % (nmax + 1)) * m : fk;
  double v[MIX_NSUM];
#pragma unroll
  for (int t = 0; t < MIX_NSUM; t++) v[t] = 0.0;
  for (long l = (long)blockIdx.x * 256 + threadIdx.x; l < m; l += (long)gridDim.x * 256) {
    const double f = fk[l];
    double dl[MIX_NMAX];
#pragma unroll
    for (int i = 0; i < MIX_NMAX; i++) dl[i] = i < nv ? f - fi[i][l] : 0.0;
    int t = 0;
#pragma unroll
    for (int i = 0; i < MIX_NMAX; i++)
#pragma unroll
      for (int j = 0; j <= i; j++) v[t++] += dl[i] * dl[j];
#pragma unroll
    for (int i = 0; i < MIX_NMAX; i++) v[MIX_NA + i] += dl[i] * f;
  }
  bsum<MIX_NSUM>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int t = 0; t < MIX_NSUM; t++) partial[(long)blockIdx.x * MIX_NSUM + t] = v[t];
}

void launch_mix_dots(const double* fs, const int* st, int nmax, long m, double* partial,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_mix_dots, dim3(MIX_BLOCKS), dim3(256), 0, s, fs, st, nmax, m, partial);
}

// theta = sum over eigenpairs (w_k > 0) of (u_k.b / w_k) u_k: the AndersonMixer's
// dsyev solve (src/AndersonMixer.cpp:113-157) with a cyclic Jacobi eigensolver
__global__ __launch_bounds__(64) void k_mix_solve(const double* __restrict__ partial,
                                                  const int* __restrict__ st, int anderson,
                                                  double* __restrict__ theta) {
  if (!st[3] || !anderson) return;
  const int nv = st[1];
  if (nv <= 0) return;
  __shared__ double sums[MIX_NSUM];
  const int t = threadIdx.x;
  if (t < MIX_NSUM) {
    double acc = 0.0;
    for (int b = 0; b < MIX_BLOCKS; b++) acc += partial[(long)b * MIX_NSUM + t];
    sums[t] = acc;
  }
  __syncthreads();
  if (t != 0) return;
  double a[MIX_NMAX * MIX_NMAX], u[MIX_NMAX * MIX_NMAX], w[MIX_NMAX], bb[MIX_NMAX];
  const int n = nv;
  {
    int q = 0;
    for (int i = 0; i < MIX_NMAX; i++)
      for (int j = 0; j <= i; j++, q++)
        if (i < n) { a[i + j * n] = sums[q]; a[j + i * n] = sums[q]; }
    for (int i = 0; i < n; i++) bb[i] = sums[MIX_NA + i];
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) u[i + j * n] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += a[i + j * n] * a[i + j * n];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[p + q * n];
        if (apq == 0.0) continue;
        const double th = (a[q + q * n] - a[p + p * n]) / (2.0 * apq);
        const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double cs = 1.0 / sqrt(tt * tt + 1.0), sn = tt * cs;
        for (int k = 0; k < n; k++) {
          const double akp = a[k + p * n], akq = a[k + q * n];
          a[k + p * n] = cs * akp - sn * akq;
          a[k + q * n] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = a[p + k * n], aqk = a[q + k * n];
          a[p + k * n] = cs * apk - sn * aqk;
          a[q + k * n] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double ukp = u[k + p * n], ukq = u[k + q * n];
          u[k + p * n] = cs * ukp - sn * ukq;
          u[k + q * n] = sn * ukp + cs * ukq;
        }
      }
  }
  double wmax = -1e300;
  for (int k = 0; k < n; k++) {
    w[k] = a[k + k * n];
    wmax = fmax(wmax, w[k]);
  }
  const double eig_ratio = 0.0;
  for (int i = 0; i < n; i++) theta[i] = 0.0;
  for (int k = 0; k < n; k++)
    if (w[k] > eig_ratio * wmax) {
      double ub = 0.0;
      for (int j = 0; j < n; j++) ub += u[j + k * n] * bb[j];
      for (int i = 0; i < n; i++) theta[i] += (ub / w[k]) * u[i + k * n];
    }
}

void launch_mix_solve(const double* partial, const int* st, int anderson, double* theta,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_mix_solve, dim3(1), dim3(64), 0, s, partial, st, anderson, theta);
}

// fbar = f_k - sum theta_i (f_k - f_i), xbar likewise (src/AndersonMixer.cpp:253-268);
// rhog_in = xbar + alpha fbar wls wk (src/BOSampleStepper.cpp:857-863); rhog = rhog_in
__global__ __launch_bounds__(256) void k_mix_apply(double* __restrict__ rhog,
                                                   double* __restrict__ rhog_in,
                                                   const double* __restrict__ wls,
                                                   const double* __restrict__ wk,
                                                   const double* __restrict__ xs,
                                                   const double* __restrict__ fs,
                                                   const int* __restrict__ st,
                                                   const double* __restrict__ theta, int nmax,
                                                   int ng, int nw, double alpha, int anderson) {
  if (!st[3] || !anderson) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ng) return;
  const int iw = i % nw;
  const int nv = st[1], k = st[2];
  const long m = 2L * ng;
  double out[2];
#pragma unroll
  for (int cpt = 0; cpt < 2; cpt++) {
    const long l = 2L * i + cpt;
    const double fk = fs[(long)k * m + l], xk = xs[(long)k * m + l];
    double fb = fk, xb = xk;
    for (int ii = 0; ii < nv; ii++) {
      const long kmi = (k - ii + nmax) % (nmax + 1);
      const double th = theta[ii];
      fb -= th * (fk - fs[kmi * m + l]);
      xb -= th * (xk - xs[kmi * m + l]);
    }
    const double drb = fb * wls[iw];
    out[cpt] = xb + (alpha * drb) * wk[iw];
  }
  const double2 rin = {out[0], out[1]};
  reinterpret_cast<double2*>(rhog_in)[i] = rin;
  reinterpret_cast<double2*>(rhog)[i] = rin;
}

void launch_mix_apply(double* rhog, double* rhog_in, const double* wls, const double* wk,
                      const double* xs, const double* fs, const int* st, const double* theta,
                      int nmax, int ng, int nw, double alpha, int anderson, hipStream_t s) {
  hipLaunchKernelGGL(k_mix_apply, dim3((ng + 255) / 256), dim3(256), 0, s, rhog, rhog_in, wls, wk,
                     xs, fs, st, theta, nmax, ng, nw, alpha, anderson);
}

__global__ __launch_bounds__(256) void k_mix_select(double* __restrict__ rhor,
                                                    const double* __restrict__ rhor_new, long n,
                                                    const int* __restrict__ st) {
  if (!st[3]) return;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) rhor[i] = rhor_new[i];
}

void launch_mix_select(double* rhor, const double* rhor_new, long n, const int* st,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_mix_select, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rhor,
                     rhor_new, n, st);
}

}  // namespace qbg

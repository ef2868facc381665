// kpoint.hip -- the complex (k != 0) wavefunction basis carried on a real
// band grid.
//
// psi(r) = sum_{G in S} c(G) e^{iGr}, S = {G : |k + G| < sqrt(2 ecut)} (the
// Bloch factor e^{ikr} is left out: |psi|^2 and v(r) psi(r) do not see it),
// is a complex function.  Its real and imaginary parts are real functions
// whose coefficients live on the half sphere of a real basis U holding every
// G of S and its -G:
//   c1(G) = (c(G) + conj c(-G)) / 2,   c2(G) = (c(G) - conj c(-G)) / (2i)
// (c = 0 outside S).  The real band grid's pair transform of (c1, c2) forms
// c1 + i c2 = c(G) at +G and conj c1 + i conj c2 = c(-G) at -G, i.e. psi
// itself, one band per transform as the reference's complex FourierTransform
// (src/FourierTransform.cpp:147-171); its forward pair transform of v psi
// returns (h1, h2) with h(G) = h1 + i h2 and h(-G) = conj h1 + i conj h2.
// Both maps are gathers over precomputed index tables (fold: U -> S pairs,
// unfold: S -> U), one coalesced pass over the coefficients each.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include "energy.h"
#include "internal.h"

namespace qbg {

// out[p + ldu*2n] = c1, out[p + ldu*(2n+1)] = c2 for the ngwu vectors p of U;
// fold[2p], fold[2p+1]: the S index of G_p and of -G_p, or -1
__global__ __launch_bounds__(256) void k_fold(const double* __restrict__ c, long ldc, int ngwu,
                                              const int32_t* __restrict__ fold,
                                              double* __restrict__ out, long ldu) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= ngwu) return;
  const int n = blockIdx.y;
  const int2 j = reinterpret_cast<const int2*>(fold)[p];
  const double* cn = c + 2 * ldc * n;
  double2 a = {0.0, 0.0}, b = {0.0, 0.0};
  if (j.x >= 0) a = reinterpret_cast<const double2*>(cn)[j.x];
  if (j.y >= 0) b = reinterpret_cast<const double2*>(cn)[j.y];
  double2* o1 = reinterpret_cast<double2*>(out + 2 * ldu * (2L * n));
  double2* o2 = reinterpret_cast<double2*>(out + 2 * ldu * (2L * n + 1));
  o1[p] = double2{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
  o2[p] = double2{0.5 * (a.y + b.y), 0.5 * (b.x - a.x)};
}

// hpsi[j + ldh*n] += h1 + i h2 (G_j at p = unfold[j]) or conj h1 + i conj h2
// (-G_j at p = -1 - unfold[j]), (h1, h2) = H[. + ldu*2n], H[. + ldu*(2n+1)]
__global__ __launch_bounds__(256) void k_unfold_add(const double* __restrict__ H, long ldu, int ngw,
                                                    const int32_t* __restrict__ unfold,
                                                    double* __restrict__ hpsi, long ldh) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= ngw) return;
  const int n = blockIdx.y;
  const int u = unfold[j];
  const int p = u >= 0 ? u : -1 - u;
  const double2 h1 = reinterpret_cast<const double2*>(H + 2 * ldu * (2L * n))[p];
  const double2 h2 = reinterpret_cast<const double2*>(H + 2 * ldu * (2L * n + 1))[p];
  double2* d = reinterpret_cast<double2*>(hpsi + 2 * ldh * n) + j;
  double2 v = *d;
  if (u >= 0) { v.x += h1.x - h2.y; v.y += h1.y + h2.x; }
  else { v.x += h1.x + h2.y; v.y += h2.x - h1.y; }
  *d = v;
}

void launch_fold(const double* c, long ldc, int nst, int ngwu, const int32_t* fold, double* out,
                 long ldu, hipStream_t s) {
  if (nst <= 0 || ngwu <= 0) return;
  hipLaunchKernelGGL(k_fold, dim3((ngwu + 255) / 256, nst), dim3(256), 0, s, c, ldc, ngwu, fold,
                     out, ldu);
}

// the scalars of a k-point set: S_KIN (8), S_ENL (1), S_NLS (6) = the sums
// over the k-points of their kacc slots [16 ik + {0..7, 8, 9..14}]
__global__ void k_kacc_sum(const double* __restrict__ kacc, int nkp, double* __restrict__ kin,
                           double* __restrict__ enl, double* __restrict__ nls) {
  const int t = threadIdx.x;
  if (t >= 15) return;
  double v = 0.0;
  for (int k = 0; k < nkp; k++) v += kacc[16 * k + t];
  if (t < 8) kin[t] = v;
  else if (t == 8) enl[0] = v;
  else nls[t - 9] = v;
}

void launch_kacc_sum(const double* kacc, int nkp, double* kin, double* enl, double* nls,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_kacc_sum, dim3(1), dim3(64), 0, s, kacc, nkp, kin, enl, nls);
}

// rows [r0, r1) of the ncols columns of a (ld) set to zero: a k-point set's
// FP32 band copy below the largest basis (the split products reduce over
// every padded row)
__global__ __launch_bounds__(256) void k_zero_rows_f32(float* __restrict__ a, long ld, long r0,
                                                       long r1, int ncols) {
  const long nr = r1 - r0;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nr * ncols) return;
  const long col = t / nr;
  a[col * ld + r0 + (t - col * nr)] = 0.0f;
}

void launch_zero_rows_f32(float* a, long ld, long r0, long r1, int ncols, hipStream_t s) {
  if (r1 <= r0 || ncols <= 0) return;
  const long tot = (r1 - r0) * ncols;
  hipLaunchKernelGGL(k_zero_rows_f32, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, a, ld,
                     r0, r1, ncols);
}

void launch_unfold_add(const double* H, long ldu, int nst, int ngw, const int32_t* unfold,
                       double* hpsi, long ldh, hipStream_t s) {
  if (nst <= 0 || ngw <= 0) return;
  hipLaunchKernelGGL(k_unfold_add, dim3((ngw + 255) / 256, nst), dim3(256), 0, s, H, ldu, ngw,
                     unfold, hpsi, ldh);
}

// ------------------------------------------------------------------------
// SlaterDet::gram of the complex basis (src/SlaterDet.cpp:523-532):
// S = C^H C (herk), S = L L^H (potrf, lower), C <- C L^-H (trsm).  With P
// the real view of C (2 ngw x nst) and P' that of i C, Re S = P^T P and
// Im S = -P^T P' (two real products on the FP64 MFMA kernels); the complex
// Cholesky runs column by column (a column kernel and a trailing update per
// column), L^-1 one block per column with the column in LDS, and
// C L^-H = P Re(X) + P' Im(X), X = L^-H, is two real NN products.

constexpr int ZG_NMAX = 2048;   // columns of L^-1 held in LDS (32 KiB)

// out = the real view of i c: (re, im) -> (-im, re)
__global__ __launch_bounds__(256) void k_times_i(const double* __restrict__ c, long ldc2, long ng,
                                                 int n, double* __restrict__ out) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= ng * n) return;
  const int col = (int)(t / ng);
  const long g = t - (long)col * ng;
  const double2 v = reinterpret_cast<const double2*>(c + col * ldc2)[g];
  reinterpret_cast<double2*>(out + col * 2 * ng)[g] = double2{-v.y, v.x};
}

// S[i + j n] = (Sr, Si) (complex, column-major)
__global__ __launch_bounds__(256) void k_zjoin(const double* __restrict__ Sr,
                                               const double* __restrict__ Si, int n,
                                               double* __restrict__ S) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)n * n) return;
  reinterpret_cast<double2*>(S)[t] = double2{Sr[t], Si[t]};
}

// column j of the Cholesky factor: L_jj = sqrt(Re S_jj), L_ij = S_ij / L_jj;
// a non-positive pivot sets info = j + 1 and every later launch does nothing
__global__ __launch_bounds__(256) void k_zchol_col(double* __restrict__ S, int n, int j,
                                                   int* __restrict__ info) {
  if (*info) return;
  __shared__ double d;
  double2* A = reinterpret_cast<double2*>(S);
  if (threadIdx.x == 0) {
    const double v = A[j + (long)j * n].x;
    d = v > 0.0 ? std::sqrt(v) : 0.0;
    A[j + (long)j * n] = double2{d, 0.0};
  }
  __syncthreads();
  if (!(d > 0.0)) {
    if (threadIdx.x == 0) *info = j + 1;
    return;
  }
  const double di = 1.0 / d;
  for (int i = j + 1 + threadIdx.x; i < n; i += 256) {
    double2 v = A[i + (long)j * n];
    A[i + (long)j * n] = double2{v.x * di, v.y * di};
  }
}

// trailing update after column j: S_ik -= L_ij conj(L_kj), n > i >= k > j
__global__ __launch_bounds__(256) void k_zchol_upd(double* __restrict__ S, int n, int j,
                                                   const int* __restrict__ info) {
  if (*info) return;
  const int k = j + 1 + blockIdx.y;
  const int i = k + blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double2* A = reinterpret_cast<double2*>(S);
  const double2 li = A[i + (long)j * n], lk = A[k + (long)j * n];
  double2 v = A[i + (long)k * n];
  v.x -= li.x * lk.x + li.y * lk.y;
  v.y -= li.y * lk.x - li.x * lk.y;
  A[i + (long)k * n] = v;
}

// column j of L^-1 (lower): x_j = 1 / L_jj, x_i = -(sum_{j<=k<i} L_ik x_k) / L_ii,
// written as X = (L^-1)^H split: Xr[m + p n] = Re conj(Linv_pm), Xi = Im conj(Linv_pm)
__global__ __launch_bounds__(256) void k_ztrinv_adj(const double* __restrict__ S, int n,
                                                    const int* __restrict__ info,
                                                    double* __restrict__ Xr,
                                                    double* __restrict__ Xi) {
  if (*info) return;
  __shared__ double2 x[ZG_NMAX];
  __shared__ double2 red[256];
  const double2* L = reinterpret_cast<const double2*>(S);
  const int j = blockIdx.x, tid = threadIdx.x;
  if (tid == 0) x[j] = double2{1.0 / L[j + (long)j * n].x, 0.0};
  __syncthreads();
  for (int i = j + 1; i < n; i++) {
    double2 p = {0.0, 0.0};
    for (int k = j + tid; k < i; k += 256) {
      const double2 a = L[i + (long)k * n], b = x[k];
      p.x += a.x * b.x - a.y * b.y;
      p.y += a.x * b.y + a.y * b.x;
    }
    red[tid] = p;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (tid < st) { red[tid].x += red[tid + st].x; red[tid].y += red[tid + st].y; }
      __syncthreads();
    }
    if (tid == 0) {
      const double dinv = 1.0 / L[i + (long)i * n].x;
      x[i] = double2{-red[0].x * dinv, -red[0].y * dinv};
    }
    __syncthreads();
  }
  // row j of X = conj of column j of L^-1: X[j][i] = conj(Linv_ij) (zero for i < j)
  for (int i = tid; i < n; i += 256) {
    const double2 v = i >= j ? x[i] : double2{0.0, 0.0};
    Xr[j + (long)i * n] = v.x;
    Xi[j + (long)i * n] = -v.y;
  }
}

}  // namespace qbg

extern "C" int qbg_gram_complex(qbg_ctx* c, double* cc, int ldc, int ngw, int nst) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !cc || ngw <= 0 || nst <= 0 || ldc < ngw) return QBG_E_INVAL;
    if (nst > qbg::ZG_NMAX) return c->fail(QBG_E_INVAL, "gram_complex: nst > 2048");
    QBG_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const long K = 2L * ngw, n2 = (long)nst * nst;
    const size_t work = (size_t)64 * n2 + ((size_t)1 << 20);
    double *Sr = nullptr, *Si = nullptr, *S = nullptr, *Xr = nullptr, *Xi = nullptr,
           *Pi = nullptr, *T = nullptr, *W = nullptr;
    int* info = nullptr;
    auto release = [&]() {
      for (void* p : {(void*)Sr, (void*)Si, (void*)S, (void*)Xr, (void*)Xi, (void*)Pi, (void*)T,
                      (void*)W, (void*)info})
        if (p) (void)hipFree(p);
    };
    if (hipMalloc((void**)&Sr, 8 * n2) != hipSuccess || hipMalloc((void**)&Si, 8 * n2) != hipSuccess ||
        hipMalloc((void**)&S, 16 * n2) != hipSuccess || hipMalloc((void**)&Xr, 8 * n2) != hipSuccess ||
        hipMalloc((void**)&Xi, 8 * n2) != hipSuccess ||
        hipMalloc((void**)&Pi, sizeof(double) * K * nst) != hipSuccess ||
        hipMalloc((void**)&T, sizeof(double) * K * nst) != hipSuccess ||
        hipMalloc((void**)&W, sizeof(double) * work) != hipSuccess ||
        hipMalloc((void**)&info, sizeof(int)) != hipSuccess) {
      release();
      return c->fail(QBG_E_NOMEM, "hipMalloc failed (gram_complex workspace)");
    }
    c->prof_begin(QBG_K_GRAM);
    (void)hipMemsetAsync(info, 0, sizeof(int), s);
    const long tot = (long)ngw * nst;
    hipLaunchKernelGGL(qbg::k_times_i, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, cc,
                       2L * ldc, (long)ngw, nst, Pi);
    // Re S = P^T P, Im S = -P^T P'
    if (qbg::gemm_syrk(cc, 2L * ldc, nst, K, 1.0, 0.0, Sr, nst, W, work, s) ||
        qbg::gemm_tn_splitk(cc, 2L * ldc, Pi, K, nst, nst, K, -1.0, 0.0, Si, nst, W, work,
                            QBG_PREC_FP64, s)) {
      (void)hipStreamSynchronize(s);
      release();
      return c->fail(QBG_E_INVAL, "gram_complex workspace");
    }
    hipLaunchKernelGGL(qbg::k_zjoin, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, Sr, Si,
                       nst, S);
    for (int j = 0; j < nst; j++) {
      hipLaunchKernelGGL(qbg::k_zchol_col, dim3(1), dim3(256), 0, s, S, nst, j, info);
      const int m = nst - j - 1;
      if (m > 0)
        hipLaunchKernelGGL(qbg::k_zchol_upd, dim3((unsigned)((m + 255) / 256), (unsigned)m),
                           dim3(256), 0, s, S, nst, j, info);
    }
    c->prof_end();
    // the coefficients are only touched once the factorization succeeded
    int hinfo = 0;
    (void)hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t er = hipStreamSynchronize(s);
    if (er != hipSuccess || hinfo) {
      release();
      if (er != hipSuccess) return c->fail(QBG_E_HIP, hipGetErrorString(er));
      return c->fail(QBG_E_NOT_SPD, "gram_complex: overlap matrix not positive definite");
    }
    c->prof_begin(QBG_K_GRAM);
    hipLaunchKernelGGL(qbg::k_ztrinv_adj, dim3(nst), dim3(256), 0, s, S, nst, info, Xr, Xi);
    // C L^-H = P Re X + P' Im X
    qbg::gemm_nn(cc, 2L * ldc, Xr, nst, K, nst, nst, T, K, 0, QBG_PREC_FP64, s);
    qbg::gemm_nn(Pi, K, Xi, nst, K, nst, nst, T, K, 1, QBG_PREC_FP64, s);
    (void)hipMemcpy2DAsync(cc, sizeof(double) * 2 * ldc, T, sizeof(double) * K, sizeof(double) * K,
                           nst, hipMemcpyDeviceToDevice, s);
    c->prof_end();
    er = hipStreamSynchronize(s);
    release();
    if (er != hipSuccess) return c->fail(QBG_E_HIP, hipGetErrorString(er));
    return QBG_OK;
  });
}

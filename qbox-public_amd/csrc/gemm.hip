// gemm.hip -- FP64 MFMA GEMMs and the small dense factorizations of the
// projector path (NonLocalPotential::energy dgemm calls,
// src/NonLocalPotential.cpp:1411-1428, 1611, 1716) and of SlaterDet::gram
// (syrk/syr/potrf/trsm, src/SlaterDet.cpp:487-520).
//
// v_mfma_f64_16x16x4f64 on a 64-wide wave: lane l holds A[l%16][l/16] and
// B[l/16][l%16]; the 16x16 result is held register-major: accumulator r
// of lane l is D[4*r + l/16][l%16], r<4.
#include <hip/hip_runtime.h>
#include "energy.h"

namespace qbg {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int BT = 64;    // output tile edge
constexpr int KC = 16;    // reduction chunk
constexpr int LP = BT + 2;   // LDS row pitch (doubles)

// ------------------------------------------------------------- TN split-K
// partial[s][n][m] = sum_{k in split s} A[k + m*lda] B[k + n*ldb]
__global__ __launch_bounds__(256) void k_gemm_tn(const double* __restrict__ A, long lda,
                                                 const double* __restrict__ B, long ldb, int M,
                                                 int N, long K, long kper,
                                                 double* __restrict__ partial) {
  __shared__ double As[KC][LP];
  __shared__ double Bs[KC][LP];
  const int m0 = blockIdx.x * BT, n0 = blockIdx.y * BT;
  const long kb = (long)blockIdx.z * kper;
  const long ke = min(K, kb + kper);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = v4d{0, 0, 0, 0};
  // loader: column c = tid/4 of the tile, 4 consecutive k at kq
  const int lc = tid >> 2, kq = (tid & 3) * 4;
  const bool am = (m0 + lc) < M, bn = (n0 + lc) < N;
  const double* Ap = A + (long)(m0 + lc) * lda;
  const double* Bp = B + (long)(n0 + lc) * ldb;
  for (long k0 = kb; k0 < ke; k0 += KC) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const long k = k0 + kq + q;
      As[kq + q][lc] = (am && k < ke) ? Ap[k] : 0.0;
      Bs[kq + q][lc] = (bn && k < ke) ? Bp[k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      const double a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
      const double b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
      acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
      acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
      acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
      acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  double* P = partial + (long)blockIdx.z * M * N;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + i * 16 + 4 * r + (lane >> 4);
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m < M && n < N) P[(long)n * M + m] = acc[i][j][r];
      }
}

// out[m + n*ldo] = alpha * sum_s partial[s][n][m] + gamma * A[m*lda] * B[n*ldb]
__global__ __launch_bounds__(256) void k_splitk_reduce(const double* __restrict__ partial,
                                                       int S, int M, int N, double alpha,
                                                       double gamma, const double* __restrict__ A,
                                                       long lda, const double* __restrict__ B,
                                                       long ldb, double* __restrict__ out,
                                                       long ldo) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)M * N) return;
  const int m = (int)(t % M), n = (int)(t / M);
  double s = 0.0;
  for (int q = 0; q < S; q++) s += partial[(long)q * M * N + t];
  double v = alpha * s;
  if (gamma != 0.0) v += gamma * A[(long)m * lda] * B[(long)n * ldb];
  out[m + (long)n * ldo] = v;
}

int gemm_tn_splitk(const double* A, long lda, const double* B, long ldb, int M, int N, long K,
                   double alpha, double gamma, double* out, long ldo, double* work,
                   size_t work_doubles, hipStream_t s) {
  const int tm = (M + BT - 1) / BT, tn = (N + BT - 1) / BT;
  const long ksteps = (K + KC - 1) / KC;
  int S = (int)std::max<long>(1, std::min<long>(ksteps / 8, (2048 + tm * tn - 1) / (tm * tn)));
  while (S > 1 && (size_t)S * M * N > work_doubles) S--;
  if ((size_t)S * M * N > work_doubles) return -1;
  const long kper = ((ksteps + S - 1) / S) * KC;
  S = (int)((K + kper - 1) / kper);
  hipLaunchKernelGGL(k_gemm_tn, dim3(tm, tn, S), dim3(256), 0, s, A, lda, B, ldb, M, N, K, kper,
                     work);
  const long tot = (long)M * N;
  hipLaunchKernelGGL(k_splitk_reduce, dim3((tot + 255) / 256), dim3(256), 0, s, work, S, M, N,
                     alpha, gamma, A, lda, B, ldb, out, ldo);
  return 0;
}

// ------------------------------------------------------------------- NN
// C[k + n*ldc] (+)= sum_m A[k + m*lda] F[m + n*ldf]   (K rows, N cols, M sum)
__global__ __launch_bounds__(256) void k_gemm_nn(const double* __restrict__ A, long lda,
                                                 const double* __restrict__ F, long ldf, long K,
                                                 int N, int M, double* __restrict__ C, long ldc,
                                                 int accumulate) {
  __shared__ double As[KC][LP];
  __shared__ double Fs[KC][LP];
  const long k0 = (long)blockIdx.x * BT;
  const int n0 = blockIdx.y * BT;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wk = (w & 1) * 32, wn = (w >> 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = v4d{0, 0, 0, 0};
  for (int mb = 0; mb < M; mb += KC) {
    // A tile: KC columns (m) x 64 rows (k), 4 per thread along k
    {
      const int mm = tid >> 4, kq = (tid & 15) * 4;
      const int m = mb + mm;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long k = k0 + kq + q;
        As[mm][kq + q] = (m < M && k < K) ? A[k + (long)m * lda] : 0.0;
      }
    }
    // F tile: KC rows (m) x 64 cols (n)
    {
      const int nn = tid >> 2, mq = (tid & 3) * 4;
      const int n = n0 + nn;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int m = mb + mq + q;
        Fs[mq + q][nn] = (m < M && n < N) ? F[m + (long)n * ldf] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      const double a0 = As[kr][wk + c], a1 = As[kr][wk + 16 + c];
      const double b0 = Fs[kr][wn + c], b1 = Fs[kr][wn + 16 + c];
      acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
      acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
      acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
      acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long k = k0 + wk + i * 16 + 4 * r + (lane >> 4);
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (k < K && n < N) {
          double* p = C + k + (long)n * ldc;
          *p = accumulate ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
}

void gemm_nn(const double* A, long lda, const double* F, long ldf, long K, int N, int M,
             double* C, long ldc, int accumulate, hipStream_t s) {
  dim3 grid((unsigned)((K + BT - 1) / BT), (N + BT - 1) / BT);
  hipLaunchKernelGGL(k_gemm_nn, grid, dim3(256), 0, s, A, lda, F, ldf, K, N, M, C, ldc,
                     accumulate);
}

// ------------------------------------------------------- Cholesky, inverse
// In-place lower Cholesky of the column-major n x n matrix S (right-looking,
// one work-group).  info[0] = 0, or j+1 if pivot j is not positive
// (src/Matrix.cpp:2308-2334 returns info from dpotrf the same way).
__global__ __launch_bounds__(1024) void k_potrf(double* __restrict__ S, int n, int lds,
                                                int* __restrict__ info) {
  __shared__ double piv;
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < n; j++) {
    if (threadIdx.x == 0) {
      const double d = S[j + (long)j * lds];
      if (!(d > 0.0)) { bad = j + 1; piv = 1.0; }
      else piv = sqrt(d);
      S[j + (long)j * lds] = piv;
    }
    __syncthreads();
    if (bad) break;
    const double p = piv;
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) S[i + (long)j * lds] /= p;
    __syncthreads();
    // trailing update of the lower triangle: S[i][c] -= L[i][j] L[c][j], c<=i
    const int m = n - j - 1;
    const long tot = (long)m * (m + 1) / 2;
    for (long t = threadIdx.x; t < tot; t += blockDim.x) {
      // t -> (i, c) with c <= i in the trailing block, row-wise packing
      int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
      while ((long)i * (i + 1) / 2 > t) i--;
      while ((long)(i + 1) * (i + 2) / 2 <= t) i++;
      const int cc = (int)(t - (long)i * (i + 1) / 2);
      const int gi = j + 1 + i, gc = j + 1 + cc;
      S[gi + (long)gc * lds] -= S[gi + (long)j * lds] * S[gc + (long)j * lds];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) info[0] = bad;
}

// X = L^{-1} written row-major (X[i*n + c]) by forward substitution on the
// identity, one work-group.  Row-major X is L^{-T} in column-major order.
__global__ __launch_bounds__(1024) void k_trtri(const double* __restrict__ L, int n, int ldl,
                                                double* __restrict__ X) {
  for (long t = threadIdx.x; t < (long)n * n; t += blockDim.x)
    X[t] = ((t / n) == (t % n)) ? 1.0 : 0.0;
  __syncthreads();
  for (int j = 0; j < n; j++) {
    const double d = L[j + (long)j * ldl];
    for (int c = threadIdx.x; c <= j; c += blockDim.x) X[(long)j * n + c] /= d;
    __syncthreads();
    // rows i > j: X[i][c] -= L[i][j] X[j][c], c <= j
    const int rows = n - j - 1, cols = j + 1;
    for (long t = threadIdx.x; t < (long)rows * cols; t += blockDim.x) {
      const int i = j + 1 + (int)(t / cols), c = (int)(t % cols);
      X[(long)i * n + c] -= L[i + (long)j * ldl] * X[(long)j * n + c];
    }
    __syncthreads();
  }
}

void potrf(double* S, int n, int lds, int* info, hipStream_t s) {
  hipLaunchKernelGGL(k_potrf, dim3(1), dim3(1024), 0, s, S, n, lds, info);
}
void trtri(const double* L, int n, int ldl, double* X, hipStream_t s) {
  hipLaunchKernelGGL(k_trtri, dim3(1), dim3(1024), 0, s, L, n, ldl, X);
}

// zero Im c(G=0) of every band (SlaterDet::cleanup, src/SlaterDet.cpp:1335-1356)
__global__ void k_cleanup(double* c, long ldc, int nst) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n < nst) c[2 * ldc * n + 1] = 0.0;
}
void cleanup(double* c, long ldc, int nst, hipStream_t s) {
  hipLaunchKernelGGL(k_cleanup, dim3((nst + 255) / 256), dim3(256), 0, s, c, ldc, nst);
}

}  // namespace qbg

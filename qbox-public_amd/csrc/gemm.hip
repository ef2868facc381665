// gemm.hip -- MFMA GEMMs and the small dense factorizations of the
// projector path (NonLocalPotential::energy dgemm calls,
// src/NonLocalPotential.cpp:1411-1428, 1611, 1716) and of SlaterDet::gram
// (syrk/syr/potrf/trsm, src/SlaterDet.cpp:487-520).
//
// One kernel shape serves every product: D[p][q] = sum_r P(r,p) Q(r,q) on a
// 128x128 (p x q) tile per block of 256 threads (2x2 waves of 64x64, each a
// 4x4 grid of 16x16 MFMA blocks), the reduction staged through LDS in
// chunks of KC with the next chunk prefetched into registers.  P is read
// along r (P(r,p) = P[r + p*ldp]); Q along r (TN form, Q[r + q*ldq]) or
// along q (NN form, Q[q + r*ldq]).  q is the contiguous index of the
// output, so the 16 lanes of an MFMA column group store 16 consecutive
// elements.  The compute type is FP64 (v_mfma_f64_16x16x4f64) or FP32
// (v_mfma_f32_16x16x4f32); operands are converted while they are staged.
//
// MFMA lane layouts on a 64-wide wave (16x16x4): lane l holds A[l%16][l/16]
// and B[l/16][l%16]; accumulator e of lane l is D[4e + l/16][l%16] for f64
// and D[4(l/16) + e][l%16] for f32.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "energy.h"
#include "qbg.h"

namespace qbg {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename TC>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int e, int lane) { return 4 * e + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef v4f acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int e, int lane) { return 4 * (lane >> 4) + e; }
};

constexpr int TP = 128;       // output tile rows (p)
constexpr int TQ = 128;       // output tile columns (q)
constexpr int KC = 16;        // reduction chunk
constexpr int PAD = 16;       // LDS row padding (elements)

struct GemmArgs {
  const void* P;
  long ldp;
  const void* Q;
  long ldq;
  int np, nq;       // output extent
  long nr;          // reduction extent
  long rper;        // reduction length of one split (multiple of KC)
  int tp, tq;       // tiles along p and q
  int pfast;        // block order: p tiles fastest (1) or q tiles fastest (0)
  void* out;        // out[split*sstride + q + p*ldo]
  long ldo, sstride;
  int accumulate;
  int tri;          // 1: only tiles with q >= p (symmetric product);
                    // 2: P(r,p) = 0 for r > p (reduction stops at the tile's last p)
  const int* skip;  // if set and *skip != 0 the launch does nothing (a step of
                    // a captured iteration that has already converged)
  int nblk;         // tiles x splits; the grid is that rounded up to 8
};

template <typename TPm, typename TQm, typename TC, bool QR, typename TO>
__global__ __launch_bounds__(256, 2) void k_gemm(GemmArgs g) {
  typedef Mfma<TC> MF;
  __shared__ TC Ps[KC][TP + PAD];
  __shared__ TC Qs[KC][TQ + PAD];
  if (g.skip && *g.skip) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // blocks are dealt to the 8 XCDs round robin: renumber so that runs of
  // consecutive tiles (which share operand panels) stay on one XCD's L2; the
  // grid is nblk rounded up to a multiple of 8, the blocks past nblk return
  const int b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (b >= g.nblk) return;
  // tri == 1 (symmetric, tp == tq): only the tiles q >= p, row by row
  const int ntile = g.tri == 1 ? g.tp * (g.tp + 1) / 2 : g.tp * g.tq;
  const int split = b / ntile, tt = b - split * ntile;
  int ip, iq;
  if (g.tri == 1) {
    int r = tt;
    ip = 0;
    while (r >= g.tp - ip) { r -= g.tp - ip; ip++; }
    iq = ip + r;
  } else if (g.pfast) { ip = tt % g.tp; iq = tt / g.tp; }
  else { iq = tt % g.tq; ip = tt / g.tq; }
  const int p0 = ip * TP, q0 = iq * TQ;
  const long rb = (long)split * g.rper;
  long re = min(g.nr, rb + g.rper);
  if (g.tri == 2) re = min(re, (long)p0 + TP);
  const TPm* __restrict__ Pp = (const TPm*)g.P;
  const TQm* __restrict__ Qp = (const TQm*)g.Q;
  TC pv[8], qv[8];
  auto load = [&](long r0) {
    const long r = r0 + (tid & 15);
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int p = p0 + (tid >> 4) + 16 * j;
      pv[j] = (p < g.np && r < re) ? (TC)Pp[r + (long)p * g.ldp] : TC(0);
    }
    if (QR) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int q = q0 + (tid >> 4) + 16 * j;
        qv[j] = (q < g.nq && r < re) ? (TC)Qp[r + (long)q * g.ldq] : TC(0);
      }
    } else {
      const int q = q0 + (tid & 127);
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const long rr = r0 + (tid >> 7) + 2 * j;
        qv[j] = (q < g.nq && rr < re) ? (TC)Qp[q + rr * g.ldq] : TC(0);
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int j = 0; j < 8; j++) Ps[tid & 15][(tid >> 4) + 16 * j] = pv[j];
    if (QR) {
#pragma unroll
      for (int j = 0; j < 8; j++) Qs[tid & 15][(tid >> 4) + 16 * j] = qv[j];
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++) Qs[(tid >> 7) + 2 * j][tid & 127] = qv[j];
    }
  };
  const int wp = (w & 1) * 64, wq = (w >> 1) * 64;
  typename MF::acc acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = typename MF::acc{0, 0, 0, 0};
  if (rb < re) load(rb);
  for (long r0 = rb; r0 < re; r0 += KC) {
    store();
    __syncthreads();
    if (r0 + KC < re) load(r0 + KC);
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      TC a[4], bq[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = Ps[kr][wp + 16 * i + c];
#pragma unroll
      for (int j = 0; j < 4; j++) bq[j] = Qs[kr][wq + 16 * j + c];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = MF::mma(a[i], bq[j], acc[i][j]);
    }
    __syncthreads();
  }
  TO* o = (TO*)g.out + (long)split * g.sstride;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int p = p0 + wp + 16 * i + MF::row(e, lane);
        const int q = q0 + wq + 16 * j + (lane & 15);
        if (p < g.np && q < g.nq) {
          TO* d = o + q + (long)p * g.ldo;
          const TO v = (TO)acc[i][j][e];
          *d = g.accumulate ? (TO)(*d + v) : v;
        }
      }
}

// out[q + p*ldo] = alpha * sum_s partial[s][p][q] + gamma * Q[q*ldq] * P[p*ldp]
template <typename TQm, typename TPm>
__global__ __launch_bounds__(256) void k_split_reduce(const double* __restrict__ partial, int S,
                                                      int nq, int np, double alpha, double gamma,
                                                      const TQm* __restrict__ Q, long ldq,
                                                      const TPm* __restrict__ P, long ldp,
                                                      double* __restrict__ out, long ldo,
                                                      int sym, const int* __restrict__ skip) {
  if (skip && *skip) return;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nq * np) return;
  const int q = (int)(t % nq), p = (int)(t / nq);
  // symmetric products hold only q >= p: the upper part mirrors it
  const long src = (sym && q < p) ? (long)q * nq + p : t;
  double s = 0.0;
  for (int k = 0; k < S; k++) s += partial[(long)k * nq * np + src];
  double v = alpha * s;
  if (gamma != 0.0) v += gamma * (double)Q[(long)q * ldq] * (double)P[(long)p * ldp];
  out[q + (long)p * ldo] = v;
}

namespace {

// compute units of the current device, cached per device
int cu_count() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cache[dev]) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 256;
    cache[dev] = std::max(1, ncu);
  }
  return cache[dev];
}

// the split count in [ceil(smax/2), smax] whose ntile*S blocks spread most
// evenly over the CUs (blocks / (cus * rounds)); the smaller on ties
long even_split(long smax, int ntile) {
  if (smax <= 1) return smax;
  const long cus = cu_count();
  long best = smax;
  double eff = 0.0;
  for (long S = (smax + 1) / 2; S <= smax; S++) {
    const long b = S * ntile;
    const double e = (double)b / (double)(cus * ((b + cus - 1) / cus));
    if (e > eff + 1e-9) { eff = e; best = S; }
  }
  return best;
}

template <typename TPm, typename TQm, typename TC, bool QR, typename TO>
void launch_gemm(GemmArgs g, hipStream_t s) {
  const int ntile = g.tri == 1 ? g.tp * (g.tp + 1) / 2 : g.tp * g.tq;
  g.nblk = ntile * (int)((g.nr + g.rper - 1) / g.rper);
  const unsigned grid = (unsigned)((g.nblk + 7) & ~7);
  hipLaunchKernelGGL((k_gemm<TPm, TQm, TC, QR, TO>), dim3(grid), dim3(256), 0, s, g);
}

// split-K TN product; partials in FP64, reduced in a fixed order.  sym: the
// product is symmetric (A = B, M = N): only the lower tiles are formed.
template <typename TQm>
int tn_splitk(const TQm* A, long lda, const double* B, long ldb, int M, int N, long K,
              double alpha, double gamma, double* out, long ldo, double* work,
              size_t work_doubles, int prec, int sym, hipStream_t s, const int* skip) {
  if (M <= 0 || N <= 0 || K <= 0) return -1;
  GemmArgs g{};
  g.skip = skip;
  g.P = B; g.ldp = ldb; g.Q = A; g.ldq = lda;
  g.np = N; g.nq = M; g.nr = K;
  g.tp = (N + TP - 1) / TP; g.tq = (M + TQ - 1) / TQ;
  g.pfast = 0;
  g.tri = sym ? 1 : 0;
  const long ksteps = (K + KC - 1) / KC;
  // tiles that do work: the upper triangle (q >= p) of a symmetric product
  const int ntile = sym && g.tp == g.tq ? g.tp * (g.tp + 1) / 2 : g.tp * g.tq;
  // about four blocks per CU; for large outputs the partial traffic is kept
  // well below the operand traffic (small outputs split freely); at most 32
  // splits for a general product (64 measured no faster on 512 x 512
  // outputs); a symmetric product takes the count in [smax/2, smax] whose
  // blocks spread most evenly over the CUs (ten 128-tiles at 512 bands: 51)
  long S = std::max<long>(1, (1024 + ntile - 1) / ntile);
  S = std::min<long>(S, std::max<long>(1, ksteps / 4));
  if ((long)M * N > (1L << 20))
    S = std::min<long>(S, std::max<long>(1, (long)((double)K * (M + N) / (4.0 * M * N))));
  if (!sym) S = std::min<long>(S, 32);
  while (S > 1 && (size_t)S * M * N > work_doubles) S--;
  if ((size_t)S * M * N > work_doubles) return -1;
  if (sym) S = even_split(S, ntile);
  g.rper = ((ksteps + S - 1) / S) * KC;
  S = (K + g.rper - 1) / g.rper;
  g.out = work; g.ldo = M; g.sstride = (long)M * N; g.accumulate = 0;
  if (prec == QBG_PREC_FP64) launch_gemm<double, TQm, double, true, double>(g, s);
  else launch_gemm<double, TQm, float, true, double>(g, s);
  const long tot = (long)M * N;
  hipLaunchKernelGGL((k_split_reduce<TQm, double>), dim3((unsigned)((tot + 255) / 256)),
                     dim3(256), 0, s, work, (int)S, M, N, alpha, gamma, A, lda, B, ldb, out, ldo,
                     sym, skip);
  return 0;
}

template <typename TQm, typename TO>
void nn(const TQm* A, long lda, const double* F, long ldf, long K, int N, int M, TO* C, long ldc,
        int accumulate, int prec, int tri, hipStream_t s, const int* skip) {
  if (K <= 0 || N <= 0 || M <= 0) return;
  GemmArgs g{};
  g.skip = skip;
  g.P = F; g.ldp = ldf; g.Q = A; g.ldq = lda;
  g.np = N; g.nq = (int)K; g.nr = M;
  g.tp = (N + TP - 1) / TP; g.tq = (int)((K + TQ - 1) / TQ);
  g.pfast = 1;
  g.tri = tri ? 2 : 0;
  g.rper = ((M + KC - 1) / KC) * KC;
  g.out = C; g.ldo = ldc; g.sstride = 0; g.accumulate = accumulate;
  if (prec == QBG_PREC_FP64) launch_gemm<double, TQm, double, false, TO>(g, s);
  else launch_gemm<double, TQm, float, false, TO>(g, s);
}

}  // namespace

// NN with the reduction (M) split over blocks, partials in work, for products
// whose output has too few 128x128 tiles to fill the device:
// C[k + n*ldc] = sum_m A[k + m*lda] F[m + n*ldf]
int gemm_nn_splitk(const double* A, long lda, const double* F, long ldf, long K, int N, int M,
                   double* C, long ldc, double* work, size_t work_doubles, hipStream_t s,
                   const int* skip) {
  if (K <= 0 || N <= 0 || M <= 0) return -1;
  GemmArgs g{};
  g.skip = skip;
  g.P = F; g.ldp = ldf; g.Q = A; g.ldq = lda;
  g.np = N; g.nq = (int)K; g.nr = M;
  g.tp = (N + TP - 1) / TP; g.tq = (int)((K + TQ - 1) / TQ);
  g.pfast = 1;
  g.tri = 0;
  const long msteps = (M + KC - 1) / KC;
  const int ntile = g.tp * g.tq;
  long S = std::max<long>(1, (1024 + ntile - 1) / ntile);
  S = std::min<long>(S, std::max<long>(1, msteps / 4));
  while (S > 1 && (size_t)S * K * N > work_doubles) S--;
  if ((size_t)S * K * N > work_doubles) return -1;
  g.rper = ((msteps + S - 1) / S) * KC;
  S = (M + g.rper - 1) / g.rper;
  g.out = work; g.ldo = K; g.sstride = K * N; g.accumulate = 0;
  launch_gemm<double, double, double, false, double>(g, s);
  const long tot = K * N;
  hipLaunchKernelGGL((k_split_reduce<double, double>), dim3((unsigned)((tot + 255) / 256)),
                     dim3(256), 0, s, work, (int)S, (int)K, N, 1.0, 0.0, A, lda, F, ldf, C, ldc,
                     0, skip);
  return 0;
}

int gemm_tn_splitk(const double* A, long lda, const double* B, long ldb, int M, int N, long K,
                   double alpha, double gamma, double* out, long ldo, double* work,
                   size_t work_doubles, int prec, hipStream_t s, const int* skip) {
  return tn_splitk(A, lda, B, ldb, M, N, K, alpha, gamma, out, ldo, work, work_doubles, prec, 0,
                   s, skip);
}
int gemm_tn_splitk(const float* A, long lda, const double* B, long ldb, int M, int N, long K,
                   double alpha, double gamma, double* out, long ldo, double* work,
                   size_t work_doubles, int prec, hipStream_t s) {
  return tn_splitk(A, lda, B, ldb, M, N, K, alpha, gamma, out, ldo, work, work_doubles, prec, 0,
                   s, nullptr);
}
int gemm_syrk(const double* A, long lda, int N, long K, double alpha, double gamma, double* out,
              long ldo, double* work, size_t work_doubles, hipStream_t s, const int* skip) {
  return tn_splitk(A, lda, A, lda, N, N, K, alpha, gamma, out, ldo, work, work_doubles,
                   QBG_PREC_FP64, 1, s, skip);
}
int gemm_nn_tile_p() { return TP; }
void gemm_nn_triu(const double* A, long lda, const double* F, long ldf, long K, int N, double* C,
                  long ldc, hipStream_t s) {
  nn(A, lda, F, ldf, K, N, N, C, ldc, 0, QBG_PREC_FP64, 1, s, nullptr);
}
void gemm_nn(const double* A, long lda, const double* F, long ldf, long K, int N, int M,
             double* C, long ldc, int accumulate, int prec, hipStream_t s, const int* skip) {
  nn(A, lda, F, ldf, K, N, M, C, ldc, accumulate, prec, 0, s, skip);
}
void gemm_nn(const float* A, long lda, const double* F, long ldf, long K, int N, int M,
             double* C, long ldc, int accumulate, int prec, hipStream_t s) {
  nn(A, lda, F, ldf, K, N, M, C, ldc, accumulate, prec, 0, s, nullptr);
}
void gemm_nn(const double* A, long lda, const double* F, long ldf, long K, int N, int M,
             float* C, long ldc, int accumulate, int prec, hipStream_t s) {
  nn(A, lda, F, ldf, K, N, M, C, ldc, accumulate, prec, 0, s, nullptr);
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

// Blocked lower Cholesky in one work-group: panels of PB columns.  Per
// panel the PB x PB diagonal block is factored in LDS, the rows below are
// solved against it (one thread per row) and kept in LDS, and the trailing
// lower triangle is updated from LDS.  info as for k_potrf.
constexpr int PB = 8;
constexpr int POTRF_NMAX = 960;    // rows of the LDS panel

__global__ __launch_bounds__(1024) void k_potrf_blk(double* __restrict__ S, int n, int lds,
                                                    int* __restrict__ info) {
  __shared__ double D[PB][PB + 1];
  __shared__ double P[POTRF_NMAX * PB];
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int j0 = 0; j0 < n; j0 += PB) {
    const int nb = min(PB, n - j0), j1 = j0 + nb, rows = n - j1;
    if (tid < nb * nb) D[tid / nb][tid % nb] = S[(j0 + tid / nb) + (long)(j0 + tid % nb) * lds];
    __syncthreads();
    if (tid == 0) {                        // unblocked factor of the diagonal block
      for (int c = 0; c < nb && !bad; c++) {
        double d = D[c][c];
        for (int t = 0; t < c; t++) d -= D[c][t] * D[c][t];
        if (!(d > 0.0)) { bad = j0 + c + 1; break; }
        d = sqrt(d);
        D[c][c] = d;
        for (int r = c + 1; r < nb; r++) {
          double v = D[r][c];
          for (int t = 0; t < c; t++) v -= D[r][t] * D[c][t];
          D[r][c] = v / d;
        }
      }
    }
    __syncthreads();
    if (bad) break;
    if (tid < nb * nb && tid / nb >= tid % nb)
      S[(j0 + tid / nb) + (long)(j0 + tid % nb) * lds] = D[tid / nb][tid % nb];
    // rows below: L_i = A_i D^-T (forward substitution)
    for (int i = tid; i < rows; i += blockDim.x) {
      double x[PB];
#pragma unroll
      for (int c = 0; c < PB; c++) {
        if (c < nb) {
          double v = S[(j1 + i) + (long)(j0 + c) * lds];
          for (int t = 0; t < c; t++) v -= x[t] * D[c][t];
          x[c] = v / D[c][c];
          S[(j1 + i) + (long)(j0 + c) * lds] = x[c];
          P[i * PB + c] = x[c];
        }
      }
    }
    __syncthreads();
    // trailing lower triangle: A_ik -= L_i . L_k, k <= i
    for (long t = tid; t < (long)rows * rows; t += blockDim.x) {
      const int i = (int)(t / rows), k = (int)(t - (long)i * rows);
      if (k <= i) {
        double v = 0.0;
        for (int c = 0; c < nb; c++) v += P[i * PB + c] * P[k * PB + c];
        S[(j1 + i) + (long)(j1 + k) * lds] -= v;
      }
    }
    __syncthreads();
  }
  if (tid == 0) info[0] = bad;
}

// Right-looking blocked Cholesky over the whole device, panels of CB
// columns: the diagonal block factored in LDS by one work-group, the rows
// below solved against it (one row per thread), and the trailing matrix
// updated by an MFMA GEMM with the panel's negated transpose kept in ws.
// info as for k_potrf; after a failure the remaining launches do nothing.
constexpr int CB = 64;

// info = 0, or -1 when *skip (every later launch of the factorization then
// does nothing)
__global__ void k_init_info(int* p, const int* skip) {
  if (threadIdx.x == 0) *p = (skip && *skip) ? -1 : 0;
}

__global__ __launch_bounds__(256) void k_chol_diag(double* __restrict__ S, int lds, int j0,
                                                   int jb, int* __restrict__ info) {
  if (*info) return;
  __shared__ double D[CB][CB + 1];
  __shared__ int bad;
  const int tid = threadIdx.x;
  for (int e = tid; e < jb * jb; e += 256) D[e % jb][e / jb] = S[(j0 + e % jb) + (long)(j0 + e / jb) * lds];
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int c = 0; c < jb; c++) {
    const double d = D[c][c];
    if (!(d > 0.0)) {                      // uniform: every thread read D[c][c]
      if (tid == 0) *info = j0 + c + 1;
      return;
    }
    const double sd = sqrt(d);
    __syncthreads();
    if (tid == 0) D[c][c] = sd;
    for (int r = c + 1 + tid; r < jb; r += 256) D[r][c] /= sd;
    __syncthreads();
    // D[r][k] -= D[r][c] D[k][c], c < k <= r
    const int m = jb - c - 1;
    for (int e = tid; e < m * m; e += 256) {
      const int r = c + 1 + e / m, k = c + 1 + e % m;
      if (k <= r) D[r][k] -= D[r][c] * D[k][c];
    }
    __syncthreads();
  }
  for (int e = tid; e < jb * jb; e += 256) {
    const int r = e % jb, k = e / jb;
    if (k <= r) S[(j0 + r) + (long)(j0 + k) * lds] = D[r][k];
  }
}

__global__ __launch_bounds__(64) void k_chol_panel(double* __restrict__ S, int n, int lds, int j0,
                                                   int jb, const int* __restrict__ info,
                                                   double* __restrict__ lt) {
  if (*info) return;
  __shared__ double D[CB][CB + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < jb * jb; e += 64) D[e % jb][e / jb] = S[(j0 + e % jb) + (long)(j0 + e / jb) * lds];
  __syncthreads();
  const int j1 = j0 + jb, i = blockIdx.x * 64 + tid;
  if (j1 + i >= n) return;
  double x[CB];
#pragma unroll
  for (int c = 0; c < CB; c++) {
    if (c < jb) {
      double v = S[(j1 + i) + (long)(j0 + c) * lds];
#pragma unroll
      for (int t = 0; t < c; t++) v -= x[t] * D[c][t];
      x[c] = v / D[c][c];
      S[(j1 + i) + (long)(j0 + c) * lds] = x[c];
      lt[c + (long)i * CB] = -x[c];
    }
  }
}

static void potrf_blocked(double* S, int n, int lds, int* info, double* ws, hipStream_t s,
                          const int* skip) {
  hipLaunchKernelGGL(k_init_info, dim3(1), dim3(64), 0, s, info, skip);
  for (int j0 = 0; j0 < n; j0 += CB) {
    const int jb = std::min(CB, n - j0), rows = n - j0 - jb;
    hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(256), 0, s, S, lds, j0, jb, info);
    if (rows <= 0) break;
    hipLaunchKernelGGL(k_chol_panel, dim3((rows + 63) / 64), dim3(64), 0, s, S, n, lds, j0, jb,
                       info, ws);
    // A22 += L21 (-L21^T)
    nn(S + (j0 + jb) + (long)j0 * lds, lds, ws, CB, rows, rows, jb, S + (j0 + jb) +
       (long)(j0 + jb) * lds, lds, 1, QBG_PREC_FP64, 0, s, info);
  }
}

void potrf_device(double* S, int n, int lds, int* info, double* ws, hipStream_t s,
                  const int* skip) {
  potrf_blocked(S, n, lds, info, ws, s, skip);
}

// X = L^{-1} (lower), written row-major (X[i*n + c], upper part zero):
// block column J of TB columns per work-group, X_IJ = L_II^{-1} (E_IJ -
// sum_{J <= K < I} L_IK X_KJ) for I = J, J+1, ... with 32 x 32 tiles staged
// in LDS; one wave solves the diagonal-block system column by column.
constexpr int TB = 32;
__global__ __launch_bounds__(256) void k_trtri_blk(const double* __restrict__ L, int n, int ldl,
                                                   double* __restrict__ X,
                                                   const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double R[TB][TB + 1];
  __shared__ double La[TB][TB + 1];
  __shared__ double Xb[TB][TB + 1];
  const int tid = threadIdx.x, J = blockIdx.x, nbk = (n + TB - 1) / TB;
  const int c0 = J * TB;
  // upper blocks of this block column are zero
  for (long t = tid; t < (long)c0 * TB; t += 256) {
    const int r = (int)(t / TB), c = c0 + (int)(t % TB);
    if (c < n) X[(long)r * n + c] = 0.0;
  }
  for (int I = J; I < nbk; I++) {
    const int r0 = I * TB;
    for (int e = tid; e < TB * TB; e += 256) {
      const int r = e / TB, c = e % TB;
      R[r][c] = (I == J && r == c) ? 1.0 : 0.0;
    }
    for (int K = J; K < I; K++) {
      const int k0 = K * TB;
      __syncthreads();
      for (int e = tid; e < TB * TB; e += 256) {
        const int r = e / TB, c = e % TB;
        La[r][c] = (r0 + r < n && k0 + c < n) ? L[(r0 + r) + (long)(k0 + c) * ldl] : 0.0;
        Xb[r][c] = (k0 + r < n && c0 + c < n) ? X[(long)(k0 + r) * n + c0 + c] : 0.0;
      }
      __syncthreads();
      for (int e = tid; e < TB * TB; e += 256) {
        const int r = e / TB, c = e % TB;
        double v = 0.0;
#pragma unroll 8
        for (int t = 0; t < TB; t++) v += La[r][t] * Xb[t][c];
        R[r][c] -= v;
      }
    }
    __syncthreads();
    for (int e = tid; e < TB * TB; e += 256) {
      const int r = e / TB, c = e % TB;
      La[r][c] = (r0 + r < n && r0 + c < n) ? L[(r0 + r) + (long)(r0 + c) * ldl]
                                            : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    if (tid < TB) {                       // column tid: forward substitution
      const int c = tid;
      for (int r = 0; r < TB; r++) {
        double v = R[r][c];
        for (int t = 0; t < r; t++) v -= La[r][t] * R[t][c];
        R[r][c] = v / La[r][r];
      }
    }
    __syncthreads();
    for (int e = tid; e < TB * TB; e += 256) {
      const int r = e / TB, c = e % TB;
      if (r0 + r < n && c0 + c < n) X[(long)(r0 + r) * n + c0 + c] = R[r][c];
    }
    __syncthreads();                      // X of this step is read in the next
  }
}

void potrf(double* S, int n, int lds, int* info, hipStream_t s, double* ws) {
  if (ws && n > 2 * CB) {
    potrf_blocked(S, n, lds, info, ws, s, nullptr);
    return;
  }
  if (n <= POTRF_NMAX)
    hipLaunchKernelGGL(k_potrf_blk, dim3(1), dim3(1024), 0, s, S, n, lds, info);
  else
    hipLaunchKernelGGL(k_potrf, dim3(1), dim3(1024), 0, s, S, n, lds, info);
}
void trtri(const double* L, int n, int ldl, double* X, hipStream_t s, const int* skip) {
  hipLaunchKernelGGL(k_trtri_blk, dim3((n + TB - 1) / TB), dim3(256), 0, s, L, n, ldl, X, skip);
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

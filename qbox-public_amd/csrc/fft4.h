// fft4.h -- four-step register FFT for the pass kernels (device templates).
//
// A line of length N = N1*N2 is transformed by N2 threads (phase A: an
// N1-point DFT over the stride-N2 elements, held in registers, then the
// twiddle W_N^{n2 k1}) and N1 threads (phase B: an N2-point DFT), with one
// exchange through LDS in between.  The small DFTs are fully unrolled with
// compile-time constants (twiddles.h), so a transform touches LDS once per
// element instead of once per radix stage.
#pragma once
#include <hip/hip_runtime.h>
#include "internal.h"
#include "twiddles.h"

namespace qbg {
namespace f4 {

// factor split chosen for a line length: the divisor pair n1 >= n2 with the
// smallest n1 (both <= 32); n1 = 0 if none
__host__ __device__ constexpr int split_n1(int n) {
  for (int a = 1; a <= 32; a++)
    if (n % a == 0 && n / a <= a && n / a >= 1) return a;
  return 0;
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
// x with a x = 1 (mod m), m > 1, gcd(a, m) = 1
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; x++)
    if ((a * x) % m == 1) return x;
  return 1;
}

// Index maps of an N = N1 N2 line split into phase A (N1-point DFTs, thread
// n2) and phase B (N2-point DFTs, thread k1).  For coprime N1, N2 the
// Good-Thomas (prime factor) maps make the split exact without twiddles:
// input n = (N2 n1 + N1 n2) mod N, output k = CRT(k1, k2) = (k1 E1 + k2 E2)
// mod N with E1 = 1 (mod N1), 0 (mod N2) and E2 = 0 (mod N1), 1 (mod N2).
// Otherwise the four-step maps n = N2 n1 + n2, k = k1 + N1 k2 with the
// twiddle W_N^{n2 k1} between the phases.  A forward transform reads its
// input in the output order of the backward one (out_b) and writes in the
// input order (in_a, with n1 the N1-DFT output and n2 the N2-DFT output).
template <int N1, int N2>
struct Map {
  static constexpr int N = N1 * N2;
  static constexpr bool pfa = N1 > 1 && N2 > 1 && gcd_c(N1, N2) == 1;
  static constexpr int E1 = pfa ? (N2 * inv_mod(N2 % N1, N1)) % N : 0;
  static constexpr int E2 = pfa ? (N1 * inv_mod(N1 % N2, N2)) % N : 0;
  static __host__ __device__ constexpr int in_a(int n1, int n2) {
    return pfa ? (N2 * n1 + N1 * n2) % N : N2 * n1 + n2;
  }
  static __host__ __device__ constexpr int out_b(int k1, int k2) {
    return pfa ? (k1 * E1 + k2 * E2) % N : k1 + N1 * k2;
  }
};
constexpr int smallest_prime(int n) {
  for (int p = 2; p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }

// exp(SIGN 2 pi i m / R) from the constant tables
template <typename T, int R, int SIGN>
__host__ __device__ __forceinline__ cpx<T> wr(int m) {
  return {(T)Tw<R>::c[m], (T)(SIGN * Tw<R>::s[m])};
}

// in-register DFT of size R: v[k] <- sum_j v[j] exp(SIGN 2 pi i j k / R)
template <typename T, int R, int SIGN>
__host__ __device__ __forceinline__ void rdft(cpx<T>* v) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
  } else if constexpr (R == 4) {
    const cpx<T> a0 = add(v[0], v[2]), a1 = sub(v[0], v[2]);
    const cpx<T> b0 = add(v[1], v[3]), b1 = sub(v[1], v[3]);
    const cpx<T> jb1 = SIGN > 0 ? cpx<T>{-b1.y, b1.x} : cpx<T>{b1.y, -b1.x};
    v[0] = add(a0, b0);
    v[2] = sub(a0, b0);
    v[1] = add(a1, jb1);
    v[3] = sub(a1, jb1);
  } else if constexpr (smallest_prime(R) == R) {
    // odd prime: symmetric form over the pair sums ap_j = v_j + v_{R-j} and
    // differences am_j = v_j - v_{R-j}; out[k] and out[R-k] share the cosine
    // sums, and out[0] = v_0 + sum ap_j reuses the pair sums
    constexpr int H = (R - 1) / 2;
    cpx<T> ap[H + 1], am[H + 1];
#pragma unroll
    for (int j = 1; j <= H; j++) {
      ap[j] = add(v[j], v[R - j]);
      am[j] = sub(v[j], v[R - j]);
    }
    cpx<T> out[R];
    cpx<T> s0 = v[0];
#pragma unroll
    for (int j = 1; j <= H; j++) s0 = add(s0, ap[j]);
    out[0] = s0;
#pragma unroll
    for (int k = 1; k <= H; k++) {
      T ra = v[0].x, ia = v[0].y;
#pragma unroll
      for (int j = 1; j <= H; j++) {
        const T c = (T)Tw<R>::c[(j * k) % R];
        ra += c * ap[j].x;
        ia += c * ap[j].y;
      }
      // SIGN i sum_j s_jk am_j = (-SIGN ib, SIGN rb)
      if constexpr (H == 1) {
        // one sine term: fused into the output multiply-adds
        const T s = (T)(SIGN * Tw<R>::s[k % R]);
        out[k] = {ra - s * am[1].y, ia + s * am[1].x};
        out[R - k] = {ra + s * am[1].y, ia - s * am[1].x};
      } else {
        T rb = 0, ib = 0;
#pragma unroll
        for (int j = 1; j <= H; j++) {
          const T s = (T)Tw<R>::s[(j * k) % R];
          rb += s * am[j].x;
          ib += s * am[j].y;
        }
        out[k] = {ra - SIGN * ib, ia + SIGN * rb};
        out[R - k] = {ra + SIGN * ib, ia - SIGN * rb};
      }
    }
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = out[k];
  } else if constexpr (Map<(R % 4 == 0) ? 4 : smallest_prime(R),
                           R / ((R % 4 == 0) ? 4 : smallest_prime(R))>::pfa) {
    // R = p q coprime, prime factor algorithm: DFT_q over m of
    // v[(q r + p m) mod R] for each r, then DFT_p over r, output at CRT
    constexpr int p = (R % 4 == 0) ? 4 : smallest_prime(R);
    constexpr int q = R / p;
    using M = Map<p, q>;
    cpx<T> t[R];
#pragma unroll
    for (int r = 0; r < p; r++) {
      cpx<T> sv[q];
#pragma unroll
      for (int m = 0; m < q; m++) sv[m] = v[M::in_a(r, m)];
      rdft<T, q, SIGN>(sv);
#pragma unroll
      for (int k = 0; k < q; k++) t[r * q + k] = sv[k];
    }
#pragma unroll
    for (int k = 0; k < q; k++) {
      cpx<T> sv[p];
#pragma unroll
      for (int r = 0; r < p; r++) sv[r] = t[r * q + k];
      rdft<T, p, SIGN>(sv);
#pragma unroll
      for (int u = 0; u < p; u++) v[M::out_b(u, k)] = sv[u];
    }
  } else {
    // R = p q, decimation in time: X[k + q u] = sum_r W_p^{ru} W_R^{rk} Y_r[k],
    // Y_r = DFT_q of x[r + p m]
    constexpr int p = (R % 4 == 0) ? 4 : smallest_prime(R);
    constexpr int q = R / p;
    cpx<T> t[R];
#pragma unroll
    for (int r = 0; r < p; r++) {
      cpx<T> s[q];
#pragma unroll
      for (int m = 0; m < q; m++) s[m] = v[r + p * m];
      rdft<T, q, SIGN>(s);
#pragma unroll
      for (int k = 0; k < q; k++) t[r * q + k] = s[k];
    }
#pragma unroll
    for (int k = 0; k < q; k++) {
      cpx<T> s[p];
#pragma unroll
      for (int r = 0; r < p; r++)
        s[r] = (r * k == 0) ? t[r * q + k] : mul(t[r * q + k], wr<T, R, SIGN>((r * k) % R));
      rdft<T, p, SIGN>(s);
#pragma unroll
      for (int u = 0; u < p; u++) v[k + q * u] = s[u];
    }
  }
}

// Raw buffer resource over [base, base + bytes): a load at a byte offset at
// or past `bytes` returns zero and a store there is dropped (the hardware's
// range check of an unswizzled buffer), so the elements a line does not
// hold need no branch.  `bytes` stays below OOB.
constexpr uint32_t OOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_res(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}
template <typename T>
__device__ __forceinline__ cpx<T> buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  else
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, cpx<T> v) {
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, (int)off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), r, (int)off, 0, 0);
}

// twiddle W_N^{SIGN m} from the LDS table tw[m] = exp(-2 pi i m / N)
template <typename T, int SIGN>
__device__ __forceinline__ cpx<T> twl(const cpx<T>* tw, int m) {
  cpx<T> w = tw[m];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

}  // namespace f4
}  // namespace qbg

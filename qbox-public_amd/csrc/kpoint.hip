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
#include "energy.h"

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

void launch_unfold_add(const double* H, long ldu, int nst, int ngw, const int32_t* unfold,
                       double* hpsi, long ldh, hipStream_t s) {
  if (nst <= 0 || ngw <= 0) return;
  hipLaunchKernelGGL(k_unfold_add, dim3((ngw + 255) / 256, nst), dim3(256), 0, s, H, ldu, ngw,
                     unfold, hpsi, ldh);
}

}  // namespace qbg

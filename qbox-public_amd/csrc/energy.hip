// energy.hip -- pointwise, reduction and structure-factor kernels of the
// energy evaluation (EnergyFunctional::update_vhxc / energy / atoms_moved,
// ChargeDensity::update_density, NonLocalPotential projector set-up).
//
// Every sum is a fixed-order tree: a block reduces 256 values in LDS and
// writes one partial per block; a second kernel adds the partials in block
// order.  Results are therefore bit-reproducible run to run.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "energy.h"

namespace qbg {

// --------------------------------------------------------------- reduction
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* __restrict__ out) {
  __shared__ double red[NV][256];
#pragma unroll
  for (int q = 0; q < NV; q++) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < NV; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < NV) out[threadIdx.x] = red[threadIdx.x][0];
}

// the same sums through wave shuffles, then the 4 wave totals in order
template <int NV>
__device__ __forceinline__ void block_reduce_shfl(double (&v)[NV], double* __restrict__ out) {
  __shared__ double red[NV][4];
#pragma unroll
  for (int q = 0; q < NV; q++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) red[q][w] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    out[q] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
  }
}

// out[q] = sum_b partial[b*nv + q] in block order; one block of 256 threads
// per q-group, each thread summing a strided slice, then a tree.
__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ partial,
                                                         int nblocks, int nv,
                                                         double* __restrict__ out) {
  __shared__ double red[256];
  const int q = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) s += partial[(long)b * nv + q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = red[0];
}

void reduce_partials(const double* partial, int nblocks, int nv, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(nv), dim3(256), 0, s, partial, nblocks, nv, out);
}

// ------------------------------------------------------- kinetic + ekin
// src/EnergyFunctional.cpp:441-571 (ekin, sigma_ekin) and :818-827
// (cp += 0.5 kpg2 c); psi2sum(G) = sum_n occ_n |c_nG|^2 (fac = 1, real basis)
// partial layout per block: [ekin, s_xx, s_yy, s_zz, s_xy, s_yz, s_xz, 0]
// blockIdx.y: a chunk of KIN_BANDS bands (the sums below are linear in the
// bands, so the chunks' partials add up); c32 (if given): the FP32 copy of
// the coefficients for the split GEMMs, c32[2*ig + {0,1} + n*ldk], written in
// the same pass over c
constexpr int KIN_BANDS = 32;
__global__ __launch_bounds__(256) void k_kinetic(const double* __restrict__ c, long ldc, int nst,
                                                 const double* __restrict__ occ, int ngw,
                                                 const double* __restrict__ g2,
                                                 const double* __restrict__ gx,
                                                 double* __restrict__ hpsi, long ldh, int store,
                                                 float* __restrict__ c32, long ldk,
                                                 double* __restrict__ partial) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  const int n0 = blockIdx.y * KIN_BANDS, n1 = min(nst, n0 + KIN_BANDS);
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (ig < ngw) {
    const double gg = g2[ig];
    double p = 0.0;
#pragma unroll 4
    for (int n = n0; n < n1; n++) {
      const double2 cv = *reinterpret_cast<const double2*>(c + 2 * (ig + ldc * n));
      p += occ[n] * (cv.x * cv.x + cv.y * cv.y);
      if (store)
        *reinterpret_cast<double2*>(hpsi + 2 * (ig + ldh * n)) =
            double2{0.5 * gg * cv.x, 0.5 * gg * cv.y};
      if (c32) *reinterpret_cast<float2*>(c32 + 2 * ig + ldk * n) = float2{(float)cv.x, (float)cv.y};
    }
    const double x = gx[ig], y = gx[ngw + ig], z = gx[2 * ngw + ig];
    const double f = 2.0 * p;
    v[0] = p * gg;
    v[1] = f * x * x; v[2] = f * y * y; v[3] = f * z * z;
    v[4] = f * x * y; v[5] = f * y * z; v[6] = f * x * z;
  }
  block_reduce_store<8>(v, partial + ((long)blockIdx.y * gridDim.x + blockIdx.x) * 8);
}

int launch_kinetic(const double* c, long ldc, int nst, const double* occ, int ngw,
                   const double* g2, const double* gx, double* hpsi, long ldh, int store,
                   double* partial, double* out8, hipStream_t s, float* c32, long ldk) {
  const int nb = (ngw + 255) / 256, nch = (nst + KIN_BANDS - 1) / KIN_BANDS;
  hipLaunchKernelGGL(k_kinetic, dim3(nb, nch), dim3(256), 0, s, c, ldc, nst, occ, ngw, g2, gx,
                     hpsi, ldh, store, c32, ldk, partial);
  reduce_partials(partial, nb * nch, 8, out8, s);
  return nb * nch;
}

// ------------------------------------------------------------------ LDA
// src/LDAFunctional.cpp:92-151 (Perdew-Zunger fit of Ceperley-Alder)
__device__ __forceinline__ void lda_pz(double rh, double& ee, double& vv) {
  const double c1 = 0.6203504908994001;
  const double c3 = -0.610887057711;
  const double A = 0.0311, B = -0.048, b1 = 1.0529, b2 = 0.3334, G = -0.1423;
  const double D = G / (1.0 + b1 + b2) - B;
  const double C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)));
  ee = 0.0;
  vv = 0.0;
  if (rh > 0.0) {
    const double rs = c1 / cbrt(rh);
    const double vx = c3 / rs;
    const double ex = 0.75 * vx;
    double ec, vc;
    if (rs < 1.0) {
      const double lr = log(rs);
      ec = A * lr + B + C * rs * lr + D * rs;
      vc = A * lr + (B - A / 3.0) + (2.0 / 3.0) * C * rs * lr + ((2.0 * D - C) / 3.0) * rs;
    } else {
      const double sr = sqrt(rs);
      const double den = 1.0 + b1 * sr + b2 * rs;
      ec = G / den;
      vc = ec * (1.0 + (7.0 / 6.0) * b1 * sr + (4.0 / 3.0) * b2 * rs) / den;
    }
    ee = ex + ec;
    vv = vx + vc;
  }
}

// vxc_r and partial sums [sum rho e, sum rho (e - v), sum rho, 0]
// (src/XCPotential.cpp:193-229, src/ChargeDensity.cpp:178-187); with a core
// density rc (or null) the functional sees rho + rc (src/XCPotential.cpp:
// 154-179) while the third sum stays the valence charge
__global__ __launch_bounds__(256) void k_lda(const double* __restrict__ rho,
                                             const double* __restrict__ rc, long n,
                                             double* __restrict__ vxc,
                                             double* __restrict__ partial) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  double v[4] = {0, 0, 0, 0};
  if (i < n) {
    double e, vv;
    const double r = rc ? rho[i] + rc[i] : rho[i];
    lda_pz(r, e, vv);
    vxc[i] = vv;
    v[0] = r * e;
    v[1] = r * (e - vv);
    v[2] = rho[i];
  }
  block_reduce_store<4>(v, partial + (long)blockIdx.x * 4);
}

int launch_lda(const double* rho, const double* rc, long n, double* vxc, double* partial,
               double* out4, hipStream_t s) {
  const int nb = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_lda, dim3(nb), dim3(256), 0, s, rho, rc, n, vxc, partial);
  reduce_partials(partial, nb, 4, out4, s);
  return nb;
}

// fully polarized Perdew-Zunger fit (src/LDAFunctional.cpp:153-213):
// exchange 2^(1/3) times the paramagnetic one, the polarized correlation
// parameters, C and D by continuity at rs = 1
__device__ __forceinline__ void lda_pz_ferro(double rh, double& ee, double& vv) {
  const double c1 = 0.6203504908994001;
  const double cx = -0.769669463118;
  const double A = 0.01555, B = -0.0269, b1 = 1.3981, b2 = 0.2611, G = -0.0843;
  const double D = G / (1.0 + b1 + b2) - B;
  const double C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)));
  ee = 0.0;
  vv = 0.0;
  if (rh > 0.0) {
    const double rs = c1 / cbrt(rh);
    const double vx = cx / rs, ex = 0.75 * vx;
    double ec, vc;
    if (rs < 1.0) {
      const double lr = log(rs);
      ec = A * lr + B + C * rs * lr + D * rs;
      vc = A * lr + (B - A / 3.0) + (2.0 / 3.0) * C * rs * lr + ((2.0 * D - C) / 3.0) * rs;
    } else {
      const double sr = sqrt(rs);
      const double den = 1.0 + b1 * sr + b2 * rs;
      ec = G / den;
      vc = ec * (1.0 + (7.0 / 6.0) * b1 * sr + (4.0 / 3.0) * b2 * rs) / den;
    }
    ee = ex + ec;
    vv = vx + vc;
  }
}

// spin-polarized LDA (src/LDAFunctional.cpp:42-89): interpolation between
// the paramagnetic and ferromagnetic fits with f(zeta); partial sums
// [sum rho e, sum (rho e - rho_up v_up - rho_dn v_dn), sum rho, 0]
// (src/XCPotential.cpp:206-218)
__global__ __launch_bounds__(256) void k_lsda(const double* __restrict__ rup,
                                              const double* __restrict__ rdn,
                                              const double* __restrict__ rc, long n,
                                              double* __restrict__ vup,
                                              double* __restrict__ vdn,
                                              double* __restrict__ partial) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  double v[4] = {0, 0, 0, 0};
  if (i < n) {
    // each spin with its share of the core density rc (or null)
    const double ci = rc ? rc[i] : 0.0;
    const double ru = rup[i] + ci, rd = rdn[i] + ci, r = ru + rd;
    double e = 0.0, vu = 0.0, vd = 0.0;
    if (r > 0.0) {
      const double kf = 1.0 / (2.0 * cbrt(2.0) - 2.0);
      const double z = (ru - rd) / r;
      const double a = 1.0 + z, b = 1.0 - z;
      const double a3 = cbrt(a), b3 = cbrt(b);
      const double f = kf * (a3 * a + b3 * b - 2.0);
      const double df = (4.0 / 3.0) * kf * (a3 - b3);
      double eu, vvu, ep, vvp;
      lda_pz(r, eu, vvu);
      lda_pz_ferro(r, ep, vvp);
      const double de = ep - eu;
      e = eu + f * de;
      const double vm = vvu + f * (vvp - vvu);
      vu = vm + de * (1.0 - z) * df;
      vd = vm + de * (-1.0 - z) * df;
    }
    vup[i] = vu;
    vdn[i] = vd;
    v[0] = r * e;
    v[1] = r * e - ru * vu - rd * vd;
    v[2] = rup[i] + rdn[i];
  }
  block_reduce_store<4>(v, partial + (long)blockIdx.x * 4);
}

int launch_lsda(const double* rup, const double* rdn, const double* rc, long n, double* vup,
                double* vdn, double* partial, double* out4, hipStream_t s) {
  const int nb = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_lsda, dim3(nb), dim3(256), 0, s, rup, rdn, rc, n, vup, vdn, partial);
  reduce_partials(partial, nb, 4, out4, s);
  return nb;
}

// LDA pointwise values for a test vector (no reduction)
__global__ void k_lda_points(const double* __restrict__ rho, int n, double* e, double* v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) lda_pz(rho[i], e[i], v[i]);
}
void launch_lda_points(const double* rho, int n, double* e, double* v, hipStream_t s) {
  hipLaunchKernelGGL(k_lda_points, dim3((n + 255) / 256), dim3(256), 0, s, rho, n, e, v);
}

// ---------------------------------------------------------- vhxc in G
// src/EnergyFunctional.cpp:243-383: rhoelg = rhog/omega; eps, Hartree sums;
// rhogt = rhoelg + rhopst; vlocal_g = vion + 4 pi rhogt g2i.
// partial per block: [sum Re(conj r vion), ehsum, ehesum, ehepsum, ehpsum,
//                     6 x hartree-stress gg sums, 6 x eps-stress sums, 0]
__global__ __launch_bounds__(256) void k_vhxc_g(
    const double* __restrict__ rhog, int ngv, double omega_inv, const double* __restrict__ vion,
    const double* __restrict__ dvion, const double* __restrict__ rhopst,
    const double* __restrict__ g2i, const double* __restrict__ gi, const double* __restrict__ gx,
    double* __restrict__ rhoelg, double* __restrict__ rhogt, double* __restrict__ vlocal,
    int stress, double* __restrict__ partial) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  double v[18];
#pragma unroll
  for (int q = 0; q < 18; q++) v[q] = 0.0;
  if (ig < ngv) {
    const double rr = omega_inv * rhog[2 * ig], ri = omega_inv * rhog[2 * ig + 1];
    rhoelg[2 * ig] = rr;
    rhoelg[2 * ig + 1] = ri;
    const double pr = rhopst[2 * ig], pi = rhopst[2 * ig + 1];
    const double vr = vion[2 * ig], vi = vion[2 * ig + 1];
    const double g2i_ = g2i[ig];
    v[0] = rr * vr + ri * vi;
    const double tr = rr + pr, ti = ri + pi;
    v[1] = (tr * tr + ti * ti) * g2i_;
    v[2] = (rr * rr + ri * ri) * g2i_;
    v[3] = 2.0 * ((rr * pr + ri * pi) * g2i_);
    v[4] = (pr * pr + pi * pi) * g2i_;
    rhogt[2 * ig] = tr;
    rhogt[2 * ig + 1] = ti;
    const double fpi = 4.0 * M_PI;
    vlocal[2 * ig] = vr + fpi * tr * g2i_;
    vlocal[2 * ig + 1] = vi + fpi * ti * g2i_;
    if (stress) {
      const double x = gx[ig], y = gx[ngv + ig], z = gx[2 * ngv + ig];
      // sigma_ehart gg part (:634-647): |rhogt|^2 g2i^2 G_i G_j
      const double t = (tr * tr + ti * ti) * g2i_ * g2i_;
      v[5] = t * x * x; v[6] = t * y * y; v[7] = t * z * z;
      v[8] = t * x * y; v[9] = t * y * z; v[10] = t * x * z;
      // sigma_eps (:580-597): 2 gi Re(conj(rhoelg) dvion) G_i G_j
      const double f = 2.0 * gi[ig] * (rr * dvion[2 * ig] + ri * dvion[2 * ig + 1]);
      v[11] = f * x * x; v[12] = f * y * y; v[13] = f * z * z;
      v[14] = f * x * y; v[15] = f * y * z; v[16] = f * x * z;
    }
  }
  block_reduce_store<18>(v, partial + (long)blockIdx.x * 18);
}

int launch_vhxc_g(const double* rhog, int ngv, double omega_inv, const double* vion,
                  const double* dvion, const double* rhopst, const double* g2i, const double* gi,
                  const double* gx, double* rhoelg, double* rhogt, double* vlocal, int stress,
                  double* partial, double* out18, hipStream_t s) {
  const int nb = (ngv + 255) / 256;
  hipLaunchKernelGGL(k_vhxc_g, dim3(nb), dim3(256), 0, s, rhog, ngv, omega_inv, vion, dvion,
                     rhopst, g2i, gi, gx, rhoelg, rhogt, vlocal, stress, partial);
  reduce_partials(partial, nb, 18, out18, s);
  return nb;
}

// ----------------------------------------------------- structure factor
// src/StructureFactor.cpp:70-162: per atom and axis cos/sin(k * b_axis.tau)
// for k in [kmin, kmax]; tables laid out tab[((a*3 + axis)*krange + k-kmin)*2]
__global__ void k_sf_tables(const double* __restrict__ tau, int nat, const double* __restrict__ b,
                            int kmin0, int kmin1, int kmin2, int krange,
                            double* __restrict__ tab) {
  const int a = blockIdx.x;
  const int kmins[3] = {kmin0, kmin1, kmin2};
  const double t0 = tau[3 * a], t1 = tau[3 * a + 1], t2 = tau[3 * a + 2];
  for (int ax = 0; ax < 3; ax++) {
    const double fac = b[3 * ax] * t0 + b[3 * ax + 1] * t1 + b[3 * ax + 2] * t2;
    for (int k = threadIdx.x; k < krange; k += blockDim.x) {
      const double arg = (double)(k + kmins[ax]) * fac;
      double sn, cs;
      sincos(arg, &sn, &cs);
      double* p = tab + ((long)(a * 3 + ax) * krange + k) * 2;
      p[0] = cs;
      p[1] = sn;
    }
  }
}

void launch_sf_tables(const double* tau, int nat, const double* b, const int* kmin, int krange,
                      double* tab, hipStream_t s) {
  hipLaunchKernelGGL(k_sf_tables, dim3(nat), dim3(128), 0, s, tau, nat, b, kmin[0], kmin[1],
                     kmin[2], krange, tab);
}

// exp(-i G.tau_a) = (ca - i sa)(cb - i sb)(cc - i sc) as in :144-150
__device__ __forceinline__ void eigr_from_tables(const double* __restrict__ tab, int a, int krange,
                                                 int kx, int ky, int kz, double& re, double& im) {
  const double* p0 = tab + ((long)(a * 3 + 0) * krange + kx) * 2;
  const double* p1 = tab + ((long)(a * 3 + 1) * krange + ky) * 2;
  const double* p2 = tab + ((long)(a * 3 + 2) * krange + kz) * 2;
  const double ca = p0[0], sa = p0[1], cb = p1[0], sb = p1[1], cc = p2[0], sc = p2[1];
  re = ca * cb * cc - sa * sb * cc - sa * cb * sc - ca * sb * sc;
  im = sa * sb * sc - sa * cb * cc - ca * sb * cc - ca * cb * sc;
}

// sfac[is][ig] = sum_{a in is} exp(-iG.tau_a); then (src/EnergyFunctional.cpp:1085-1095)
// rhopst = sum_s sfac rhops, vion = sum_s sfac vps, dvion = sum_s sfac dvps
__global__ __launch_bounds__(256) void k_sfac(
    const int32_t* __restrict__ idx, int ngv, const double* __restrict__ tab, int krange,
    int kmin0, int kmin1, int kmin2, int nsp, const int32_t* __restrict__ sp_first,
    const double* __restrict__ vps, const double* __restrict__ dvps,
    const double* __restrict__ rhops, double* __restrict__ sfac, double* __restrict__ vion,
    double* __restrict__ dvion, double* __restrict__ rhopst) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngv) return;
  const int kx = idx[3 * ig] - kmin0, ky = idx[3 * ig + 1] - kmin1, kz = idx[3 * ig + 2] - kmin2;
  double vr = 0, vi = 0, dr = 0, di = 0, pr = 0, pi = 0;
  for (int is = 0; is < nsp; is++) {
    double sr = 0.0, si = 0.0;
    for (int a = sp_first[is]; a < sp_first[is + 1]; a++) {
      double re, im;
      eigr_from_tables(tab, a, krange, kx, ky, kz, re, im);
      sr += re;
      si += im;
    }
    sfac[2 * ((long)is * ngv + ig)] = sr;
    sfac[2 * ((long)is * ngv + ig) + 1] = si;
    const long o = (long)is * ngv + ig;
    pr += sr * rhops[o]; pi += si * rhops[o];
    vr += sr * vps[o]; vi += si * vps[o];
    dr += sr * dvps[o]; di += si * dvps[o];
  }
  vion[2 * ig] = vr; vion[2 * ig + 1] = vi;
  dvion[2 * ig] = dr; dvion[2 * ig + 1] = di;
  rhopst[2 * ig] = pr; rhopst[2 * ig + 1] = pi;
}

void launch_sfac(const int32_t* idx, int ngv, const double* tab, int krange, const int* kmin,
                 int nsp, const int32_t* sp_first, const double* vps, const double* dvps,
                 const double* rhops, double* sfac, double* vion, double* dvion, double* rhopst,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_sfac, dim3((ngv + 255) / 256), dim3(256), 0, s, idx, ngv, tab, krange,
                     kmin[0], kmin[1], kmin[2], nsp, sp_first, vps, dvps, rhops, sfac, vion,
                     dvion, rhopst);
}

// ---------------------------------------------------- core density (NLCC)
// rhocore_g = spin_fac sum_s sfac_s rhoc_s with rhoc_s = rhocoreg/omega and
// spin_fac = omega / nspin (src/EnergyFunctional.cpp:1064-1077)
__global__ __launch_bounds__(256) void k_core_g(int ngv, int nsp, const double* __restrict__ sfac,
                                                const double* __restrict__ rhoc, double spin_fac,
                                                double* __restrict__ rhocg) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngv) return;
  double re = 0.0, im = 0.0;
  for (int is = 0; is < nsp; is++) {
    const long o = (long)is * ngv + ig;
    re += sfac[2 * o] * rhoc[o];
    im += sfac[2 * o + 1] * rhoc[o];
  }
  rhocg[2 * ig] = spin_fac * re;
  rhocg[2 * ig + 1] = spin_fac * im;
}

void launch_core_g(int ngv, int nsp, const double* sfac, const double* rhoc, double spin_fac,
                   double* rhocg, hipStream_t s) {
  hipLaunchKernelGGL(k_core_g, dim3((ngv + 255) / 256), dim3(256), 0, s, ngv, nsp, sfac, rhoc,
                     spin_fac, rhocg);
}

// sum_G w_G Re(conj(vxc_g) rhocore_g), w = 1 at G = 0 (index 0) and 2
// otherwise: the half-sphere form of the core term of dxc
// (src/EnergyFunctional.cpp:290-302)
__global__ __launch_bounds__(256) void k_core_dxc(int ngv, const double* __restrict__ vxcg,
                                                  const double* __restrict__ rhocg,
                                                  double* __restrict__ partial) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  double v[1] = {0.0};
  if (ig < ngv)
    v[0] = (ig == 0 ? 1.0 : 2.0) *
           (vxcg[2 * ig] * rhocg[2 * ig] + vxcg[2 * ig + 1] * rhocg[2 * ig + 1]);
  block_reduce_store<1>(v, partial + blockIdx.x);
}

void launch_core_dxc(int ngv, const double* vxcg, const double* rhocg, double* partial,
                     double* out1, hipStream_t s) {
  const int nb = (ngv + 255) / 256;
  hipLaunchKernelGGL(k_core_dxc, dim3(nb), dim3(256), 0, s, ngv, vxcg, rhocg, partial);
  reduce_partials(partial, nb, 1, out1, s);
}

// ---------------------------------------------------------- local forces
// src/EnergyFunctional.cpp:841-918: per atom a of species s,
//   vtemp_s = 4 pi rhops_s g2i conj(rhogt) + vps_s conj(rhoelg)
//   f_a,j = -2 omega sum_G (Re e * Im vtemp_s + Im e * Re vtemp_s) G_j
// vtemp_s is formed once per species; a block then takes a chunk of G and a
// group of up to LF_AG atoms of one species, with the atoms' per-axis phase
// tables in LDS, so the G data is read once per group instead of per atom.
constexpr int LF_AG = 8;

//   (+ conj(vxc_g) rhoc_s with a core density, :855-856)
__global__ __launch_bounds__(256) void k_vtemp(int ngv, int nsp, const double* __restrict__ vps,
                                               const double* __restrict__ rhops,
                                               const double* __restrict__ g2i,
                                               const double* __restrict__ rhogt,
                                               const double* __restrict__ rhoelg,
                                               const double* __restrict__ vxcg,
                                               const double* __restrict__ rhoc,
                                               double* __restrict__ vt) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngv) return;
  const double fpi = 4.0 * M_PI;
  const double gtr = rhogt[2 * ig], gti = rhogt[2 * ig + 1];
  const double elr = rhoelg[2 * ig], eli = rhoelg[2 * ig + 1];
  const double vxr = vxcg ? vxcg[2 * ig] : 0.0, vxi = vxcg ? vxcg[2 * ig + 1] : 0.0;
  for (int is = 0; is < nsp; is++) {
    const long o = (long)is * ngv + ig;
    const double tmp = fpi * rhops[o] * g2i[ig];
    // conj: (x + iy)* = x - iy
    double re = tmp * gtr + vps[o] * elr;
    double im = -tmp * gti - vps[o] * eli;
    if (vxcg) {
      re += vxr * rhoc[o];
      im -= vxi * rhoc[o];
    }
    vt[2 * o] = re;
    vt[2 * o + 1] = im;
  }
}

// grid (chunks, groups); partial[(a*nchunk + chunk)*3 + j]
__global__ __launch_bounds__(256) void k_local_forces(
    const int32_t* __restrict__ idx, int ngv, const double* __restrict__ tab, int krange,
    int kmin0, int kmin1, int kmin2, const int32_t* __restrict__ grp, const double* __restrict__ vt,
    const double* __restrict__ gx, int chunk, double* __restrict__ partial) {
  extern __shared__ __align__(16) double ltab[];   // [LF_AG][3][krange][2]
  const int a0 = grp[3 * blockIdx.y], na = grp[3 * blockIdx.y + 1], is = grp[3 * blockIdx.y + 2];
  const int tsz = 3 * krange * 2;
  for (int i = threadIdx.x; i < na * tsz; i += 256) ltab[i] = tab[(long)a0 * tsz + i];
  __syncthreads();
  const int g0 = blockIdx.x * chunk, g1 = min(ngv, g0 + chunk);
  double v[3 * LF_AG];
#pragma unroll
  for (int q = 0; q < 3 * LF_AG; q++) v[q] = 0.0;
  const double* vts = vt + 2 * (long)is * ngv;
  for (int ig = g0 + threadIdx.x; ig < g1; ig += 256) {
    const double vtr = vts[2 * ig], vti = vts[2 * ig + 1];
    const double gxx = gx[ig], gyy = gx[ngv + ig], gzz = gx[2 * ngv + ig];
    const int kx = idx[3 * ig] - kmin0, ky = idx[3 * ig + 1] - kmin1, kz = idx[3 * ig + 2] - kmin2;
#pragma unroll
    for (int q = 0; q < LF_AG; q++) {
      if (q < na) {
        double er, ei;
        eigr_from_tables(ltab + q * tsz, 0, krange, kx, ky, kz, er, ei);
        const double t = er * vti + ei * vtr;
        v[3 * q] += t * gxx;
        v[3 * q + 1] += t * gyy;
        v[3 * q + 2] += t * gzz;
      }
    }
  }
  __shared__ double res[3 * LF_AG];
  block_reduce_shfl<3 * LF_AG>(v, res);
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * na; i += 256)
    partial[((long)(a0 + i / 3) * gridDim.x + blockIdx.x) * 3 + i % 3] = res[i];
}

// sum chunks per atom: f[3a+j] = fion_esr[3a+j] - 2 omega sum_chunk partial
__global__ void k_local_forces_final(const double* __restrict__ partial, int nchunk, int nat,
                                     double omega, const double* __restrict__ fesr,
                                     double* __restrict__ fion) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * nat) return;
  const int a = t / 3, j = t % 3;
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += partial[((long)a * nchunk + c) * 3 + j];
  fion[t] += fesr[t] + (-2.0 * omega) * s;
}

// atoms per group: at most LF_AG, and their tables within 64 KiB of LDS
int local_forces_group_size(int krange) {
  return std::max(1, std::min(LF_AG, 65536 / (int)(sizeof(double) * 6 * krange)));
}

void launch_local_forces(const int32_t* idx, int ngv, const double* tab, int krange,
                         const int* kmin, const int32_t* grp, int ngrp, int nsp,
                         const double* vps, const double* rhops, const double* g2i,
                         const double* gx, const double* rhogt, const double* rhoelg,
                         const double* vxcg, const double* rhoc,
                         double omega, const double* fesr, int nat, double* vt, double* partial,
                         int nchunk, double* fion, hipStream_t s) {
  hipLaunchKernelGGL(k_vtemp, dim3((ngv + 255) / 256), dim3(256), 0, s, ngv, nsp, vps, rhops, g2i,
                     rhogt, rhoelg, vxcg, rhoc, vt);
  const int chunk = (ngv + nchunk - 1) / nchunk;
  const size_t lds = sizeof(double) * (size_t)local_forces_group_size(krange) * 3 * krange * 2;
  hipLaunchKernelGGL(k_local_forces, dim3(nchunk, ngrp), dim3(256), lds, s, idx, ngv, tab,
                     krange, kmin[0], kmin[1], kmin[2], grp, vt, gx, chunk, partial);
  hipLaunchKernelGGL(k_local_forces_final, dim3((3 * nat + 255) / 256), dim3(256), 0, s,
                     partial, nchunk, nat, omega, fesr, fion);
}

// ------------------------------------------------------- projectors anl
// src/NonLocalPotential.cpp:1314-1395: anl = twnl (-i)^l e^{-iG.tau} for
// l = 0, 1, 2.  One thread per (G, atom) forms the phase once and writes the
// atom's projector columns m = a_m0[a] + ipr*a_stride[a]:
// A[m*lda + 2*ig + {0,1}].
constexpr int ANL_ATOMS = 8;
template <typename T>
struct alignas(2 * sizeof(T)) cpx2 {
  T x, y;
};
template <typename T>
__global__ __launch_bounds__(256) void k_anl(int ngw, int nat, long lda, const double* __restrict__ gx,
                                             const double* __restrict__ tau,
                                             const int32_t* __restrict__ a_m0,
                                             const int32_t* __restrict__ a_npr,
                                             const int32_t* __restrict__ a_stride,
                                             const int32_t* __restrict__ m_twnl,
                                             const int32_t* __restrict__ m_l,
                                             const double* __restrict__ twnl,
                                             T* __restrict__ A) {
  // ANL_ATOMS atoms per thread: G read once, the atoms' stores in flight
  // together; each complex entry is one vector store
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngw) return;
  const int a0 = blockIdx.y * ANL_ATOMS, a1 = min(nat, a0 + ANL_ATOMS);
  const double gxx = gx[ig], gyy = gx[ngw + ig], gzz = gx[2 * ngw + ig];
  for (int a = a0; a < a1; a++) {
    const double arg = -(gxx * tau[3 * a] + gyy * tau[3 * a + 1] + gzz * tau[3 * a + 2]);
    double s, c;
    sincos(arg, &s, &c);
    const int m0 = a_m0[a], npr = a_npr[a], st = a_stride[a];
    for (int ipr = 0; ipr < npr; ipr++) {
      const int m = m0 + ipr * st;
      const double t = twnl[(long)m_twnl[m] + ig];
      // (-i)^l (c + i s): l = 0 (c, s), 1 (s, -c), 2 (-c, -s), 3 (-s, c)
      cpx2<T> v;
      const int l = m_l[m];
      if (l == 0) { v.x = (T)(t * c); v.y = (T)(t * s); }
      else if (l == 1) { v.x = (T)(t * s); v.y = (T)(-t * c); }
      else if (l == 2) { v.x = (T)(-t * c); v.y = (T)(-t * s); }
      else { v.x = (T)(-t * s); v.y = (T)(t * c); }
      reinterpret_cast<cpx2<T>*>(A + (long)m * lda)[ig] = v;
    }
  }
}

void launch_anl(int ngw, int nat, long lda, const double* gx, const double* tau,
                const int32_t* a_m0, const int32_t* a_npr, const int32_t* a_stride,
                const int32_t* m_twnl, const int32_t* m_l, const double* twnl, void* A, int prec,
                hipStream_t s) {
  dim3 grid((ngw + 255) / 256, (nat + ANL_ATOMS - 1) / ANL_ATOMS);
  if (prec == 0)
    hipLaunchKernelGGL(k_anl<double>, grid, dim3(256), 0, s, ngw, nat, lda, gx, tau, a_m0,
                       a_npr, a_stride, m_twnl, m_l, twnl, (double*)A);
  else
    hipLaunchKernelGGL(k_anl<float>, grid, dim3(256), 0, s, ngw, nat, lda, gx, tau, a_m0,
                       a_npr, a_stride, m_twnl, m_l, twnl, (float*)A);
}

// ------------------------------------------------- D matrix, enl, forces
// src/NonLocalPotential.cpp:1469-1571.  F (M x nst) holds fnl after the
// G=0 correction and factor 2.  For column m (atom a, projector ipr of
// species s): F'_m = sum_jpr D_s[ipr][jpr] F_{m(a,jpr)};
// enl += occ_n wt/omega F F'; fout = wt/omega F'.
// m_dbase[m]: offset of the species D block; m_npr[m]; m_stride[m] = na_s
// (columns of one atom are ipr*na_s apart); m_ipr[m].
__global__ __launch_bounds__(256) void k_dmat_enl(
    const double* __restrict__ F, int M, int nst, const double* __restrict__ occ,
    const int32_t* __restrict__ m_ipr, const int32_t* __restrict__ m_npr,
    const int32_t* __restrict__ m_stride, const int32_t* __restrict__ m_dbase,
    const double* __restrict__ dmat, const double* __restrict__ m_wt, double omega_inv,
    double* __restrict__ fout, double* __restrict__ partial) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  double v[1] = {0.0};
  if (t < (long)M * nst) {
    const int m = (int)(t % M), n = (int)(t / M);
    const int ipr = m_ipr[m], npr = m_npr[m], st = m_stride[m];
    const int m0 = m - ipr * st;   // column of projector 0 of this atom
    const double* D = dmat + m_dbase[m] + (long)ipr * npr;
    double fp = 0.0;
    for (int j = 0; j < npr; j++) fp += D[j] * F[(long)n * M + m0 + (long)j * st];
    const double fac = m_wt[m] * omega_inv;
    v[0] = fac * occ[n] * F[t] * fp;
    fout[t] = fac * fp;
  }
  block_reduce_store<1>(v, partial + blockIdx.x);
}

int launch_dmat_enl(const double* F, int M, int nst, const double* occ, const int32_t* m_ipr,
                    const int32_t* m_npr, const int32_t* m_stride, const int32_t* m_dbase,
                    const double* dmat, const double* m_wt, double omega_inv, double* fout,
                    double* partial, double* out1, hipStream_t s) {
  const long tot = (long)M * nst;
  const int nb = (int)((tot + 255) / 256);
  hipLaunchKernelGGL(k_dmat_enl, dim3(nb), dim3(256), 0, s, F, M, nst, occ, m_ipr, m_npr,
                     m_stride, m_dbase, dmat, m_wt, omega_inv, fout, partial);
  reduce_partials(partial, nb, 1, out1, s);
  return nb;
}

// fion[3a+j] -= sum_{m of atom a} sum_n 2 occ_n fout[m,n] dF_j[m,n]
// (src/NonLocalPotential.cpp:1734-1751); one block per (atom, j)
__global__ __launch_bounds__(256) void k_nl_forces(const double* __restrict__ fout,
                                                   const double* __restrict__ dF, int M, int nst,
                                                   const double* __restrict__ occ,
                                                   const int32_t* __restrict__ a_m0,
                                                   const int32_t* __restrict__ a_npr,
                                                   const int32_t* __restrict__ a_stride,
                                                   double* __restrict__ fion) {
  // dF is the 3M x nst block of the force GEMM: dF[(j*M + m) + n*3M]
  const int a = blockIdx.x, j = blockIdx.y;
  const int m0 = a_m0[a], npr = a_npr[a], st = a_stride[a];
  double v[1] = {0.0};
  for (int t = threadIdx.x; t < npr * nst; t += 256) {
    const int ipr = t % npr, n = t / npr;
    const int m = m0 + ipr * st;
    v[0] += 2.0 * occ[n] * fout[(long)n * M + m] * dF[(long)n * 3 * M + (long)j * M + m];
  }
  __shared__ double res[1];
  block_reduce_store<1>(v, res);
  __syncthreads();
  if (threadIdx.x == 0) fion[3 * a + j] -= res[0];
}

void launch_nl_forces(const double* fout, const double* dF, int M, int nst, const double* occ,
                      const int32_t* a_m0, const int32_t* a_npr, const int32_t* a_stride,
                      int nat, double* fion, hipStream_t s) {
  hipLaunchKernelGGL(k_nl_forces, dim3(nat, 3), dim3(256), 0, s, fout, dF, M, nst, occ, a_m0,
                     a_npr, a_stride, fion);
}

// Forces by reassociation of src/NonLocalPotential.cpp:1703-1751:
//   f_a,j = -sum_{m in a} sum_n 2 occ_n fout[m,n] dF_j[m,n],  dF_j = 2 anl_j^T c,
//         = -4 sum_{m in a} sum_K anl_j[K,m] Y[K,m],  Y = c W,  W[n,m] = occ_n fout[m,n],
// with anl_j = -i G_j anl, i.e. in the real view anl_j[2ig] = G_j a_im,
// anl_j[2ig+1] = -G_j a_re.  One GEMM instead of three.
__global__ void k_force_w(const double* __restrict__ fout, const double* __restrict__ occ,
                          int M, int nst, double* __restrict__ W) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)M * nst) return;
  const int nb = (int)(t % nst), m = (int)(t / nst);
  W[t] = occ[nb] * fout[m + (long)nb * M];     // W[nb + m*nst]
}

void launch_force_w(const double* fout, const double* occ, int M, int nst, double* W,
                    hipStream_t s) {
  const long tot = (long)M * nst;
  hipLaunchKernelGGL(k_force_w, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, fout, occ,
                     M, nst, W);
}

// partial[(a*nchunk + chunk)*3 + j] = sum over the chunk's G and the atom's
// projectors of G_j (a_im Y_re - a_re Y_im), accumulated in FP64 for FP32
// operands too
template <typename T>
__global__ __launch_bounds__(256) void k_nl_forces_y(
    const T* __restrict__ A, const T* __restrict__ Y, long lda, int ngw,
    const double* __restrict__ gx, const int32_t* __restrict__ a_m0,
    const int32_t* __restrict__ a_npr, const int32_t* __restrict__ a_stride, int chunk,
    double* __restrict__ partial) {
  const int a = blockIdx.y;
  const int m0 = a_m0[a], npr = a_npr[a], st = a_stride[a];
  const int g0 = blockIdx.x * chunk, g1 = min(ngw, g0 + chunk);
  double v[3] = {0, 0, 0};
  for (int ip = 0; ip < npr; ip++) {
    const long o = (long)(m0 + ip * st) * lda;
    for (int ig = g0 + threadIdx.x; ig < g1; ig += 256) {
      const double ar = (double)A[o + 2 * ig], ai = (double)A[o + 2 * ig + 1];
      const double yr = (double)Y[o + 2 * ig], yi = (double)Y[o + 2 * ig + 1];
      const double t = ai * yr - ar * yi;
      v[0] += gx[ig] * t;
      v[1] += gx[ngw + ig] * t;
      v[2] += gx[2 * ngw + ig] * t;
    }
  }
  block_reduce_store<3>(v, partial + ((long)a * gridDim.x + blockIdx.x) * 3);
}

__global__ void k_nl_forces_final(const double* __restrict__ partial, int nchunk, int nat,
                                  double* __restrict__ fion) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * nat) return;
  const int a = t / 3, j = t % 3;
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += partial[((long)a * nchunk + c) * 3 + j];
  fion[t] -= 4.0 * s;
}

// the fused force product's partials: sum over the row tiles in a fixed-order
// tree (one block per projector and component), then over each atom's
// projectors: fion[3a+j] -= 4 sum (src/NonLocalPotential.cpp:1560-1770)
__global__ __launch_bounds__(256) void k_fpart_sum(const double* __restrict__ fpart, int ntq,
                                                   double* __restrict__ red) {
  double v[1] = {0.0};
  for (int t = threadIdx.x; t < ntq; t += 256) v[0] += fpart[(long)blockIdx.x * ntq + t];
  block_reduce_store<1>(v, red + blockIdx.x);
}
__global__ void k_nl_forces_proj(const double* __restrict__ red,
                                 const int32_t* __restrict__ a_m0,
                                 const int32_t* __restrict__ a_npr,
                                 const int32_t* __restrict__ a_stride, int nat,
                                 double* __restrict__ fion) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * nat) return;
  const int a = t / 3, j = t - 3 * a;
  double sum = 0.0;
  for (int ip = 0; ip < a_npr[a]; ip++) sum += red[(long)(a_m0[a] + ip * a_stride[a]) * 3 + j];
  fion[t] -= 4.0 * sum;
}
void launch_nl_forces_tiles(const double* fpart, int M, int ntq, double* red,
                            const int32_t* a_m0, const int32_t* a_npr,
                            const int32_t* a_stride, int nat, double* fion, hipStream_t s) {
  hipLaunchKernelGGL(k_fpart_sum, dim3(3 * M), dim3(256), 0, s, fpart, ntq, red);
  hipLaunchKernelGGL(k_nl_forces_proj, dim3((3 * nat + 255) / 256), dim3(256), 0, s, red, a_m0,
                     a_npr, a_stride, nat, fion);
}

template <typename T>
void launch_nl_forces_y(const T* A, const T* Y, long lda, int ngw, const double* gx,
                        const int32_t* a_m0, const int32_t* a_npr, const int32_t* a_stride,
                        int nat, double* partial, int nchunk, double* fion, hipStream_t s) {
  const int chunk = (ngw + nchunk - 1) / nchunk;
  hipLaunchKernelGGL(k_nl_forces_y<T>, dim3(nchunk, nat), dim3(256), 0, s, A, Y, lda, ngw, gx, a_m0,
                     a_npr, a_stride, chunk, partial);
  hipLaunchKernelGGL(k_nl_forces_final, dim3((3 * nat + 255) / 256), dim3(256), 0, s, partial,
                     nchunk, nat, fion);
}
template void launch_nl_forces_y<double>(const double*, const double*, long, int, const double*,
                                         const int32_t*, const int32_t*, const int32_t*, int,
                                         double*, int, double*, hipStream_t);
template void launch_nl_forces_y<float>(const float*, const float*, long, int, const double*,
                                        const int32_t*, const int32_t*, const int32_t*, int,
                                        double*, int, double*, hipStream_t);

}  // namespace qbg

namespace qbg {

// ------------------------------------------------------------ NL stress
// Reassociation of src/NonLocalPotential.cpp:1786-2037 (the six dfnl_ij =
// 2 danl_ij^T c products): with Y = c W, W[n,m] = occ_n fout[m,n] (the
// matrix of the force GEMM),
//   tsum_ij = sum_{m,n} 2 occ_n fout[m,n] dfnl_ij[m,n]
//           = 4 sum_m sum_K danl_ij[K,m] Y[K,m],
// danl_ij = dtwnl_ij (-i)^l e^{-iG.tau}.  Summing the atoms of a species
// first, Z_ipr(G) = sum_a Re-part of (phase_a(G) * Y_{m(a,ipr)}(G)), gives
// tsum_ij = 4 sum_s sum_ipr sum_G dtwnl_s,ipr,ij(G) Z_ipr(G).
// Block (chunk of G, species); partial[(s*nchunk + chunk)*6 + ij].  The
// projectors of a species are taken 16 at a time (Z in registers), each with
// the l of its column (m_l).
template <typename T>
__global__ __launch_bounds__(256) void k_nl_stress_y(
    const T* __restrict__ Y, long lda, int ngw, const double* __restrict__ gx,
    const double* __restrict__ tau, const int32_t* __restrict__ sp_first,
    const int32_t* __restrict__ sp_npr, const int64_t* __restrict__ dt_off,
    const double* __restrict__ dtwnl, const int32_t* __restrict__ a_m0,
    const int32_t* __restrict__ a_stride, const int32_t* __restrict__ m_l, int chunk,
    int halves, long mhalf, double* __restrict__ partial) {
  constexpr int ZB = 16;
  const int s = blockIdx.y;
  const int a0 = sp_first[s], a1 = sp_first[s + 1], npr = sp_npr[s];
  const double* dt = dtwnl + dt_off[s];
  const int g0 = blockIdx.x * chunk, g1 = min(ngw, g0 + chunk);
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int ig = g0 + threadIdx.x; ig < g1; ig += 256) {
    const double x = gx[ig], y = gx[ngw + ig], z = gx[2 * ngw + ig];
    for (int p0 = 0; p0 < npr; p0 += ZB) {
      double Z[ZB];
#pragma unroll
      for (int k = 0; k < ZB; k++) Z[k] = 0.0;
      for (int a = a0; a < a1; a++) {
        double sn, cs;
        sincos(-(x * tau[3 * a] + y * tau[3 * a + 1] + z * tau[3 * a + 2]), &sn, &cs);
        const long st = a_stride[a];
        // halves = 2 (k != 0): the i anl columns mhalf further, with their
        // own phase index in m_l
        for (int h = 0; h < halves; h++) {
        const long mb = a_m0[a] + h * mhalf;
#pragma unroll
        for (int k = 0; k < ZB; k++) {
          const int ipr = p0 + k;
          if (ipr < npr) {
            const long m = mb + ipr * st;
            const T* yp = Y + m * lda + 2 * ig;
            const double yr = (double)yp[0], yi = (double)yp[1];
            // Re(conj(phase) Y) for the phase (-i)^l (c + i s):
            // l = 0 (c, s), 1 (s, -c), 2 (-c, -s), 3 (-s, c)
            const int l = m_l[m];
            Z[k] += l == 0 ? cs * yr + sn * yi
                           : (l == 1 ? sn * yr - cs * yi
                                     : (l == 2 ? -(cs * yr + sn * yi) : cs * yi - sn * yr));
          }
        }
        }
      }
#pragma unroll
      for (int k = 0; k < ZB; k++)
        if (p0 + k < npr)
#pragma unroll
          for (int ij = 0; ij < 6; ij++)
            v[ij] += dt[(long)(ij + 6 * (p0 + k)) * ngw + ig] * Z[k];
    }
  }
  block_reduce_store<6>(v, partial + ((long)s * gridDim.x + blockIdx.x) * 6);
}

// out6[ij] = 4 * sum of the partials (fixed order)
__global__ void k_nl_stress_final(const double* __restrict__ partial, int nblk,
                                  double* __restrict__ out6) {
  const int ij = threadIdx.x;
  if (ij >= 6) return;
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += partial[(long)b * 6 + ij];
  out6[ij] = 4.0 * s;
}

template <typename T>
void launch_nl_stress_y(const T* Y, long lda, int ngw, const double* gx, const double* tau,
                        int nsp, const int32_t* sp_first, const int32_t* sp_npr,
                        const int64_t* dt_off, const double* dtwnl, const int32_t* a_m0,
                        const int32_t* a_stride, const int32_t* m_l, double* partial,
                        int nchunk, double* out6, hipStream_t s, int halves, long mhalf) {
  const int chunk = (ngw + nchunk - 1) / nchunk;
  hipLaunchKernelGGL(k_nl_stress_y<T>, dim3(nchunk, nsp), dim3(256), 0, s, Y, lda, ngw, gx, tau,
                     sp_first, sp_npr, dt_off, dtwnl, a_m0, a_stride, m_l, chunk, halves, mhalf,
                     partial);
  hipLaunchKernelGGL(k_nl_stress_final, dim3(1), dim3(64), 0, s, partial, nchunk * nsp, out6);
}
template void launch_nl_stress_y<double>(const double*, long, int, const double*, const double*,
                                         int, const int32_t*, const int32_t*, const int64_t*,
                                         const double*, const int32_t*, const int32_t*,
                                         const int32_t*, double*, int, double*, hipStream_t,
                                         int, long);
template void launch_nl_stress_y<float>(const float*, long, int, const double*, const double*,
                                        int, const int32_t*, const int32_t*, const int64_t*,
                                        const double*, const int32_t*, const int32_t*,
                                        const int32_t*, double*, int, double*, hipStream_t,
                                        int, long);

// ------------------------------------------- Hartree pseudocharge stress
// The species part of sigma_ehart (src/EnergyFunctional.cpp:649-676):
// sum_s 2 (rcps_s^2 / 4) Re(conj(rhogt) sfac_s) rhops_s g2i G_i G_j, and with
// a core density - sum_s Re(conj(vxc_g) sfac_s) drhoc_s G_i G_j, where
// drhoc_s = drhocoreg gi / (4 pi omega) (:650-659).
__global__ __launch_bounds__(256) void k_hart_ps_stress(
    const double* __restrict__ rhogt, const double* __restrict__ sfac,
    const double* __restrict__ rhops, const double* __restrict__ g2i,
    const double* __restrict__ gx, int ngv, int nsp, const double* __restrict__ fac2,
    const double* __restrict__ vxcg, const double* __restrict__ drhoc,
    double* __restrict__ partial) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (ig < ngv) {
    const double tr = rhogt[2 * ig], ti = rhogt[2 * ig + 1];
    double t = 0.0;
    for (int is = 0; is < nsp; is++) {
      const double* sf = sfac + 2 * (long)is * ngv;
      t += fac2[is] * (tr * sf[2 * ig] + ti * sf[2 * ig + 1]) * rhops[(long)is * ngv + ig];
    }
    t *= g2i[ig];
    if (vxcg) {
      const double vr = vxcg[2 * ig], vi = vxcg[2 * ig + 1];
      for (int is = 0; is < nsp; is++) {
        const double* sf = sfac + 2 * (long)is * ngv;
        t -= (vr * sf[2 * ig] + vi * sf[2 * ig + 1]) * drhoc[(long)is * ngv + ig];
      }
    }
    const double x = gx[ig], y = gx[ngv + ig], z = gx[2 * ngv + ig];
    v[0] = t * x * x; v[1] = t * y * y; v[2] = t * z * z;
    v[3] = t * x * y; v[4] = t * y * z; v[5] = t * x * z;
  }
  block_reduce_store<6>(v, partial + (long)blockIdx.x * 6);
}

void launch_hart_ps_stress(const double* rhogt, const double* sfac, const double* rhops,
                           const double* g2i, const double* gx, int ngv, int nsp,
                           const double* fac2, const double* vxcg, const double* drhoc,
                           double* partial, double* out6, hipStream_t s) {
  const int nb = (ngv + 255) / 256;
  hipLaunchKernelGGL(k_hart_ps_stress, dim3(nb), dim3(256), 0, s, rhogt, sfac, rhops, g2i, gx,
                     ngv, nsp, fac2, vxcg, drhoc, partial);
  reduce_partials(partial, nb, 6, out6, s);
}

}  // namespace qbg

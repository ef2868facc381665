// energy.cpp -- oracle restatement of EnergyFunctional, NonLocalPotential,
// StructureFactor, ChargeDensity and SlaterDet (density, rs_mul_add, gram).
// TEST INFRASTRUCTURE ONLY (see qbo.h).
#include "qbo.h"
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <algorithm>

namespace qbo {

static const double fpi = 4.0 * M_PI;

// ---------------------------------------------------------------- setup
void System::setup() {
  const Cell* ref = has_ref ? &refcell : nullptr;
  if (kpts.size() / 3 > 1) {           // a set: every k-point on the complex basis
    for (int i = 0; i < 3; i++) kpoint[i] = kpts[i];
    force_complex = true;
  }
  wb.resize(cell, ref, ecut, kpoint, force_complex);
  vb.resize(cell, ref, 4.0 * ecut);          // src/ChargeDensity.cpp:37
  int n[3];
  for (int i = 0; i < 3; i++) {              // src/ChargeDensity.cpp:44-49
    n[i] = vb.np[i] + 2;
    while (!factorizable(n[i])) n[i] += 2;
    if (grid[i] > 0) {                       // FourierTransform(basis, np0, np1, np2)
      if (grid[i] < vb.np[i] || !factorizable(grid[i]))
        throw std::invalid_argument("forced grid does not hold the density basis");
      n[i] = grid[i];
    }
  }
  vft.init(vb, n[0], n[1], n[2]);
  ft.init(wb, n[0], n[1], n[2]);             // src/ChargeDensity.cpp:98
  // further k-points of a set: their bases and transforms on the same grid
  kset.clear();
  const int nkp = (int)kpts.size() / 3;
  if (nkp > 1) {
    kset.resize(nkp);
    for (int ik = 0; ik < nkp; ik++) kset[ik].weight = kwts[ik];
    for (int ik = 1; ik < nkp; ik++) {
      kset[ik].wb.resize(cell, ref, ecut, &kpts[3 * ik], true);
      kset[ik].ft.init(kset[ik].wb, n[0], n[1], n[2]);
    }
  }
  for (auto& s : sp) s.initialize();
  eself_total = 0.0;
  for (size_t is = 0; is < sp.size(); is++)
    eself_total += (tau[is].size() / 3) * sp[is].eself();
  const long N = (long)n[0] * n[1] * n[2];
  rhor.assign(N, 0.0);
  v_r.assign(N, 0.0);
  vxc_r.assign(N, 0.0);
  rhor_dn.assign(N, 0.0);
  v_r_dn.assign(N, 0.0);
  vxc_r_dn.assign(N, 0.0);
  rhog.assign(vb.size, 0.0);
  rhogt.assign(vb.size, 0.0);
  rhoelg.assign(vb.size, 0.0);
  core = false;                              // src/EnergyFunctional.cpp:135-163
  for (auto& q : sp) core |= q.has_nlcc();
  rhocore_r.assign(core ? N : 0, 0.0);
  rhocore_g.assign(core ? vb.size : 0, 0.0);
  vxc_g.assign(core ? vb.size : 0, 0.0);
  cell_moved();
  atoms_moved();
}

// src/EnergyFunctional.cpp:1200-1245
void System::cell_moved() {
  const int ngv = vb.size;
  const double omega_inv = 1.0 / cell.volume;
  const int nsp = (int)sp.size();
  vps.assign(nsp, std::vector<double>(ngv));
  dvps.assign(nsp, std::vector<double>(ngv));
  rhops.assign(nsp, std::vector<double>(ngv));
  rhocore_sp.assign(core ? nsp : 0, std::vector<double>(ngv, 0.0));
  for (int is = 0; is < nsp; is++) {
#pragma omp parallel for schedule(static)
    for (int ig = 0; ig < ngv; ig++) {
      double v, dv;
      rhops[is][ig] = sp[is].rhopsg(vb.g[ig]) * omega_inv;
      sp[is].dvlocg(vb.g[ig], v, dv);
      vps[is][ig] = v * omega_inv;
      dvps[is][ig] = dv * omega_inv;
      if (core) {                               // :1224-1228
        double rc, drc;
        sp[is].drhocoreg(vb.g[ig], rc, drc);
        rhocore_sp[is][ig] = rc * omega_inv;
      }
    }
  }
  update_twnl();
  for (size_t ik = 1; ik < kset.size(); ik++) {
    swap_k(kset[ik]);
    update_twnl();
    swap_k(kset[ik]);
  }
}

// src/NonLocalPotential.cpp:54-232 (KB, wsg = 1) and :235-520 (l = 0, 1)
void System::update_twnl() {
  const int ngw = wb.size;
  const int nsp = (int)sp.size();
  const double s14pi = std::sqrt(1.0 / fpi);
  const double s34pi = std::sqrt(3.0 / fpi);
  twnl.assign(nsp, {});
  dtwnl.assign(nsp, {});
  lproj.assign(nsp, {});
  const double* kpg = wb.kpg.data();
  const double* kpgi = wb.kpgi.data();
  const double* kx = &wb.kpgx[0];
  const double* ky = &wb.kpgx[ngw];
  const double* kz = &wb.kpgx[2 * ngw];
  for (int is = 0; is < nsp; is++) {
    const Species& s = sp[is];
    const int npr = s.nlm;
    twnl[is].assign((size_t)npr * ngw, 0.0);
    dtwnl[is].assign((size_t)npr * 6 * ngw, 0.0);
    lproj[is].assign(npr, 0);
    int ilm = 0;
    for (int iop = 0; iop < s.nop; iop++) {
      const int l = iop;
      if (s.quad() && l == s.llocal) continue;   // no projectors for the local channel
      // one radial function per channel, or one per quadrature point; the
      // projector of (m, src) is ilm + m, or src + nquad (ilm + m) in the
      // quadrature form (src/NonLocalPotential.cpp:196-229)
      const int nsrc = s.quad() ? s.nquad : 1;
      for (int src = 0; src < nsrc; src++) {
      auto col = [&](int m) { return s.quad() ? src + s.nquad * (ilm + m) : ilm + m; };
      auto radial = [&](double g, double& v, double& dv) {
        if (s.quad()) s.quad_radial(l, src, g, v, dv);
        else s.dvnlg(iop, g, v, dv);
      };
      if (l == 0) {
        lproj[is][col(0)] = 0;
        double* t0 = &twnl[is][(size_t)ngw * col(0)];
        double* dt = &dtwnl[is][(size_t)ngw * 6 * col(0)];
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ngw; ig++) {
          double v, dv;
          radial(kpg[ig], v, dv);
          t0[ig] = s14pi * v;
          const double tgx = kx[ig], tgy = ky[ig], tgz = kz[ig];
          const double tmp = kpgi[ig] * s14pi * dv;
          dt[ig] = tmp * tgx * tgx;
          dt[ngw + ig] = tmp * tgy * tgy;
          dt[2 * ngw + ig] = tmp * tgz * tgz;
          dt[3 * ngw + ig] = tmp * tgx * tgy;
          dt[4 * ngw + ig] = tmp * tgy * tgz;
          dt[5 * ngw + ig] = tmp * tgx * tgz;
        }
      } else if (l == 1) {
        for (int m = 0; m < 3; m++) lproj[is][col(m)] = 1;
        double* t1 = &twnl[is][(size_t)ngw * col(0)];
        double* t2 = &twnl[is][(size_t)ngw * col(1)];
        double* t3 = &twnl[is][(size_t)ngw * col(2)];
        double* d1 = &dtwnl[is][(size_t)ngw * 6 * col(0)];
        double* d2 = &dtwnl[is][(size_t)ngw * 6 * col(1)];
        double* d3 = &dtwnl[is][(size_t)ngw * 6 * col(2)];
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ngw; ig++) {
          double v, dv;
          const double tg = kpg[ig];
          radial(tg, v, dv);
          const double tgx = kx[ig], tgy = ky[ig], tgz = kz[ig];
          const double tgx2 = tgx * tgx, tgy2 = tgy * tgy, tgz2 = tgz * tgz;
          const double tgi = kpgi[ig], tgi2 = tgi * tgi;
          const double y1 = s34pi * tgx * tgi;
          const double y2 = s34pi * tgy * tgi;
          const double y3 = s34pi * tgz * tgi;
          t1[ig] = y1 * v;
          t2[ig] = y2 * v;
          t3[ig] = y3 * v;
          const double fac1 = -y1 * (v - tg * dv) * tgi2;
          d1[ig] = fac1 * tgx2 + v * y1;
          d1[ngw + ig] = fac1 * tgy2;
          d1[2 * ngw + ig] = fac1 * tgz2;
          d1[3 * ngw + ig] = fac1 * tgx * tgy;
          d1[4 * ngw + ig] = fac1 * tgy * tgz;
          d1[5 * ngw + ig] = fac1 * tgx * tgz;
          const double fac2 = -y2 * (v - tg * dv) * tgi2;
          d2[ig] = fac2 * tgx2;
          d2[ngw + ig] = fac2 * tgy2 + v * y2;
          d2[2 * ngw + ig] = fac2 * tgz2;
          d2[3 * ngw + ig] = fac2 * tgx * tgy + v * y1;
          d2[4 * ngw + ig] = fac2 * tgy * tgz;
          d2[5 * ngw + ig] = fac2 * tgx * tgz;
          const double fac3 = -y3 * (v - tg * dv) * tgi2;
          d3[ig] = fac3 * tgx2;
          d3[ngw + ig] = fac3 * tgy2;
          d3[2 * ngw + ig] = fac3 * tgz2 + v * y3;
          d3[3 * ngw + ig] = fac3 * tgx * tgy;
          d3[4 * ngw + ig] = fac3 * tgy * tgz + v * y2;
          d3[5 * ngw + ig] = fac3 * tgx * tgz + v * y1;
        }
      } else if (l == 3) {   // (:963-1245): y9..y15 and their strain derivatives
        // The harmonics are the reference's.  Their strain derivatives take
        // the angular term with the sign of the l <= 2 branches,
        // y dv |G| u_i u_j + v (d_i y G_j + d_j y G_i) / 2 with d = d/dG
        // (the reference's l = 3 branch writes - v (...) / 2, opposite to its
        // own l = 0..2 expressions; the form here is the one the strain
        // finite differences of the energy confirm).
        const double s74pi = std::sqrt(7.0 / fpi), s32 = std::sqrt(1.5), s52 = std::sqrt(2.5),
                     s15 = std::sqrt(15.0);
        const double factor = 0.5 * s74pi;
        for (int m = 0; m < 7; m++) lproj[is][col(m)] = 3;
        double* t[7];
        double* d[7];
        for (int m = 0; m < 7; m++) {
          t[m] = &twnl[is][(size_t)ngw * col(m)];
          d[m] = &dtwnl[is][(size_t)ngw * 6 * col(m)];
        }
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ngw; ig++) {
          double v, dv;
          const double tg = kpg[ig];
          radial(tg, v, dv);
          const double tgx = kx[ig], tgy = ky[ig], tgz = kz[ig];
          const double tgi = kpgi[ig], tgi2 = tgi * tgi, tgi3 = tgi * tgi2;
          const double tgxx = tgx * tgx * tgi2, tgyy = tgy * tgy * tgi2, tgzz = tgz * tgz * tgi2;
          const double tgxy = tgx * tgy * tgi2, tgyz = tgy * tgz * tgi2, tgxz = tgx * tgz * tgi2;
          const double tgxxx = tgx * tgx * tgx * tgi3, tgyyy = tgy * tgy * tgy * tgi3,
                       tgzzz = tgz * tgz * tgz * tgi3;
          const double tgxyy = tgx * tgy * tgy * tgi3, tgxzz = tgx * tgz * tgz * tgi3,
                       tgyxx = tgy * tgx * tgx * tgi3, tgyzz = tgy * tgz * tgz * tgi3,
                       tgzxx = tgz * tgx * tgx * tgi3, tgzyy = tgz * tgy * tgy * tgi3,
                       tgxyz = tgx * tgy * tgz * tgi3;
          const double y[7] = {factor * s52 * (3.0 * tgyxx - tgyyy),
                               factor * s15 * 2.0 * tgxyz,
                               factor * s32 * (4.0 * tgyzz - tgyxx - tgyyy),
                               factor * (2.0 * tgzzz - 3.0 * (tgzxx + tgzyy)),
                               factor * s32 * (4.0 * tgxzz - tgxxx - tgxyy),
                               factor * s15 * (tgzxx - tgzyy),
                               factor * s52 * (tgxxx - 3.0 * tgxyy)};
          // gradients (d/dGx, d/dGy, d/dGz) of the cubic monomials over |G|^3
          const double dxxx[3] = {3.0 * tgxx * (1.0 - tgxx) * tgi, -3.0 * tgxx * tgxy * tgi,
                                  -3.0 * tgxx * tgxz * tgi};
          const double dyyy[3] = {-3.0 * tgyy * tgxy * tgi, 3.0 * tgyy * (1.0 - tgyy) * tgi,
                                  -3.0 * tgyy * tgyz * tgi};
          const double dzzz[3] = {-3.0 * tgzz * tgxz * tgi, -3.0 * tgzz * tgyz * tgi,
                                  3.0 * tgzz * (1.0 - tgzz) * tgi};
          const double dxyy[3] = {tgyy * (1.0 - 3.0 * tgxx) * tgi, tgxy * (2.0 - 3.0 * tgyy) * tgi,
                                  -3.0 * tgyy * tgxz * tgi};
          const double dxzz[3] = {tgzz * (1.0 - 3.0 * tgxx) * tgi, -3.0 * tgzz * tgxy * tgi,
                                  tgxz * (2.0 - 3.0 * tgzz) * tgi};
          const double dyxx[3] = {tgxy * (2.0 - 3.0 * tgxx) * tgi, tgxx * (1.0 - 3.0 * tgyy) * tgi,
                                  -3.0 * tgxx * tgyz * tgi};
          const double dyzz[3] = {-3.0 * tgzz * tgxy * tgi, tgzz * (1.0 - 3.0 * tgyy) * tgi,
                                  tgyz * (2.0 - 3.0 * tgzz) * tgi};
          const double dzxx[3] = {tgxz * (2.0 - 3.0 * tgxx) * tgi, -3.0 * tgxx * tgyz * tgi,
                                  tgxx * (1.0 - 3.0 * tgzz) * tgi};
          const double dzyy[3] = {-3.0 * tgyy * tgxz * tgi, tgyz * (2.0 - 3.0 * tgyy) * tgi,
                                  tgyy * (1.0 - 3.0 * tgzz) * tgi};
          const double dxyz[3] = {tgyz * (1 - 3.0 * tgxx) * tgi, tgxz * (1 - 3.0 * tgyy) * tgi,
                                  tgxy * (1 - 3.0 * tgzz) * tgi};
          double dyv[7][3];
          for (int c = 0; c < 3; c++) {
            dyv[0][c] = factor * s52 * (3.0 * dyxx[c] - dyyy[c]);
            dyv[1][c] = factor * s15 * 2.0 * dxyz[c];
            dyv[2][c] = factor * s32 * (4.0 * dyzz[c] - dyxx[c] - dyyy[c]);
            dyv[3][c] = factor * (2.0 * dzzz[c] - 3.0 * (dzxx[c] + dzyy[c]));
            dyv[4][c] = factor * s32 * (4.0 * dxzz[c] - dxxx[c] - dxyy[c]);
            dyv[5][c] = factor * s15 * (dzxx[c] - dzyy[c]);
            dyv[6][c] = factor * s52 * (dxxx[c] - 3.0 * dxyy[c]);
          }
          const double g3[3] = {tgx, tgy, tgz};
          const double uu[6] = {tgxx, tgyy, tgzz, tgxy, tgyz, tgxz};
          static const int IJ[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {0, 2}};
          for (int m = 0; m < 7; m++) {
            t[m][ig] = y[m] * v;
            for (int ij = 0; ij < 6; ij++) {
              const int a = IJ[ij][0], b = IJ[ij][1];
              d[m][ij * ngw + ig] = y[m] * dv * tg * uu[ij] +
                                    0.5 * v * (dyv[m][a] * g3[b] + dyv[m][b] * g3[a]);
            }
          }
        }
      } else {   // l == 2 (:627-760): y4..y8 and their strain derivatives
        const double s54pi = std::sqrt(5.0 / fpi), s3 = std::sqrt(3.0);
        const double s20pi = std::sqrt(20.0 * M_PI), s20pi3 = std::sqrt(20.0 * M_PI / 3.0);
        for (int m = 0; m < 5; m++) lproj[is][col(m)] = 2;
        double* t[5];
        double* d[5];
        for (int m = 0; m < 5; m++) {
          t[m] = &twnl[is][(size_t)ngw * col(m)];
          d[m] = &dtwnl[is][(size_t)ngw * 6 * col(m)];
        }
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ngw; ig++) {
          double v, dv;
          const double tg = kpg[ig];
          radial(tg, v, dv);
          const double tgx = kx[ig], tgy = ky[ig], tgz = kz[ig];
          const double tgi = kpgi[ig], tgi2 = tgi * tgi;
          // unit-vector products, ij order xx yy zz xy yz xz
          const double uu[6] = {tgx * tgx * tgi2, tgy * tgy * tgi2, tgz * tgz * tgi2,
                                tgx * tgy * tgi2, tgy * tgz * tgi2, tgx * tgz * tgi2};
          const double y[5] = {s54pi * 0.5 * (3.0 * uu[2] - 1.0), s54pi * 0.5 * s3 * (uu[0] - uu[1]),
                               s54pi * s3 * uu[3], s54pi * s3 * uu[4], s54pi * s3 * uu[5]};
          const double y1x = s34pi * tgx * tgi, y1y = s34pi * tgy * tgi, y1z = s34pi * tgz * tgi;
          // derivatives of the p harmonics: dX_ij = y1X uu_ij - (delta terms)
          double dx[6], dy[6], dz[6];
          for (int ij = 0; ij < 6; ij++) {
            dx[ij] = y1x * uu[ij];
            dy[ij] = y1y * uu[ij];
            dz[ij] = y1z * uu[ij];
          }
          dx[0] -= y1x;
          dy[1] -= y1y; dy[3] -= y1x;
          dz[2] -= y1z; dz[4] -= y1y; dz[5] -= y1x;
          for (int m = 0; m < 5; m++) t[m][ig] = y[m] * v;
          for (int ij = 0; ij < 6; ij++) {
            const double r = dv * tg * uu[ij];
            d[0][ij * ngw + ig] = -(v * s20pi * dz[ij] * y1z - y[0] * r);
            d[1][ij * ngw + ig] = -(v * s20pi3 * (y1x * dx[ij] - y1y * dy[ij]) - y[1] * r);
            d[2][ij * ngw + ig] = -(v * s20pi3 * (dx[ij] * y1y + y1x * dy[ij]) - y[2] * r);
            d[3][ij * ngw + ig] = -(v * s20pi3 * (dy[ij] * y1z + y1y * dz[ij]) - y[3] * r);
            d[4][ij * ngw + ig] = -(v * s20pi3 * (dx[ij] * y1z + y1x * dz[ij]) - y[4] * r);
          }
        }
      }
      }   // src
      ilm += 2 * l + 1;
    }
  }
}

// src/StructureFactor.cpp:70-162 and src/EnergyFunctional.cpp:1047-1197
void System::atoms_moved() {
  const int ngv = vb.size;
  const int nsp = (int)sp.size();
  const int* idx = vb.idx.data();
  sfac.assign(nsp, std::vector<cplx>(ngv, 0.0));
  const int k0min = vb.idxmin[0], k1min = vb.idxmin[1], k2min = vb.idxmin[2];
  const int r0 = vb.idxmax[0] - k0min + 1, r1 = vb.idxmax[1] - k1min + 1,
            r2 = vb.idxmax[2] - k2min + 1;
  for (int is = 0; is < nsp; is++) {
    const int na = (int)tau[is].size() / 3;
    std::vector<double> c0(r0), c1(r1), c2(r2), s0(r0), s1(r1), s2(r2);
    for (int ia = 0; ia < na; ia++) {
      const double* t = &tau[is][3 * ia];
      const double fac0 = cell.b[0][0] * t[0] + cell.b[0][1] * t[1] + cell.b[0][2] * t[2];
      const double fac1 = cell.b[1][0] * t[0] + cell.b[1][1] * t[1] + cell.b[1][2] * t[2];
      const double fac2 = cell.b[2][0] * t[0] + cell.b[2][1] * t[1] + cell.b[2][2] * t[2];
      for (int i = k0min; i <= vb.idxmax[0]; i++) {
        c0[i - k0min] = std::cos(i * fac0); s0[i - k0min] = std::sin(i * fac0);
      }
      for (int i = k1min; i <= vb.idxmax[1]; i++) {
        c1[i - k1min] = std::cos(i * fac1); s1[i - k1min] = std::sin(i * fac1);
      }
      for (int i = k2min; i <= vb.idxmax[2]; i++) {
        c2[i - k2min] = std::cos(i * fac2); s2[i - k2min] = std::sin(i * fac2);
      }
      cplx* sf = sfac[is].data();
#pragma omp parallel for schedule(static)
      for (int i = 0; i < ngv; i++) {
        const int kx = idx[3 * i] - k0min, ky = idx[3 * i + 1] - k1min,
                  kz = idx[3 * i + 2] - k2min;
        const double ca = c0[kx], cb = c1[ky], cc = c2[kz];
        const double sa = s0[kx], sb = s1[ky], sc = s2[kz];
        sf[i] += cplx(ca * cb * cc - sa * sb * cc - sa * cb * sc - ca * sb * sc,
                      sa * sb * sc - sa * cb * cc - ca * sb * cc - ca * cb * sc);
      }
    }
  }
  vion_local_g.assign(ngv, 0.0);
  dvion_local_g.assign(ngv, 0.0);
  rhopst.assign(ngv, 0.0);
  for (int is = 0; is < nsp; is++)
    for (int ig = 0; ig < ngv; ig++) {
      const cplx sg = sfac[is][ig];
      rhopst[ig] += sg * rhops[is][ig];
      vion_local_g[ig] += sg * vps[is][ig];
      dvion_local_g[ig] += sg * dvps[is][ig];
    }

  // esr, fion_esr, sigma_esr: :1097-1192
  const double omega_inv = 1.0 / cell.volume;
  auto len = [](const double* x) { return std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]); };
  const double d0 = 2.0 * M_PI / len(cell.b[0]);
  const double d1 = 2.0 * M_PI / len(cell.b[1]);
  const double d2 = 2.0 * M_PI / len(cell.b[2]);
  int nat = 0;
  std::vector<int> off(nsp);
  for (int is = 0; is < nsp; is++) { off[is] = nat; nat += (int)tau[is].size() / 3; }
  fion_esr.assign(3 * nat, 0.0);
  esr = 0.0;
  for (int i = 0; i < 6; i++) sigma_esr[i] = 0.0;
  for (int is1 = 0; is1 < nsp; is1++)
    for (int is2 = is1; is2 < nsp; is2++) {
      const double rc1 = sp[is1].rcps, rc2 = sp[is2].rcps;
      const double rcps12 = std::sqrt(rc1 * rc1 + rc2 * rc2);
      const int nc0 = (int)(8.0 * rcps12 / d0);
      const int nc1 = (int)(8.0 * rcps12 / d1);
      const int nc2 = (int)(8.0 * rcps12 / d2);
      const double z1 = sp[is1].zval, z2 = sp[is2].zval;
      const int na1 = (int)tau[is1].size() / 3, na2 = (int)tau[is2].size() / 3;
      for (int ia1 = 0; ia1 < na1; ia1++) {
        const int ia2min = (is1 == is2) ? ia1 : 0;
        for (int ia2 = ia2min; ia2 < na2; ia2++) {
          const bool same = (is1 == is2 && ia1 == ia2);
          double v12[3];
          for (int j = 0; j < 3; j++) v12[j] = tau[is1][3 * ia1 + j] - tau[is2][3 * ia2 + j];
          cell.fold_in_ws(v12);
          double* f1 = &fion_esr[3 * (off[is1] + ia1)];
          double* f2 = &fion_esr[3 * (off[is2] + ia2)];
          for (int ic0 = -nc0; ic0 <= nc0; ic0++)
            for (int ic1 = -nc1; ic1 <= nc1; ic1++)
              for (int ic2 = -nc2; ic2 <= nc2; ic2++) {
                if (same && ic0 == 0 && ic1 == 0 && ic2 == 0) continue;
                const double fac = same ? 0.5 : 1.0;
                double s[3];
                for (int j = 0; j < 3; j++)
                  s[j] = (ic0 * cell.a[0][j] + ic1 * cell.a[1][j]) + ic2 * cell.a[2][j];
                const double x12 = v12[0] + s[0], y12 = v12[1] + s[1], zz12 = v12[2] + s[2];
                const double r12 = std::sqrt(x12 * x12 + y12 * y12 + zz12 * zz12);
                const double arg = r12 / rcps12;
                const double esr_term = z1 * z2 * std::erfc(arg) / r12;
                esr += fac * esr_term;
                const double desr_erfc = 2.0 * z1 * z2 * std::exp(-arg * arg) / (rcps12 * std::sqrt(M_PI));
                const double desrdr = -fac * (esr_term + desr_erfc) / (r12 * r12);
                f1[0] -= desrdr * x12; f2[0] += desrdr * x12;
                f1[1] -= desrdr * y12; f2[1] += desrdr * y12;
                f1[2] -= desrdr * zz12; f2[2] += desrdr * zz12;
                sigma_esr[0] += desrdr * x12 * x12;
                sigma_esr[1] += desrdr * y12 * y12;
                sigma_esr[2] += desrdr * zz12 * zz12;
                sigma_esr[3] += desrdr * x12 * y12;
                sigma_esr[4] += desrdr * y12 * zz12;
                sigma_esr[5] += desrdr * x12 * zz12;
              }
        }
      }
    }
  for (int i = 0; i < 6; i++) sigma_esr[i] *= -omega_inv;
}

// ------------------------------------------------------------ SlaterDet
// src/SlaterDet.cpp:225-298 (real basis: two bands per transform; complex
// basis: one band per transform, rho += w occ |psi|^2 / omega)
void compute_density(const FFT3& ft, const cplx* c, int ngw, int nst,
                     const double* occ, double weight, double omega, double* rho) {
  const long N = ft.np012();
  std::vector<cplx> tmp(N);
  const double omega_inv = 1.0 / omega;
  if (!ft.basis->real) {
    for (int n = 0; n < nst; n++) {
      const double fac = weight * omega_inv * occ[n];
      if (fac > 0.0) {
        ft.backward(c + (long)n * ngw, tmp.data());
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; i++) rho[i] += fac * std::norm(tmp[i]);
      }
    }
    return;
  }
  for (int n = 0; n < nst - 1; n += 2) {
    const double fac1 = weight * omega_inv * occ[n];
    const double fac2 = weight * omega_inv * occ[n + 1];
    if (fac1 + fac2 > 0.0) {
      ft.backward(c + (long)n * ngw, c + (long)(n + 1) * ngw, tmp.data());
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; i++) {
        const double p1 = tmp[i].real(), p2 = tmp[i].imag();
        rho[i] += fac1 * p1 * p1 + fac2 * p2 * p2;
      }
    }
  }
  if (nst % 2 != 0) {
    const int n = nst - 1;
    const double fac1 = weight * omega_inv * occ[n];
    if (fac1 > 0.0) {
      ft.backward(c + (long)n * ngw, tmp.data());
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; i++) {
        const double p1 = tmp[i].real();
        rho[i] += fac1 * p1 * p1;
      }
    }
  }
}

// src/SlaterDet.cpp:410-477 (real basis: band pairs; complex basis: one
// band per transform)
void rs_mul_add(const FFT3& ft, const double* v, const cplx* c, int ngw,
                int nst, cplx* hpsi) {
  const long N = ft.np012();
  std::vector<cplx> tmp(N), c1(ngw), c2(ngw);
  if (!ft.basis->real) {
    for (int n = 0; n < nst; n++) {
      ft.backward(c + (long)n * ngw, tmp.data());
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; i++) tmp[i] *= v[i];
      ft.forward(tmp.data(), c1.data());
      cplx* h1 = hpsi + (long)n * ngw;
      for (int ig = 0; ig < ngw; ig++) h1[ig] += c1[ig];
    }
    return;
  }
  for (int n = 0; n < nst - 1; n += 2) {
    ft.backward(c + (long)n * ngw, c + (long)(n + 1) * ngw, tmp.data());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) tmp[i] *= v[i];
    ft.forward(tmp.data(), c1.data(), c2.data());
    cplx* h1 = hpsi + (long)n * ngw;
    cplx* h2 = hpsi + (long)(n + 1) * ngw;
    for (int ig = 0; ig < ngw; ig++) { h1[ig] += c1[ig]; h2[ig] += c2[ig]; }
  }
  if (nst % 2 != 0) {
    const int n = nst - 1;
    ft.backward(c + (long)n * ngw, tmp.data());
    for (long i = 0; i < N; i++) tmp[i] *= v[i];
    ft.forward(tmp.data(), c1.data());
    cplx* h1 = hpsi + (long)n * ngw;
    for (int ig = 0; ig < ngw; ig++) h1[ig] += c1[ig];
  }
}

// src/SlaterDet.cpp:1335-1356
void cleanup(cplx* c, int ngw, int nst) {
  for (int n = 0; n < nst; n++) c[(long)n * ngw] = cplx(c[(long)n * ngw].real(), 0.0);
}

// src/SlaterDet.cpp:480-534 (real basis): S = 2 P^T P - p0 p0^T,
// S = L L^T, C <- C L^-T
int gram(cplx* c, int ngw, int nst) {
  cleanup(c, ngw, nst);
  const long m = 2L * ngw;
  const double* p = reinterpret_cast<const double*>(c);
  std::vector<double> S((size_t)nst * nst, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nst; i++)
    for (int j = 0; j <= i; j++) {
      const double* a = p + (long)i * m;
      const double* b = p + (long)j * m;
      double s = 0.0;
      for (long k = 0; k < m; k++) s += a[k] * b[k];
      S[(size_t)i * nst + j] = 2.0 * s - a[0] * b[0];
    }
  // Cholesky, lower, row-major S[i*nst+j] for j <= i
  for (int j = 0; j < nst; j++) {
    double d = S[(size_t)j * nst + j];
    for (int k = 0; k < j; k++) d -= S[(size_t)j * nst + k] * S[(size_t)j * nst + k];
    if (!(d > 0.0)) return j + 1;
    d = std::sqrt(d);
    S[(size_t)j * nst + j] = d;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < nst; i++) {
      double s = S[(size_t)i * nst + j];
      for (int k = 0; k < j; k++) s -= S[(size_t)i * nst + k] * S[(size_t)j * nst + k];
      S[(size_t)i * nst + j] = s / d;
    }
  }
  // X L^T = C, row by row of the real view: x_j = (c_j - sum_k<j L_jk x_k)/L_jj
  double* q = reinterpret_cast<double*>(c);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < m; r++) {
    for (int j = 0; j < nst; j++) {
      double s = q[r + (long)j * m];
      for (int k = 0; k < j; k++) s -= S[(size_t)j * nst + k] * q[r + (long)k * m];
      q[r + (long)j * m] = s / S[(size_t)j * nst + j];
    }
  }
  return 0;
}

// src/SlaterDet.cpp:523-532 (complex basis): S = C^H C (herk), S = L L^H
// (potrf, lower), C <- C L^-H (trsm 'r','l','c'); returns 0 or the failing
// pivot + 1.  c is band-major (c[ig + ngw*n] = C[ig][n]).
int gram_complex(cplx* c, int ngw, int nst) {
  std::vector<cplx> S((size_t)nst * nst, cplx(0.0));   // S[i*nst + j], j <= i
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nst; i++)
    for (int j = 0; j <= i; j++) {
      const cplx* a = c + (size_t)i * ngw;
      const cplx* b = c + (size_t)j * ngw;
      cplx sacc(0.0);
      for (int g = 0; g < ngw; g++) sacc += std::conj(a[g]) * b[g];
      S[(size_t)i * nst + j] = sacc;               // (C^H C)_ij
    }
  // Cholesky, lower: L_jj = sqrt(S_jj - sum_k |L_jk|^2),
  // L_ij = (S_ij - sum_k L_ik conj(L_jk)) / L_jj
  for (int j = 0; j < nst; j++) {
    double d = S[(size_t)j * nst + j].real();
    for (int k = 0; k < j; k++) d -= std::norm(S[(size_t)j * nst + k]);
    if (!(d > 0.0)) return j + 1;
    d = std::sqrt(d);
    S[(size_t)j * nst + j] = d;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < nst; i++) {
      cplx sacc = S[(size_t)i * nst + j];
      for (int k = 0; k < j; k++) sacc -= S[(size_t)i * nst + k] * std::conj(S[(size_t)j * nst + k]);
      S[(size_t)i * nst + j] = sacc / d;
    }
  }
  // X L^H = C row by row: x_j = (c_j - sum_{k<j} x_k conj(L_jk)) / L_jj
#pragma omp parallel for schedule(static)
  for (int g = 0; g < ngw; g++) {
    for (int j = 0; j < nst; j++) {
      cplx sacc = c[g + (size_t)j * ngw];
      for (int k = 0; k < j; k++) sacc -= c[g + (size_t)k * ngw] * std::conj(S[(size_t)j * nst + k]);
      c[g + (size_t)j * ngw] = sacc / S[(size_t)j * nst + j].real();
    }
  }
  return 0;
}

// ---------------------------------------------------------- ChargeDensity
// src/ChargeDensity.cpp:136-198
// (nspin 2: one density per spin from its bands; rhog is the transform of the
// total, rhog[0] + rhog[1] of src/EnergyFunctional.cpp:256-261)
void System::update_density(const cplx* c) {
  std::fill(rhor.begin(), rhor.end(), 0.0);
  std::fill(rhor_dn.begin(), rhor_dn.end(), 0.0);
  const int nup = nspin == 2 ? nst_up : nst;
  // each k-point with its weight (src/ChargeDensity.cpp:136-160)
  const int nk = kset.empty() ? 1 : (int)kset.size();
  std::vector<size_t> koff(nk);            // (before any k-point is swapped in)
  for (int ik = 0; ik < nk; ik++) koff[ik] = k_offset(ik);
  for (int ik = 0; ik < nk; ik++) {
    if (ik > 0) swap_k(kset[ik]);
    const double w = kset.empty() ? 1.0 : kset[ik].weight;
    const cplx* ck = c + koff[ik];
    compute_density(ft, ck, wb.size, nup, occ.data(), w, cell.volume, rhor.data());
    if (nspin == 2 && nst > nup)
      compute_density(ft, ck + (size_t)nup * wb.size, wb.size, nst - nup, occ.data() + nup, w,
                      cell.volume, rhor_dn.data());
    if (ik > 0) swap_k(kset[ik]);
  }
  const long N = ft.np012();
  const double omega = cell.volume;
  std::vector<cplx> rhotmp(N);
  double sum = 0.0;
  for (long i = 0; i < N; i++) {
    const double r = nspin == 2 ? rhor[i] + rhor_dn[i] : rhor[i];
    sum += r;
    rhotmp[i] = cplx(omega * r, 0.0);
  }
  total_charge = sum * omega / (double)N;
  vft.forward(rhotmp.data(), rhog.data());
}

// src/EnergyFunctional.cpp:231-394 and XCPotential.cpp:124-230, 467-516
// the core density of the current positions and spin count: rhocore_g =
// sum_is (omega / nspin) sfac rhocore_sp, rhocore_r = Re(backward) / omega
// (src/EnergyFunctional.cpp:1064-1083; computed there in atoms_moved)
void System::core_density() {
  const int ngv = vb.size;
  const double omega = cell.volume;
  const long N = vft.np012();
  std::fill(rhocore_g.begin(), rhocore_g.end(), cplx(0.0));
  const double spin_fac = omega / nspin;
  for (size_t is = 0; is < sp.size(); is++)
    for (int ig = 0; ig < ngv; ig++) rhocore_g[ig] += spin_fac * sfac[is][ig] * rhocore_sp[is][ig];
  std::vector<cplx> tmp(N);
  vft.backward(rhocore_g.data(), tmp.data());
  for (long i = 0; i < N; i++) rhocore_r[i] = tmp[i].real() / omega;
}

void System::update_vhxc(bool compute_stress) {
  const int ngv = vb.size;
  const double omega = cell.volume, omega_inv = 1.0 / omega;
  const long N = vft.np012();
  for (int ig = 0; ig < ngv; ig++) rhoelg[ig] = omega_inv * rhog[ig];
  if (core) core_density();
  // LDA, or LSDA for nspin 2 (XCPotential.cpp:181-226), on the valence
  // density plus each spin's share of the core density (:154-179)
  auto rc = [&](long i) { return core ? rhocore_r[i] : 0.0; };
  double esum = 0.0, dsum = 0.0;
  if (nspin == 2) {
    for (long i = 0; i < N; i++) {
      double e, vu, vd;
      const double ru = rhor[i] + rc(i), rd = rhor_dn[i] + rc(i);
      lsda(ru, rd, e, vu, vd);
      const double r = ru + rd;
      esum += r * e;
      dsum += r * e - ru * vu - rd * vd;
      vxc_r[i] = vu;
      vxc_r_dn[i] = vd;
    }
  } else {
    for (long i = 0; i < N; i++) {
      double e, v;
      const double r = rhor[i] + rc(i);
      lda_unpolarized(r, e, v);
      esum += r * e;
      dsum += r * (e - v);
      vxc_r[i] = v;
    }
  }
  exc = esum * omega / (double)N;
  dxc = dsum * omega / (double)N;
  if (core) {
    // vxc_g (spin average for nspin 2) and the core term of dxc:
    // nspin (2 sum_G Re(vxc_g* rhocore_g) - the G = 0 term)
    // (src/EnergyFunctional.cpp:275-306)
    std::vector<cplx> tmp(N), vg(ngv);
    for (long i = 0; i < N; i++) tmp[i] = vxc_r[i];
    vft.forward(tmp.data(), vxc_g.data());
    if (nspin == 2) {
      for (long i = 0; i < N; i++) tmp[i] = vxc_r_dn[i];
      vft.forward(tmp.data(), vg.data());
      for (int ig = 0; ig < ngv; ig++) vxc_g[ig] = 0.5 * (vxc_g[ig] + vg[ig]);
    }
    double sum = 0.0;
    for (int ig = 0; ig < ngv; ig++)
      sum += vxc_g[ig].real() * rhocore_g[ig].real() + vxc_g[ig].imag() * rhocore_g[ig].imag();
    sum = 2.0 * sum - std::real(std::conj(vxc_g[0]) * rhocore_g[0]);
    dxc += nspin * sum;
  }
  for (int i = 0; i < 6; i++) sigma_exc[i] = 0.0;
  if (compute_stress) {
    const double t = -dsum / (double)N;
    sigma_exc[0] = sigma_exc[1] = sigma_exc[2] = t;
  }
  // eps
  double s0 = 0.0;
  for (int ig = 0; ig < ngv; ig++)
    s0 += rhoelg[ig].real() * vion_local_g[ig].real() +
          rhoelg[ig].imag() * vion_local_g[ig].imag();
  s0 = 2.0 * s0 - std::real(std::conj(rhoelg[0]) * vion_local_g[0]);
  eps = s0 * omega;
  double ehsum = 0, ehesum = 0, ehepsum = 0, ehpsum = 0;
  const double* g2i = vb.g2i.data();
  for (int ig = 0; ig < ngv; ig++) {
    const cplx r = rhoelg[ig], rp = rhopst[ig];
    ehsum += std::norm(r + rp) * g2i[ig];
    ehesum += std::norm(r) * g2i[ig];
    ehepsum += 2.0 * std::real(std::conj(r) * rp * g2i[ig]);
    ehpsum += std::norm(rp) * g2i[ig];
    rhogt[ig] = r + rp;
  }
  ehart = omega * fpi * ehsum;
  ehart_e = omega * fpi * ehesum;
  ehart_ep = omega * fpi * ehepsum;
  ehart_p = omega * fpi * ehpsum;
  std::vector<cplx> vlocal_g(ngv), tmp_r(N);
  for (int ig = 0; ig < ngv; ig++) vlocal_g[ig] = vion_local_g[ig] + fpi * rhogt[ig] * g2i[ig];
  vft.backward(vlocal_g.data(), tmp_r.data());
  for (long i = 0; i < N; i++) v_r[i] = vxc_r[i] + tmp_r[i].real();
  if (nspin == 2)                               // :383-389
    for (long i = 0; i < N; i++) v_r_dn[i] = vxc_r_dn[i] + tmp_r[i].real();
}

// ---------------------------------------------------- blocked GEMM helpers
// C[i + j*ldc] (+)= alpha * sum_k A[k + i*lda] * B[k + j*ldb]   (TN, K-contiguous)
static void gemm_tn(int M, int N, long K, double alpha, const double* A, long lda,
                    const double* B, long ldb, double beta, double* C, int ldc) {
  std::vector<double> acc((size_t)M * N, 0.0);
  const long KB = 4096;
  for (long k0 = 0; k0 < K; k0 += KB) {
    const long k1 = std::min(K, k0 + KB);
#pragma omp parallel for schedule(static) collapse(2)
    for (int i = 0; i < M; i++)
      for (int j = 0; j < N; j++) {
        const double* a = A + i * lda;
        const double* b = B + j * ldb;
        double s = 0.0;
        for (long k = k0; k < k1; k++) s += a[k] * b[k];
        acc[(size_t)i + (size_t)j * M] += s;
      }
  }
  for (int j = 0; j < N; j++)
    for (int i = 0; i < M; i++)
      C[i + (long)j * ldc] = alpha * acc[(size_t)i + (size_t)j * M] +
                             (beta == 0.0 ? 0.0 : beta * C[i + (long)j * ldc]);
}

// C[k + j*ldc] += sum_i A[k + i*lda] * B[i + j*ldb]   (NN, C += A B)
static void gemm_nn_acc(long K, int N, int M, const double* A, long lda,
                        const double* B, int ldb, double* C, long ldc) {
  const long KB = 1024;
  const long nkb = (K + KB - 1) / KB;
#pragma omp parallel for schedule(static)
  for (long kb = 0; kb < nkb; kb++) {
    const long k0 = kb * KB, k1 = std::min(K, k0 + KB);
    for (int j = 0; j < N; j++) {
      double* c = C + j * ldc;
      for (int i = 0; i < M; i++) {
        const double f = B[i + (long)j * ldb];
        const double* a = A + i * lda;
        for (long k = k0; k < k1; k++) c[k] += a[k] * f;
      }
    }
  }
}

// -------------------------------------------------- NonLocalPotential
// src/NonLocalPotential.cpp:1258-2069 restated for the real basis, KB
// projectors with l <= 3.  The reference processes atoms in blocks of 32
// (:1274); the block size only reorders independent columns and is not
// restated.
// t (-i)^l (c + i s) as (re, im)
static inline void phase_l(int l, double t, double c, double s, double& re, double& im) {
  switch (l & 3) {
    case 0: re = t * c; im = t * s; break;
    case 1: re = t * s; im = -t * c; break;
    case 2: re = -t * c; im = -t * s; break;
    default: re = -t * s; im = t * c; break;
  }
}

// The complex basis (k != 0): anl and fnl complex, fnl = anl^H c (zgemm,
// :1429-1445), enl = sum occ wt/omega Re(conj(fnl) D fnl) (:1572-1595),
// H psi += anl fnl' (:1616-1630), forces from dfnl = (-i (k+G)_j anl)^H c
// (:1722-1775), stress from dfnl = (dtwnl_ij phase)^H c (:1990-2037, with
// dtwnl at k + G).
double System::nl_energy_complex(const cplx* c, bool compute_hpsi, cplx* hpsi,
                                 bool compute_forces, std::vector<double>& fion_enl,
                                 bool compute_stress, double* sigma_enl) {
  const int ngw = wb.size;
  const double omega_inv = 1.0 / cell.volume;
  const int nsp = (int)sp.size();
  double enl = 0.0;
  int nat = 0;
  for (int is = 0; is < nsp; is++) nat += (int)tau[is].size() / 3;
  fion_enl.assign(3 * nat, 0.0);
  const double* kx[3] = {&wb.kpgx[0], &wb.kpgx[ngw], &wb.kpgx[2 * ngw]};
  double tsum[6] = {0, 0, 0, 0, 0, 0};
  int atom_off = 0;
  for (int is = 0; is < nsp; is++) {
    const Species& s = sp[is];
    const int na = (int)tau[is].size() / 3;
    const int npr = s.nlm;
    const int M = npr * na;
    // anl[ig + ngw*(ia + ipr*na)] = t (-i)^l e^{-i(k+G).tau}; dir 0..2:
    // -i (k+G)_dir anl; dir 3 + ij: dtwnl_ij (-i)^l e^{-i(k+G).tau}
    std::vector<cplx> anl((size_t)ngw * M);
    auto build_anl = [&](int dir) {
#pragma omp parallel for schedule(static) collapse(2)
      for (int ipr = 0; ipr < npr; ipr++)
        for (int ia = 0; ia < na; ia++) {
          const double* t = dir < 3 ? &twnl[is][(size_t)ngw * ipr]
                                    : &dtwnl[is][(size_t)ngw * ((dir - 3) + 6 * ipr)];
          const double* tu = &tau[is][3 * ia];
          const int l = lproj[is][ipr];
          cplx* a = &anl[(size_t)ngw * (ia + ipr * na)];
          for (int ig = 0; ig < ngw; ig++) {
            const double arg = -(kx[0][ig] * tu[0] + kx[1][ig] * tu[1] + kx[2][ig] * tu[2]);
            const bool force = dir >= 0 && dir < 3;
            const int ll = force ? l + 1 : l;
            const double tt = force ? kx[dir][ig] * t[ig] : t[ig];
            double re, im;
            phase_l(ll, tt, std::cos(arg), std::sin(arg), re, im);
            a[ig] = cplx(re, im);
          }
        }
    };
    auto project = [&](std::vector<cplx>& f) {   // f = anl^H c
      f.assign((size_t)M * nst, cplx(0.0));
#pragma omp parallel for schedule(static) collapse(2)
      for (int n = 0; n < nst; n++)
        for (int i = 0; i < M; i++) {
          const cplx* a = &anl[(size_t)ngw * i];
          const cplx* cn = c + (size_t)ngw * n;
          cplx sacc(0.0);
          for (int ig = 0; ig < ngw; ig++) sacc += std::conj(a[ig]) * cn[ig];
          f[i + (size_t)n * M] = sacc;
        }
    };
    build_anl(-1);
    std::vector<cplx> fnl, fbuf((size_t)M * nst);
    project(fnl);
    for (int n = 0; n < nst; n++)
      for (int ipr = 0; ipr < npr; ipr++)
        for (int ia = 0; ia < na; ia++) {
          cplx sacc(0.0);
          for (int jpr = 0; jpr < npr; jpr++)
            sacc += s.dmatrix(ipr, jpr) * fnl[(ia + jpr * na) + (size_t)n * M];
          fbuf[(ia + ipr * na) + (size_t)n * M] = sacc;
        }
    for (int ipr = 0; ipr < npr; ipr++) {
      const double fac = 1.0 * omega_inv;
      for (int n = 0; n < nst; n++) {
        const double facn = fac * occ[n];
        for (int ia = 0; ia < na; ia++) {
          const size_t i = ia + ipr * na + (size_t)n * M;
          enl += facn * (fnl[i].real() * fbuf[i].real() + fnl[i].imag() * fbuf[i].imag());
          fnl[i] = fac * fbuf[i];
        }
      }
    }
    if (compute_hpsi) {
#pragma omp parallel for schedule(static)
      for (int n = 0; n < nst; n++)
        for (int i = 0; i < M; i++) {
          const cplx f = fnl[i + (size_t)n * M];
          const cplx* a = &anl[(size_t)ngw * i];
          cplx* h = hpsi + (size_t)ngw * n;
          for (int ig = 0; ig < ngw; ig++) h[ig] += a[ig] * f;
        }
    }
    if (compute_forces) {
      std::vector<cplx> dfnl;
      for (int j = 0; j < 3; j++) {
        build_anl(j);
        project(dfnl);
        for (int ipr = 0; ipr < npr; ipr++)
          for (int n = 0; n < nst; n++) {
            const double facn = 2.0 * occ[n];
            for (int ia = 0; ia < na; ia++) {
              const size_t i = ia + ipr * na + (size_t)n * M;
              fion_enl[3 * (atom_off + ia) + j] -=
                  facn * (fnl[i].real() * dfnl[i].real() + fnl[i].imag() * dfnl[i].imag());
            }
          }
      }
    }
    if (compute_stress) {
      std::vector<cplx> dfnl;
      for (int ij = 0; ij < 6; ij++) {
        build_anl(3 + ij);
        project(dfnl);
        for (int n = 0; n < nst; n++) {
          const double facn = 2.0 * occ[n];
          for (int i = 0; i < M; i++) {
            const size_t q = i + (size_t)n * M;
            tsum[ij] += facn * (fnl[q].real() * dfnl[q].real() + fnl[q].imag() * dfnl[q].imag());
          }
        }
      }
    }
    atom_off += na;
  }
  if (compute_stress) {   // :2057-2066
    for (int i = 0; i < 3; i++) sigma_enl[i] = (enl + tsum[i]) * omega_inv;
    for (int i = 3; i < 6; i++) sigma_enl[i] = tsum[i] * omega_inv;
  }
  return enl;
}

double System::nl_energy(const cplx* c, bool compute_hpsi, cplx* hpsi,
                         bool compute_forces, std::vector<double>& fion_enl,
                         bool compute_stress, double* sigma_enl) {
  if (!wb.real) {
    for (int i = 0; i < 6; i++) sigma_enl[i] = 0.0;
    return nl_energy_complex(c, compute_hpsi, hpsi, compute_forces, fion_enl, compute_stress,
                             sigma_enl);
  }
  const int ngw = wb.size;
  const long K = 2L * ngw;
  const double omega_inv = 1.0 / cell.volume;
  const int nsp = (int)sp.size();
  double enl = 0.0;
  int nat = 0;
  for (int is = 0; is < nsp; is++) nat += (int)tau[is].size() / 3;
  fion_enl.assign(3 * nat, 0.0);
  for (int i = 0; i < 6; i++) sigma_enl[i] = 0.0;
  double tsum[6] = {0, 0, 0, 0, 0, 0};
  const double* kx[3] = {&wb.kpgx[0], &wb.kpgx[ngw], &wb.kpgx[2 * ngw]};
  const double* cp = reinterpret_cast<const double*>(c);
  int atom_off = 0;
  for (int is = 0; is < nsp; is++) {
    const Species& s = sp[is];
    const int na = (int)tau[is].size() / 3;
    const int npr = s.nlm;
    const int M = npr * na;
    // eigr: gr = -kpgx . tau (:1314-1339)
    std::vector<double> cgr((size_t)na * ngw), sgr((size_t)na * ngw);
#pragma omp parallel for schedule(static)
    for (int ia = 0; ia < na; ia++) {
      const double* t = &tau[is][3 * ia];
      for (int ig = 0; ig < ngw; ig++) {
        const double arg = -(kx[0][ig] * t[0] + kx[1][ig] * t[1] + kx[2][ig] * t[2]);
        sgr[(size_t)ia * ngw + ig] = std::sin(arg);
        cgr[(size_t)ia * ngw + ig] = std::cos(arg);
      }
    }
    // anl[2*ig + K*(ia + ipr*na)] (:1341-1395)
    std::vector<double> anl((size_t)K * M);
    auto build_anl = [&](int dir) {   // dir < 0: anl; dir = 0,1,2: -i G_dir anl
#pragma omp parallel for schedule(static) collapse(2)
      for (int ipr = 0; ipr < npr; ipr++)
        for (int ia = 0; ia < na; ia++) {
          const double* t = &twnl[is][(size_t)ngw * ipr];
          const int l = lproj[is][ipr];
          double* a = &anl[(size_t)K * (ia + ipr * na)];
          const double* cc = &cgr[(size_t)ia * ngw];
          const double* ss = &sgr[(size_t)ia * ngw];
          for (int ig = 0; ig < ngw; ig++) {
            // anl = t (-i)^l e^{-iG.tau} (:1341-1395); -i G_dir anl for the
            // forces (:1656-1700): the phase of l + 1 with t G_dir
            const int ll = dir < 0 ? l : l + 1;
            const double tt = dir < 0 ? t[ig] : kx[dir][ig] * t[ig];
            phase_l(ll, tt, cc[ig], ss[ig], a[2 * ig], a[2 * ig + 1]);
          }
        }
    };
    build_anl(-1);
    // fnl = 2 (anl^T c) - anl0 c0 (:1399-1464)
    std::vector<double> fnl((size_t)M * nst), fbuf((size_t)M * nst);
    gemm_tn(M, nst, K, 1.0, anl.data(), K, cp, K, 0.0, fnl.data(), M);
    for (int j = 0; j < nst; j++)
      for (int i = 0; i < M; i++)
        fnl[i + (size_t)j * M] += -0.5 * anl[(size_t)K * i] * cp[(size_t)K * j];
    for (auto& x : fnl) x *= 2.0;
    // D matrix (:1469-1507): F'_I = sum_J D_IJ F_J, diagonal in atom
    for (int j = 0; j < nst; j++)
      for (int ipr = 0; ipr < npr; ipr++)
        for (int ia = 0; ia < na; ia++) {
          double sacc = 0.0;
          for (int jpr = 0; jpr < npr; jpr++)
            sacc += s.dmatrix(ipr, jpr) * fnl[(ia + jpr * na) + (size_t)j * M];
          fbuf[(ia + ipr * na) + (size_t)j * M] = sacc;
        }
    // enl and fnl <- wt/omega * F' (:1550-1571), wt = 1
    for (int ipr = 0; ipr < npr; ipr++) {
      const double fac = 1.0 * omega_inv;
      for (int n = 0; n < nst; n++) {
        const double facn = fac * occ[n];
        for (int ia = 0; ia < na; ia++) {
          const size_t i = ia + ipr * na + (size_t)n * M;
          enl += facn * fnl[i] * fbuf[i];
          fnl[i] = fac * fbuf[i];
        }
      }
    }
    if (compute_hpsi)   // :1599-1614  cp += anl * fnl
      gemm_nn_acc(K, nst, M, anl.data(), K, fnl.data(), M,
                  reinterpret_cast<double*>(hpsi), K);
    if (compute_forces) {   // :1634-1784
      std::vector<double> dfnl((size_t)M * nst);
      for (int j = 0; j < 3; j++) {
        build_anl(j);
        gemm_tn(M, nst, K, 2.0, anl.data(), K, cp, K, 0.0, dfnl.data(), M);
        for (int ipr = 0; ipr < npr; ipr++)
          for (int n = 0; n < nst; n++) {
            const double facn = 2.0 * occ[n];
            for (int ia = 0; ia < na; ia++) {
              const size_t i = ia + ipr * na + (size_t)n * M;
              fion_enl[3 * (atom_off + ia) + j] -= facn * fnl[i] * dfnl[i];
            }
          }
      }
    }
    if (compute_stress) {   // :1786-2037, KB l = 0, 1
      std::vector<double> dfnl((size_t)M * nst);
      for (int ij = 0; ij < 6; ij++) {
        // danl = dtwnl_ij (-i)^l e^{-iG.tau}, the phases of anl
#pragma omp parallel for schedule(static) collapse(2)
        for (int ipr = 0; ipr < npr; ipr++)
          for (int ia = 0; ia < na; ia++) {
            const double* d = &dtwnl[is][(size_t)ngw * (ij + 6 * ipr)];
            const int l = lproj[is][ipr];
            double* a = &anl[(size_t)K * (ia + ipr * na)];
            const double* cc = &cgr[(size_t)ia * ngw];
            const double* ss = &sgr[(size_t)ia * ngw];
            for (int ig = 0; ig < ngw; ig++)
              phase_l(l, d[ig], cc[ig], ss[ig], a[2 * ig], a[2 * ig + 1]);
          }
        // dfnl = 2 danl^T c, no G=0 correction (dtwnl(G=0) = 0)
        gemm_tn(M, nst, K, 2.0, anl.data(), K, cp, K, 0.0, dfnl.data(), M);
        for (int n = 0; n < nst; n++) {
          const double facn = 2.0 * occ[n];
          for (int i = 0; i < M; i++)
            tsum[ij] += facn * fnl[i + (size_t)n * M] * dfnl[i + (size_t)n * M];
        }
      }
    }
    atom_off += na;
  }
  if (compute_stress) {   // :2057-2066
    for (int i = 0; i < 3; i++) sigma_enl[i] = (enl + tsum[i]) * omega_inv;
    for (int i = 3; i < 6; i++) sigma_enl[i] = tsum[i] * omega_inv;
  }
  return enl;
}

// src/EnergyFunctional.cpp:397-1044 (nspin = 1, one k-point, LDA, no
// confinement, no external potential or fields)
// the k-point set: exchange k-point k's basis, transform and tables with the
// members (fixing the transforms' basis pointers)
void System::swap_k(KState& k) {
  std::swap(wb, k.wb);
  std::swap(ft, k.ft);
  std::swap(twnl, k.twnl);
  std::swap(dtwnl, k.dtwnl);
  std::swap(lproj, k.lproj);
  ft.basis = &wb;
  k.ft.basis = &k.wb;
}

// (with no k-point swapped in: the members hold k-point 0)
size_t System::k_offset(int ik) const {
  if (kset.empty()) return 0;
  size_t off = 0;
  for (int j = 0; j < ik; j++) off += (size_t)nst * (j == 0 ? wb.size : kset[j].wb.size);
  return off;
}

double System::energy(const cplx* c, bool compute_hpsi, cplx* hpsi,
                      bool compute_forces, bool compute_stress, Results& r) {
  const int ngv = vb.size;
  const double omega = cell.volume, omega_inv = 1.0 / omega;
  int nat = 0;
  for (auto& t : tau) nat += (int)t.size() / 3;
  std::vector<double> fion(3 * nat, 0.0);
  const int nk = kset.empty() ? 1 : (int)kset.size();
  if (compute_hpsi) std::fill(hpsi, hpsi + k_offset(nk - 1) + (size_t)nst *
                                  (nk == 1 ? wb.size : kset[nk - 1].wb.size), cplx(0.0));
  // the parts of each k-point, weighted by its weight (:446-543, 703-716):
  // ekin (:441-571; fac = 1 for the real basis, G and -G times 1/2, and 1/2
  // for the complex basis), the nonlocal energy, forces and stress, and
  // H psi (unweighted, :789-838)
  double ts[7] = {0, 0, 0, 0, 0, 0, 0};
  double enl = 0.0, sig_enl[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> fenl(3 * nat, 0.0);
  std::vector<size_t> koff(nk);            // (before any k-point is swapped in)
  for (int ik = 0; ik < nk; ik++) koff[ik] = k_offset(ik);
  for (int ik = 0; ik < nk; ik++) {
    if (ik > 0) swap_k(kset[ik]);
    const double w = kset.empty() ? 1.0 : kset[ik].weight;
    const int ngw = wb.size;
    const cplx* ck = c + koff[ik];
    cplx* hk = compute_hpsi ? hpsi + koff[ik] : nullptr;
    const double kfac = wb.real ? 1.0 : 0.5;
    std::vector<double> psi2sum(ngw);
#pragma omp parallel for schedule(static)
    for (int ig = 0; ig < ngw; ig++) {
      double t = 0.0;
      for (int n = 0; n < nst; n++) t += kfac * occ[n] * std::norm(ck[ig + (size_t)n * ngw]);
      psi2sum[ig] = t;
    }
    double tk[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int ig = 0; ig < ngw; ig++) {
      const double p = psi2sum[ig];
      tk[0] += p * wb.kpg2[ig];
      if (compute_stress) {
        const double gx = wb.kpgx[ig], gy = wb.kpgx[ngw + ig], gz = wb.kpgx[2 * ngw + ig];
        const double f = 2.0 * p;
        tk[1] += f * gx * gx; tk[2] += f * gy * gy; tk[3] += f * gz * gz;
        tk[4] += f * gx * gy; tk[5] += f * gy * gz; tk[6] += f * gx * gz;
      }
    }
    for (int i = 0; i < 7; i++) ts[i] += w * tk[i];
    // NL (:686-742)
    std::vector<double> fk;
    double sk[6];
    enl += w * nl_energy(ck, compute_hpsi, hk, compute_forces, fk, compute_stress, sk);
    if (compute_forces)
      for (int i = 0; i < 3 * nat; i++) fenl[i] += w * fk[i];
    for (int i = 0; i < 6; i++) sig_enl[i] += w * sk[i];
    // H psi (:789-838)
    if (compute_hpsi) {
#pragma omp parallel for schedule(static)
      for (int n = 0; n < nst; n++)
        for (int ig = 0; ig < ngw; ig++)
          hk[ig + (size_t)n * ngw] += 0.5 * wb.kpg2[ig] * ck[ig + (size_t)n * ngw];
      // rs_mul_add with the potential of each band's spin (:831)
      const int nup = nspin == 2 ? nst_up : nst;
      rs_mul_add(ft, v_r.data(), ck, ngw, nup, hk);
      if (nspin == 2 && nst > nup)
        rs_mul_add(ft, v_r_dn.data(), ck + (size_t)nup * ngw, ngw, nst - nup,
                   hk + (size_t)nup * ngw);
    }
    if (ik > 0) swap_k(kset[ik]);
  }
  r.ekin = ts[0];
  for (int i = 0; i < 6; i++) r.sigma_ekin[i] = ts[i + 1] * omega_inv;
  for (int i = 0; i < 6; i++) { r.sigma_eps[i] = 0.0; r.sigma_ehart[i] = 0.0; }
  if (compute_stress) {   // :573-683
    double sm[6] = {0, 0, 0, 0, 0, 0};
    for (int ig = 0; ig < ngv; ig++) {
      const double fac = 2.0 * vb.gi[ig] * std::real(std::conj(rhoelg[ig]) * dvion_local_g[ig]);
      const double gx = vb.gx[ig], gy = vb.gx[ngv + ig], gz = vb.gx[2 * ngv + ig];
      sm[0] += fac * gx * gx; sm[1] += fac * gy * gy; sm[2] += fac * gz * gz;
      sm[3] += fac * gx * gy; sm[4] += fac * gy * gz; sm[5] += fac * gx * gz;
    }
    for (int i = 0; i < 6; i++) r.sigma_eps[i] = sm[i] + (i < 3 ? eps * omega_inv : 0.0);
    double sh[6] = {0, 0, 0, 0, 0, 0};
    for (int ig = 0; ig < ngv; ig++) {
      const double temp = std::norm(rhogt[ig]) * vb.g2i[ig] * vb.g2i[ig];
      const double gx = vb.gx[ig], gy = vb.gx[ngv + ig], gz = vb.gx[2 * ngv + ig];
      sh[0] += temp * gx * gx; sh[1] += temp * gy * gy; sh[2] += temp * gz * gz;
      sh[3] += temp * gx * gy; sh[4] += temp * gy * gz; sh[5] += temp * gx * gz;
    }
    for (size_t is = 0; is < sp.size(); is++) {
      const double fac2 = 2.0 * 0.25 * sp[is].rcps * sp[is].rcps;
      for (int ig = 0; ig < ngv; ig++) {
        const cplx sg = sfac[is][ig], rg = rhogt[ig];
        double temp = fac2 * (rg.real() * sg.real() + rg.imag() * sg.imag()) *
                      rhops[is][ig] * vb.g2i[ig];
        if (core) {                                 // :650-659
          double rhoc, drhoc;
          sp[is].drhocoreg(vb.g[ig], rhoc, drhoc);
          temp -= (vxc_g[ig].real() * sg.real() + vxc_g[ig].imag() * sg.imag()) * drhoc *
                  vb.gi[ig] * omega_inv / fpi;
        }
        const double gx = vb.gx[ig], gy = vb.gx[ngv + ig], gz = vb.gx[2 * ngv + ig];
        sh[0] += temp * gx * gx; sh[1] += temp * gy * gy; sh[2] += temp * gz * gz;
        sh[3] += temp * gx * gy; sh[4] += temp * gy * gz; sh[5] += temp * gx * gz;
      }
    }
    for (int i = 0; i < 6; i++)
      r.sigma_ehart[i] = (i < 3 ? ehart * omega_inv : 0.0) - 2.0 * fpi * sh[i];
  }
  r.enl = enl;
  for (int i = 0; i < 6; i++) r.sigma_enl[i] = sig_enl[i];
  if (compute_forces)
    for (int i = 0; i < 3 * nat; i++) fion[i] += fenl[i];
  r.eps = eps; r.ehart = ehart; r.ehart_e = ehart_e; r.ehart_ep = ehart_ep;
  r.ehart_p = ehart_p; r.exc = exc; r.dxc = dxc; r.esr = esr; r.eself = eself_total;
  r.ecoul = ehart + esr - eself_total;
  r.etotal = r.ekin + 0.0 + eps + r.enl + r.ecoul + exc + 0.0 + 0.0;
  // local forces (:841-930), eigr as products of per-axis tables
  if (compute_forces) {
    const int* idx = vb.idx.data();
    const int k0min = vb.idxmin[0], k1min = vb.idxmin[1], k2min = vb.idxmin[2];
    const int r0 = vb.idxmax[0] - k0min + 1, r1 = vb.idxmax[1] - k1min + 1,
              r2 = vb.idxmax[2] - k2min + 1;
    std::vector<cplx> vtemp(ngv);
    int atom_off = 0;
    for (size_t is = 0; is < sp.size(); is++) {
      for (int ig = 0; ig < ngv; ig++) {
        const double tmp = fpi * rhops[is][ig] * vb.g2i[ig];
        vtemp[ig] = tmp * std::conj(rhogt[ig]) + vps[is][ig] * std::conj(rhoelg[ig]);
        if (core) vtemp[ig] += std::conj(vxc_g[ig]) * rhocore_sp[is][ig];   // :855-856
      }
      const int na = (int)tau[is].size() / 3;
#pragma omp parallel for schedule(dynamic)
      for (int ia = 0; ia < na; ia++) {
        const double* t = &tau[is][3 * ia];
        const double fac0 = cell.b[0][0] * t[0] + cell.b[0][1] * t[1] + cell.b[0][2] * t[2];
        const double fac1 = cell.b[1][0] * t[0] + cell.b[1][1] * t[1] + cell.b[1][2] * t[2];
        const double fac2 = cell.b[2][0] * t[0] + cell.b[2][1] * t[1] + cell.b[2][2] * t[2];
        std::vector<double> c0(r0), c1(r1), c2(r2), s0(r0), s1(r1), s2(r2);
        for (int i = 0; i < r0; i++) { c0[i] = std::cos((i + k0min) * fac0); s0[i] = std::sin((i + k0min) * fac0); }
        for (int i = 0; i < r1; i++) { c1[i] = std::cos((i + k1min) * fac1); s1[i] = std::sin((i + k1min) * fac1); }
        for (int i = 0; i < r2; i++) { c2[i] = std::cos((i + k2min) * fac2); s2[i] = std::sin((i + k2min) * fac2); }
        double sum0 = 0, sum1 = 0, sum2 = 0;
        for (int ig = 0; ig < ngv; ig++) {
          const int kx = idx[3 * ig] - k0min, ky = idx[3 * ig + 1] - k1min, kz = idx[3 * ig + 2] - k2min;
          const double ca = c0[kx], cb = c1[ky], cc = c2[kz];
          const double sa = s0[kx], sb = s1[ky], sc = s2[kz];
          const double re = ca * cb * cc - sa * sb * cc - sa * cb * sc - ca * sb * sc;
          const double im = sa * sb * sc - sa * cb * cc - ca * sb * cc - ca * cb * sc;
          const double tmp = re * vtemp[ig].imag() + im * vtemp[ig].real();
          sum0 += tmp * vb.gx[ig];
          sum1 += tmp * vb.gx[ngv + ig];
          sum2 += tmp * vb.gx[2 * ngv + ig];
        }
        const double fac = -2.0 * omega;
        double* f = &fion[3 * (atom_off + ia)];
        const double* fe = &fion_esr[3 * (atom_off + ia)];
        f[0] += fe[0] + fac * sum0;
        f[1] += fe[1] + fac * sum1;
        f[2] += fe[2] + fac * sum2;
      }
      atom_off += na;
    }
  }
  r.fion = fion;
  for (int i = 0; i < 6; i++) {
    r.sigma_exc[i] = sigma_exc[i];
    r.sigma_esr[i] = sigma_esr[i];
    r.sigma[i] = compute_stress ? (r.sigma_ekin[i] + r.sigma_eps[i] + r.sigma_enl[i] +
                                   r.sigma_ehart[i] + r.sigma_exc[i] + r.sigma_esr[i])
                                : 0.0;
  }
  return r.etotal;
}

}  // namespace qbo

// ef.cpp -- EnergyFunctional drop-in: a whole system (Gamma, or k-points off
// Gamma on the complex basis) resident on
// one GPU (src/EnergyFunctional.cpp, src/ChargeDensity.cpp,
// src/NonLocalPotential.cpp, src/StructureFactor.cpp), plus SlaterDet::gram.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <unordered_map>
#include "ef.h"
#include "energy.h"
#include "internal.h"
#include "setup.h"

using namespace qbg;

namespace {

// scalar slots of the device result buffer
enum {
  S_LDA = 0,      // 4: sum rho e, sum rho (e-v), sum rho, -
  S_VHXC = 4,     // 18: see k_vhxc_g
  S_G0 = 22,      // 1: Re(conj(rhoelg[0]) vion[0])
  S_KIN = 24,     // 8: ekin, sigma_ekin[6], -
  S_ENL = 32,     // 1
  S_NLS = 33,     // 6: NL stress sums tsum_ij
  S_HPS = 39,     // 6: species part of the Hartree stress
  S_FLAGS = 45,   // 2: forces, stress computed by the staged evaluation
  S_CORE = 47,    // 1: sum_G w Re(conj(vxc_g) rhocore_g) (core density systems)
  S_COUNT = 48
};

template <typename T>
int dalloc(qbg_ctx* c, T** p, size_t n) {
  *p = nullptr;
  if (hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess)
    return c->fail(QBG_E_NOMEM, "hipMalloc failed (energy functional)");
  return QBG_OK;
}

template <typename T>
int dput(qbg_ctx* c, T** p, const std::vector<T>& h) {
  int rc = dalloc(c, p, h.size());
  if (rc) return rc;
  if (!h.empty() && hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    return c->fail(QBG_E_HIP, "hipMemcpy failed (energy functional)");
  return QBG_OK;
}

// copy host values into an existing device array of the same size
template <typename T>
int dset(qbg_ctx* c, T* d, const std::vector<T>& h) {
  if (!h.empty())
    QBG_HIP(c, hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return QBG_OK;
}

}  // namespace


// band range and first pair slot of spin s (pairs never straddle the spins)
static void spin_bands(const qbg_ef* e, int s, int& b0, int& nb, int& pair0) {
  const int nup = e->nspin == 2 ? e->nst_up : e->nst;
  b0 = s == 0 ? 0 : nup;
  nb = s == 0 ? nup : e->nst - nup;
  pair0 = s == 0 ? 0 : (nup + 1) / 2;
}

// density factors occ_n / omega in pair order, spin by spin (a missing
// second band of a spin's last pair contributes 0)
static std::vector<double> density_factors(const qbg_ef* e) {
  std::vector<double> f;
  if (e->cplx) {
    // one complex band per pair: both real parts carry its occupation (times
    // its k-point's weight in a set, k-point by k-point)
    for (int ik = 0; ik < e->nkp; ik++) {
      const double w = e->nkp > 1 ? e->kset[ik].weight : 1.0;
      for (int n = 0; n < e->nst; n++) {
        f.push_back(w * e->occ[n] / e->omega);
        f.push_back(w * e->occ[n] / e->omega);
      }
    }
    return f;
  }
  for (int s = 0; s < e->nspin; s++) {
    int b0, nb, p0;
    spin_bands(e, s, b0, nb, p0);
    const size_t base = f.size();
    f.resize(base + 2 * (size_t)((nb + 1) / 2), 0.0);
    for (int n = 0; n < nb; n++) f[base + n] = e->occ[b0 + n] / e->omega;
  }
  return f;
}

static std::vector<double> occ_half(const std::vector<double>& occ) {
  std::vector<double> h(occ);
  for (double& x : h) x *= 0.5;
  return h;
}

// a k-point set: exchange k-point k's basis and tables with the members
static void kswap(qbg_ef* e, qbg_ef::KPoint& k) {
  std::swap(e->wb, k.wb);
  std::swap(e->d_g2w, k.d_g2w);
  std::swap(e->d_gxw, k.d_gxw);
  std::swap(e->d_twnl, k.d_twnl);
  std::swap(e->d_dtwnl, k.d_dtwnl);
  std::swap(e->d_occ, k.d_occ);
  std::swap(e->d_occh, k.d_occh);
  std::swap(e->d_m_twnl, k.d_m_twnl);
  std::swap(e->d_dt_off, k.d_dt_off);
  std::swap(e->d_fold, k.d_fold);
  std::swap(e->d_unfold, k.d_unfold);
}

static void ef_free(qbg_ef* e) {
  void* p[] = {e->d_g2w, e->d_gxw, e->d_idxv, e->d_g2iv, e->d_giv, e->d_gxv, e->d_vps, e->d_dvps,
               e->d_rhops, e->d_twnl, e->d_tau, e->d_b, e->d_sftab, e->d_sfac, e->d_vion,
               e->d_dvion, e->d_rhopst, e->d_sp_first, e->d_atom_sp, e->d_rhor, e->d_vxc,
               e->d_vr, e->d_rhog, e->d_rhoelg, e->d_rhogt, e->d_vlocal, e->d_occ, e->d_fesr,
               e->d_fion, e->d_scal, e->d_partial, e->d_m_atom, e->d_m_twnl, e->d_m_l,
               e->d_m_ipr, e->d_m_npr, e->d_m_stride, e->d_m_dbase, e->d_m_wt, e->d_dmat,
               e->d_a_m0, e->d_a_npr, e->d_a_stride, e->d_anl, e->d_fnl, e->d_fout, e->d_work,
               e->d_dtwnl, e->d_dt_off, e->d_sp_npr, e->d_fac2, e->b_anl, e->b_c32,
               e->b_y32, e->b_fout, e->b_w32, e->b_part, e->d_lfgrp, e->d_vtemp, e->d_dfac,
               e->d_seg_ptr, e->d_seg_n, e->d_rhor2, e->d_vxc2, e->d_vr2, e->d_vloc,
               e->d_rhot, e->b_fpart, e->b_fred, e->d_rhoc, e->d_drhoc, e->d_rhocg,
               e->d_rhocr, e->d_vxcg, e->b_yh, e->d_fold, e->d_unfold, e->d_a_m0h,
               e->d_cpair, e->d_hpair, e->d_occh};
  for (void* q : p)
    if (q) (void)hipFree(q);
  for (qbg_ef::KPoint& k : e->kset) {
    void* kp[] = {k.d_g2w, k.d_gxw, k.d_twnl, k.d_dtwnl, k.d_occ, k.d_occh, k.d_m_twnl,
                  k.d_fold, k.d_unfold, k.d_dt_off};
    for (void* q : kp)
      if (q) (void)hipFree(q);
  }
  if (e->d_kacc) (void)hipFree(e->d_kacc);
  if (e->h_res) (void)hipHostFree(e->h_res);
  if (e->gw) qbg_grid_destroy(e->gw);
  if (e->gwf) qbg_grid_destroy(e->gwf);
  if (e->gv) qbg_grid_destroy(e->gv);
}

// esr, fion_esr, sigma_esr: real-space erfc lattice sums over pairs of ions
// (src/EnergyFunctional.cpp:1097-1192); O(nat^2) host work per geometry.
static void ef_esr(qbg_ef* e) {
  e->esr = 0.0;
  for (double& x : e->sigma_esr) x = 0.0;
  e->fesr.assign(3 * e->nat, 0.0);
  double d[3];
  for (int i = 0; i < 3; i++) {
    const double* b = e->cell.b[i];
    d[i] = 2.0 * M_PI / std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  }
  const double* a = &e->cell.a[0][0];
  for (int s1 = 0; s1 < e->nsp; s1++)
    for (int s2 = s1; s2 < e->nsp; s2++) {
      const double rc = std::sqrt(e->sp[s1].rcps * e->sp[s1].rcps + e->sp[s2].rcps * e->sp[s2].rcps);
      int nc[3];
      for (int i = 0; i < 3; i++) nc[i] = (int)(8.0 * rc / d[i]);
      const double zz = (double)e->sp[s1].zval * e->sp[s2].zval;
      for (int a1 = e->sp_first[s1]; a1 < e->sp_first[s1 + 1]; a1++)
        for (int a2 = (s1 == s2 ? a1 : e->sp_first[s2]); a2 < e->sp_first[s2 + 1]; a2++) {
          const bool same = a1 == a2;
          double v[3];
          for (int j = 0; j < 3; j++) v[j] = e->tau[3 * a1 + j] - e->tau[3 * a2 + j];
          e->cell.fold_in_ws(v);
          for (int i0 = -nc[0]; i0 <= nc[0]; i0++)
            for (int i1 = -nc[1]; i1 <= nc[1]; i1++)
              for (int i2 = -nc[2]; i2 <= nc[2]; i2++) {
                if (same && i0 == 0 && i1 == 0 && i2 == 0) continue;
                const double fac = same ? 0.5 : 1.0;
                double x[3];
                for (int j = 0; j < 3; j++)
                  x[j] = v[j] + ((i0 * a[j] + i1 * a[3 + j]) + i2 * a[6 + j]);
                const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const double arg = r / rc;
                const double term = zz * std::erfc(arg) / r;
                e->esr += fac * term;
                const double derfc = 2.0 * zz * std::exp(-arg * arg) / (rc * std::sqrt(M_PI));
                const double dr = -fac * (term + derfc) / (r * r);
                for (int j = 0; j < 3; j++) {
                  e->fesr[3 * a1 + j] -= dr * x[j];
                  e->fesr[3 * a2 + j] += dr * x[j];
                }
                e->sigma_esr[0] += dr * x[0] * x[0];
                e->sigma_esr[1] += dr * x[1] * x[1];
                e->sigma_esr[2] += dr * x[2] * x[2];
                e->sigma_esr[3] += dr * x[0] * x[1];
                e->sigma_esr[4] += dr * x[1] * x[2];
                e->sigma_esr[5] += dr * x[0] * x[2];
              }
        }
    }
  for (double& x : e->sigma_esr) x *= -1.0 / e->omega;
}

// EnergyFunctional::cell_moved's tables on the density basis of the current
// cell (src/EnergyFunctional.cpp:1200-1245): rhops = rhopsg/omega,
// vps = vlocg/omega, dvps = dvlocg/omega, species-major [is][ig]
// and, with core densities, rhoc = rhocoreg/omega (:1224-1228) and
// drhoc = drhocoreg gi / (4 pi omega) (the factor of the core term of the
// Hartree stress, :650-659)
static void ef_local_tables(const qbg_ef* e, std::vector<double>& vps, std::vector<double>& dvps,
                            std::vector<double>& rhops, std::vector<double>& rhoc,
                            std::vector<double>& drhoc) {
  const int nsp = e->nsp, ngv = e->vb.size;
  vps.assign((size_t)nsp * ngv, 0.0);
  dvps.assign((size_t)nsp * ngv, 0.0);
  rhops.assign((size_t)nsp * ngv, 0.0);
  rhoc.assign(e->core ? (size_t)nsp * ngv : 0, 0.0);
  drhoc.assign(e->core ? (size_t)nsp * ngv : 0, 0.0);
  for (int is = 0; is < nsp; is++)
    for (int ig = 0; ig < ngv; ig++) {
      double v, dv;
      const double oi = 1.0 / e->omega, g = e->vb.g[ig];
      rhops[(size_t)is * ngv + ig] = e->sp[is].rhopsg(g) * oi;
      e->sp[is].dvlocg(g, v, dv);
      vps[(size_t)is * ngv + ig] = v * oi;
      dvps[(size_t)is * ngv + ig] = dv * oi;
      if (e->core) {
        double rc, drc;
        e->sp[is].drhocoreg(g, rc, drc);
        rhoc[(size_t)is * ngv + ig] = rc * oi;
        drhoc[(size_t)is * ngv + ig] = drc * e->vb.gi[ig] * oi / (4.0 * M_PI);
      }
    }
}

// the core density of the current positions and spin count: rhocore_g =
// (omega / nspin) sum_s sfac_s rhoc_s, rhocore_r = Re(backward(rhocore_g)) /
// omega on the density grid (src/EnergyFunctional.cpp:1064-1083)
static int ef_core_density(qbg_ef* e) {
  if (!e->core) return QBG_OK;
  qbg_ctx* c = e->ctx;
  hipStream_t s = c->stream;
  launch_core_g(e->vb.size, e->nsp, e->d_sfac, e->d_rhoc, e->omega / e->nspin, e->d_rhocg, s);
  qbg_grid* g = e->gv;
  int rc = grid_ensure_ws(g, 1);
  if (rc) return rc;
  rc = zbwd(g, QBG_PREC_FP64, Z_SINGLE, e->d_rhocg, 0, 0, 0, 1, s);
  if (!rc) rc = xbwd(g, QBG_PREC_FP64, 1, s);
  if (!rc)
    rc = ypass(g, QBG_PREC_FP64, Y_BWD_REAL_ADD, 1, nullptr, nullptr, nullptr, e->d_rhocr,
               1.0 / e->omega, s);
  if (rc) return c->fail(rc, "core density: density-grid transform refused");
  return QBG_OK;
}

// real spherical harmonic of index ipr = l^2 + m (0 s; 1-3 p as x, y, z;
// 4-8 d as 3z^2-1, x^2-y^2, xy, yz, xz; 9-15 f as y(3x^2-y^2), xyz,
// y(4z^2-x^2-y^2), z(2z^2-3x^2-3y^2), x(4z^2-x^2-y^2), z(x^2-y^2),
// x(x^2-3y^2), the order and normalization of
// src/NonLocalPotential.cpp:250-1245) at the unit vector u: value and
// gradient with respect to u (as a polynomial in u)
static void harmonic(int ipr, const double* u, double& y, double* dy) {
  const double s14pi = std::sqrt(1.0 / (4.0 * M_PI)), s34pi = std::sqrt(3.0 / (4.0 * M_PI));
  const double d5 = std::sqrt(5.0 / (4.0 * M_PI)), d53 = d5 * std::sqrt(3.0);
  const double f = 0.5 * std::sqrt(7.0 / (4.0 * M_PI));
  const double f52 = f * std::sqrt(2.5), f15 = f * std::sqrt(15.0), f32 = f * std::sqrt(1.5);
  const double x = u[0], yy = u[1], z = u[2];
  dy[0] = dy[1] = dy[2] = 0.0;
  switch (ipr) {
    case 0: y = s14pi; break;
    case 1: case 2: case 3: y = s34pi * u[ipr - 1]; dy[ipr - 1] = s34pi; break;
    case 4: y = 0.5 * d5 * (3.0 * z * z - 1.0); dy[2] = 3.0 * d5 * z; break;
    case 5: y = 0.5 * d53 * (x * x - yy * yy); dy[0] = d53 * x; dy[1] = -d53 * yy; break;
    case 6: y = d53 * x * yy; dy[0] = d53 * yy; dy[1] = d53 * x; break;
    case 7: y = d53 * yy * z; dy[1] = d53 * z; dy[2] = d53 * yy; break;
    case 8: y = d53 * x * z; dy[0] = d53 * z; dy[2] = d53 * x; break;
    case 9:
      y = f52 * (3.0 * yy * x * x - yy * yy * yy);
      dy[0] = 6.0 * f52 * x * yy; dy[1] = 3.0 * f52 * (x * x - yy * yy);
      break;
    case 10:
      y = 2.0 * f15 * x * yy * z;
      dy[0] = 2.0 * f15 * yy * z; dy[1] = 2.0 * f15 * x * z; dy[2] = 2.0 * f15 * x * yy;
      break;
    case 11:
      y = f32 * (4.0 * yy * z * z - yy * x * x - yy * yy * yy);
      dy[0] = -2.0 * f32 * x * yy; dy[1] = f32 * (4.0 * z * z - x * x - 3.0 * yy * yy);
      dy[2] = 8.0 * f32 * yy * z;
      break;
    case 12:
      y = f * (2.0 * z * z * z - 3.0 * (z * x * x + z * yy * yy));
      dy[0] = -6.0 * f * z * x; dy[1] = -6.0 * f * z * yy;
      dy[2] = f * (6.0 * z * z - 3.0 * (x * x + yy * yy));
      break;
    case 13:
      y = f32 * (4.0 * x * z * z - x * x * x - x * yy * yy);
      dy[0] = f32 * (4.0 * z * z - 3.0 * x * x - yy * yy); dy[1] = -2.0 * f32 * x * yy;
      dy[2] = 8.0 * f32 * x * z;
      break;
    case 14:
      y = f15 * (z * x * x - z * yy * yy);
      dy[0] = 2.0 * f15 * z * x; dy[1] = -2.0 * f15 * z * yy; dy[2] = f15 * (x * x - yy * yy);
      break;
    default:
      y = f52 * (x * x * x - 3.0 * x * yy * yy);
      dy[0] = 3.0 * f52 * (x * x - yy * yy); dy[1] = -6.0 * f52 * x * yy;
      break;
  }
}

// NonLocalPotential::update_twnl on the wavefunction basis of the current
// cell (src/NonLocalPotential.cpp:235-1245, KB projectors l <= 3): twnl[is] =
// [ipr][ig] at offset sum_{js<is} npr_js ngw, dtwnl[is] = [ipr][ij][ig] at
// six times that offset.  twnl = y(u) v_l(|q|), u = q/|q|, q = k + G; the strain
// derivative is dtwnl_ij = y dv |G| u_i u_j - v sum_a dy/du_a (u_a u_i u_j -
// delta_aj u_i), ij = xx yy zz xy yz xz, which is the reference's
// expressions for l = 0, 1, 2 (:283-306, :426-512, :627-760) written once.
// For l = 3 the reference (:1166-1236) writes the angular term with the
// opposite sign, which the strain finite differences of the energy do not
// confirm; this form (the sign of its l <= 2 branches) is used for every l.
static void ef_twnl(const qbg_ef* e, std::vector<double>& twnl, std::vector<double>& dtwnl) {
  static const int IJ[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {0, 2}};
  const int ngw = e->wb.size;
  twnl.clear();
  dtwnl.clear();
  for (int is = 0; is < e->nsp; is++) {
    const host::Species& q = e->sp[is];
    const int npr = q.nlm;
    std::vector<int> pl, ph, pq;
    std::vector<double> pd;
    q.projectors(pl, ph, pq, pd);
    const size_t toff = twnl.size(), doff = dtwnl.size();
    twnl.resize(toff + (size_t)npr * ngw);
    dtwnl.resize(doff + (size_t)npr * 6 * ngw);
    double* t = &twnl[toff];
    double* d = &dtwnl[doff];
    // radial functions per G: one per channel, or one per (channel,
    // quadrature point) in the quadrature form
    const int nq = q.quad() ? q.nquad : 1;
    std::vector<double> rv((size_t)(q.lmax + 1) * nq), rdv((size_t)(q.lmax + 1) * nq);
    for (int ig = 0; ig < ngw; ig++) {
      // k + G (G itself at k = 0)
      const double tg = e->wb.kpg[ig], tgi = e->wb.kpgi[ig];
      const double u[3] = {e->wb.kpgx[ig] * tgi, e->wb.kpgx[ngw + ig] * tgi,
                           e->wb.kpgx[2 * ngw + ig] * tgi};
      for (int l = 0; l <= q.lmax; l++) {
        if (!q.quad()) {
          q.dvnlg(l, tg, rv[l], rdv[l]);
        } else if (l != q.llocal) {
          for (int k = 0; k < nq; k++)
            q.quad_radial(l, k, tg, rv[(size_t)l * nq + k], rdv[(size_t)l * nq + k]);
        }
      }
      for (int ipr = 0; ipr < npr; ipr++) {
        const int l = pl[ipr];
        const size_t ri = (size_t)l * nq + (pq[ipr] < 0 ? 0 : pq[ipr]);
        const double v = rv[ri], dv = rdv[ri];
        double y, dy[3];
        harmonic(ph[ipr], u, y, dy);
        t[(size_t)ipr * ngw + ig] = y * v;
        for (int ij = 0; ij < 6; ij++) {
          const int a = IJ[ij][0], b = IJ[ij][1];
          double sum = 0.0;
          for (int c = 0; c < 3; c++)
            if (dy[c] != 0.0) sum += dy[c] * (u[c] * u[a] * u[b] - (c == b ? u[a] : 0.0));
          d[((size_t)ipr * 6 + ij) * ngw + ig] = y * dv * tg * u[a] * u[b] - v * sum;
        }
      }
    }
  }
}

static int ef_atoms_moved(qbg_ef* e) {
  qbg_ctx* c = e->ctx;
  hipStream_t s = c->stream;
  QBG_HIP(c, hipMemcpyAsync(e->d_tau, e->tau.data(), sizeof(double) * e->tau.size(),
                            hipMemcpyHostToDevice, s));
  c->prof_begin(QBG_K_OTHER);
  launch_sf_tables(e->d_tau, e->nat, e->d_b, e->kmin, e->krange, e->d_sftab, s);
  launch_sfac(e->d_idxv, e->vb.size, e->d_sftab, e->krange, e->kmin, e->nsp, e->d_sp_first,
              e->d_vps, e->d_dvps, e->d_rhops, e->d_sfac, e->d_vion, e->d_dvion, e->d_rhopst, s);
  c->prof_end();
  int rc = ef_core_density(e);
  if (rc) return rc;
  ef_esr(e);
  QBG_HIP(c, hipMemcpyAsync(e->d_fesr, e->fesr.data(), sizeof(double) * e->fesr.size(),
                            hipMemcpyHostToDevice, s));
  QBG_HIP(c, hipStreamSynchronize(s));
  e->have_vhxc = false;
  return QBG_OK;
}

// the index tables between the complex basis wb and the real band basis ub
// (see qbg_ef::cplx): fold[2p + {0,1}] over ub, unfold[j] over wb
static bool kpoint_maps(const host::Basis& wb, const host::Basis& ub, std::vector<int32_t>& fold,
                        std::vector<int32_t>& unfold) {
  auto key = [](int h, int k, int l) -> int64_t {
    return ((int64_t)(h + 4096) << 26) | ((int64_t)(k + 4096) << 13) | (int64_t)(l + 4096);
  };
  std::unordered_map<int64_t, int> in_wb;
  in_wb.reserve((size_t)wb.size * 2);
  for (int j = 0; j < wb.size; j++) in_wb[key(wb.idx[3 * j], wb.idx[3 * j + 1], wb.idx[3 * j + 2])] = j;
  std::unordered_map<int64_t, int> in_ub;
  in_ub.reserve((size_t)ub.size * 2);
  fold.assign(2 * (size_t)ub.size, -1);
  for (int p = 0; p < ub.size; p++) {
    const int h = ub.idx[3 * p], k = ub.idx[3 * p + 1], l = ub.idx[3 * p + 2];
    in_ub[key(h, k, l)] = p;
    auto a = in_wb.find(key(h, k, l)), b = in_wb.find(key(-h, -k, -l));
    if (a != in_wb.end()) fold[2 * p] = a->second;
    if (b != in_wb.end()) fold[2 * p + 1] = b->second;
  }
  unfold.assign(wb.size, 0);
  for (int j = 0; j < wb.size; j++) {
    const int h = wb.idx[3 * j], k = wb.idx[3 * j + 1], l = wb.idx[3 * j + 2];
    auto a = in_ub.find(key(h, k, l));
    if (a != in_ub.end()) { unfold[j] = a->second; continue; }
    auto b = in_ub.find(key(-h, -k, -l));
    if (b == in_ub.end()) return false;
    unfold[j] = -1 - b->second;
  }
  return true;
}

// EnergyFunctional's constructor (src/EnergyFunctional.cpp:67-226) for one
// k-point: kp null or 0 gives the real (Gamma) basis
static int ef_create(qbg_ctx* c, const double* cell9, const double* refcell9, double ecut,
                     const double* kp, int nsp, const qbg_species_desc* spd, const int* na,
                     const double* tau, int nst, const double* occ, qbg_ef** out,
                     int nkp = 1, const double* kwts = nullptr) {
  {
    if (!c || !cell9 || !spd || !na || !tau || !occ || !out || nsp <= 0 || nst <= 0 || !(ecut > 0))
      return QBG_E_INVAL;
    if (nkp < 1 || (nkp > 1 && (!kp || !kwts))) return QBG_E_INVAL;
    *out = nullptr;
    QBG_HIP(c, hipSetDevice(c->device));
    qbg_ef* e = new (std::nothrow) qbg_ef;
    if (!e) return QBG_E_NOMEM;
    e->ctx = c;
    e->cell.set(cell9);
    e->omega = e->cell.volume;
    e->ecut = ecut;
    e->nsp = nsp;
    e->nst = nst;
    e->occ.assign(occ, occ + nst);
    e->sp_first.assign(1, 0);
    for (int is = 0; is < nsp; is++) {
      host::Species q;
      const qbg_species_desc& d = spd[is];
      q.zval = d.zval; q.nmesh = d.nmesh; q.deltar = d.deltar; q.rcps = d.rcps;
      bool ok;
      if (d.type == QBG_SPECIES_NCPP && d.nquad > 0) {
        // semilocal quadrature form: the semilocal potentials as given
        ok = d.vps && d.nmesh > 0 && d.lmax >= 0 && d.lmax <= 3;
        if (ok) {
          q.lmax = d.lmax;
          q.llocal = d.llocal;
          q.nquad = d.nquad;
          q.rquad = d.rquad;
          q.vsl.assign(d.vps, d.vps + (size_t)(d.lmax + 1) * d.nmesh);
        }
      } else if (d.type == QBG_SPECIES_NCPP) {
        ok = d.vps && d.phi && d.nmesh > 0 &&
             host::ncpp_to_kb(d.lmax, d.llocal, d.nmesh, d.deltar, d.vps, d.phi, q.vloc, q.proj,
                              q.dl, q.lmax);
      } else {
        ok = d.type == QBG_SPECIES_SLPP && d.vloc && d.proj && d.dl && d.lmax >= 0 &&
             d.lmax <= 3 && d.nmesh > 0;
        if (ok) {
          q.lmax = d.lmax;
          q.vloc.assign(d.vloc, d.vloc + q.nmesh);
          q.proj.assign(d.proj, d.proj + (size_t)(q.lmax + 1) * q.nmesh);
          q.dl.assign(d.dl, d.dl + q.lmax + 1);
        }
      }
      if (ok && d.nlcc) q.nlcc.assign(d.nlcc, d.nlcc + q.nmesh);
      if (!ok || !q.init()) {
        delete e;
        return c->fail(QBG_E_INVAL, "species: unsupported definition (nonlocal f channel of "
                                    "a norm-conserving potential, l > 3, or a quadrature "
                                    "species without rquad or with llocal out of range?)");
      }
      e->sp.push_back(q);
      e->na.push_back(na[is]);
      e->sp_first.push_back(e->sp_first.back() + na[is]);
    }
    e->nat = e->sp_first.back();
    e->tau.assign(tau, tau + 3 * e->nat);
    for (int is = 0; is < nsp; is++) e->core |= e->sp[is].has_nlcc();   // :135-153
    for (int is = 0; is < nsp; is++) e->eself += e->na[is] * e->sp[is].eself();
    // bases and grid (src/ChargeDensity.cpp:34-98)
    host::Cell def = e->cell;
    if (refcell9) def.set(refcell9);
    e->defcell = def;
    e->has_ref = refcell9 != nullptr;
    e->cplx = nkp > 1 || (kp && (kp[0] != 0.0 || kp[1] != 0.0 || kp[2] != 0.0));
    e->nkp = nkp;
    e->wb.build(def, e->cell, ecut, e->cplx ? kp : nullptr, nkp > 1);
    if (nkp > 1) {
      e->kset.resize(nkp);
      for (int ik = 0; ik < nkp; ik++) {
        e->kset[ik].weight = kwts[ik];
        if (ik > 0) e->kset[ik].wb.build(def, e->cell, ecut, kp + 3 * ik, true);
      }
    }
    e->ngw_max = e->wb.size;
    for (int ik = 1; ik < nkp; ik++) e->ngw_max = std::max(e->ngw_max, e->kset[ik].wb.size);
    e->vb.build(def, e->cell, 4.0 * ecut);
    for (int i = 0; i < 3; i++) {
      int n = e->vb.np[i] + 2;
      while (!host::factorizable(n)) n += 2;
      e->np[i] = n;
    }
    e->N = (long)e->np[0] * e->np[1] * e->np[2];
    // the basis the band transforms run on: wb, or at k != 0 the real basis
    // of every |G| <= max over wb of |G| (norms in the defining cell)
    const host::Basis* band = &e->wb;
    std::vector<int32_t> fold, unfold;
    std::vector<std::vector<int32_t>> kfold(nkp), kunfold(nkp);
    if (e->cplx) {
      double g2max = 0.0;
      for (int ik = 0; ik < nkp; ik++) {
        const host::Basis& b = ik == 0 ? e->wb : e->kset[ik].wb;
        for (int j = 0; j < b.size; j++) {
          double t2 = 0.0;
          for (int a = 0; a < 3; a++) {
            const double t = (b.idx[3 * j] * def.b[0][a] + b.idx[3 * j + 1] * def.b[1][a]) +
                             b.idx[3 * j + 2] * def.b[2][a];
            t2 += t * t;
          }
          g2max = std::max(g2max, t2);
        }
      }
      e->ub.build(def, e->cell, 0.5 * (g2max * (1.0 + 1e-9) + 1e-12));
      band = &e->ub;
      bool ok = kpoint_maps(e->wb, e->ub, fold, unfold);
      for (int ik = 1; ik < nkp && ok; ik++)
        ok = kpoint_maps(e->kset[ik].wb, e->ub, kfold[ik], kunfold[ik]);
      if (!ok) {
        delete e;
        return c->fail(QBG_E_INVAL, "k-point basis: a vector outside the band basis");
      }
    }
    if (!e->wm.build(*band, e->np[0], e->np[1], e->np[2]) ||
        !e->vm.build(e->vb, e->np[0], e->np[1], e->np[2])) {
      delete e;
      return c->fail(QBG_E_INVAL, "basis does not fit in the grid");
    }
    int rc;
    if ((rc = qbg_grid_create(c, e->np[0], e->np[1], e->np[2], band->size, e->wm.nvec,
                              e->wm.ip.data(), e->wm.im.data(), e->wm.zvec_to_val.data(), &e->gw)) ||
        (rc = qbg_grid_create(c, e->np[0], e->np[1], e->np[2], e->vb.size, e->vm.nvec,
                              e->vm.ip.data(), e->vm.im.data(), e->vm.zvec_to_val.data(), &e->gv))) {
      ef_free(e);
      delete e;
      return rc;
    }
    const int ngw = e->wb.size, ngv = e->vb.size, ngwm = e->ngw_max;
    // cell_moved tables (src/EnergyFunctional.cpp:1200-1245)
    std::vector<double> vps, dvps, rhops, twnl, dtwnl, rhoc, drhoc;
    ef_local_tables(e, vps, dvps, rhops, rhoc, drhoc);
    ef_twnl(e, twnl, dtwnl);
    // projector layout: species-major, ipr-major within a species, atom-minor
    // (src/NonLocalPotential.cpp:1341-1395)
    std::vector<int> twnl_off(nsp);
    std::vector<int32_t> m_atom, m_twnl, m_l, m_ipr, m_npr, m_stride, m_dbase, a_m0(e->nat),
        a_npr(e->nat), a_stride(e->nat);
    std::vector<double> m_wt, dmat, fac2(nsp);
    std::vector<int64_t> dt_off(nsp);
    std::vector<int32_t> sp_npr(nsp);
    for (int is = 0, tw = 0; is < nsp; is++) {
      const host::Species& q = e->sp[is];
      const int npr = q.nlm;
      twnl_off[is] = tw;
      dt_off[is] = (int64_t)6 * tw;
      tw += npr * ngw;
      std::vector<int> lp, hp, qp;
      std::vector<double> dp;
      q.projectors(lp, hp, qp, dp);
      sp_npr[is] = npr;
      fac2[is] = 2.0 * 0.25 * q.rcps * q.rcps;
      // D matrix: diagonal, D_l in (l, m) (src/Species.cpp:1450-1464) or the
      // quadrature weights (src/NonLocalPotential.cpp:196-229)
      const int dbase = (int)dmat.size();
      for (int i = 0; i < npr; i++)
        for (int j = 0; j < npr; j++) dmat.push_back(i == j ? dp[i] : 0.0);
      const int Ms = (int)m_atom.size();
      for (int ipr = 0; ipr < npr; ipr++)
        for (int ia = 0; ia < e->na[is]; ia++) {
          const int a = e->sp_first[is] + ia;
          m_atom.push_back(a);
          m_twnl.push_back(twnl_off[is] + ipr * ngw);
          m_l.push_back(lp[ipr]);
          m_ipr.push_back(ipr);
          m_npr.push_back(npr);
          m_stride.push_back(e->na[is]);
          m_dbase.push_back(dbase);
          m_wt.push_back(1.0);   // wsg = 1 for SLPP (src/Species.cpp:733)
        }
      for (int ia = 0; ia < e->na[is]; ia++) {
        const int a = e->sp_first[is] + ia;
        a_m0[a] = Ms + ia;
        a_npr[a] = npr;
        a_stride[a] = e->na[is];
      }
    }
    e->M = e->Mh = (int)m_atom.size();
    std::vector<int32_t> a_m0h;
    if (e->cplx) {
      // complex projections as two real ones: columns Mh + m are i times
      // columns m ((-i)^(l+3) = i (-i)^l), same twnl, atom and D block
      const int Mh = e->Mh;
      for (int m = 0; m < Mh; m++) {
        m_atom.push_back(m_atom[m]);
        m_twnl.push_back(m_twnl[m]);
        m_l.push_back((m_l[m] + 3) & 3);
        m_ipr.push_back(m_ipr[m]);
        m_npr.push_back(m_npr[m]);
        m_stride.push_back(m_stride[m]);
        m_dbase.push_back(m_dbase[m]);
        m_wt.push_back(m_wt[m]);
      }
      e->M = 2 * Mh;
      for (int a = 0; a < e->nat; a++) a_m0h.push_back(a_m0[a] + Mh);
    }
    // structure factor table extents (src/StructureFactor.cpp:27-39)
    int krange = 0;
    for (int i = 0; i < 3; i++) {
      e->kmin[i] = e->vb.idxmin[i];
      krange = std::max(krange, e->vb.idxmax[i] - e->vb.idxmin[i] + 1);
    }
    e->krange = krange;
    // atom groups of the local-force kernel: consecutive atoms of one species
    std::vector<int32_t> lfgrp;
    {
      const int ag = local_forces_group_size(krange);
      for (int is = 0; is < nsp; is++)
        for (int a = e->sp_first[is]; a < e->sp_first[is + 1]; a += ag) {
          lfgrp.push_back(a);
          lfgrp.push_back(std::min(ag, e->sp_first[is + 1] - a));
          lfgrp.push_back(is);
        }
      e->nlfgrp = (int)lfgrp.size() / 3;
    }
    std::vector<double> bm(9);
    for (int i = 0; i < 9; i++) bm[i] = e->cell.b[i / 3][i % 3];
    std::vector<int32_t> atom_sp(e->nat), spf(e->sp_first.begin(), e->sp_first.end());
    for (int is = 0; is < nsp; is++)
      for (int a = e->sp_first[is]; a < e->sp_first[is + 1]; a++) atom_sp[a] = is;
    std::vector<double> g2iv(e->vb.g2i), giv(e->vb.gi);
    const long N = e->N;
    const size_t nbmax = (size_t)std::max<long>((N + 255) / 256, (ngv + 255) / 256);
    e->partial_doubles = std::max<size_t>(nbmax * 18, (size_t)e->nat * 64 * 3 + 16);
    e->partial_doubles = std::max<size_t>(e->partial_doubles, ((size_t)e->M * nst + 255) / 256 + 16);
    e->partial_doubles = std::max<size_t>(e->partial_doubles,
                                          (size_t)((ngwm + 255) / 256) * ((nst + 31) / 32) * 8);
    e->partial_doubles = std::max<size_t>(e->partial_doubles, (size_t)nsp * 64 * 6 + 16);
    // split-K partials of the fnl product: up to 40 splits of M x nst
    e->work_doubles = std::max<size_t>((size_t)1 << 22, (size_t)40 * e->M * nst);
    // occupations times the weight of k-point 0 (of a set; 1 otherwise)
    std::vector<double> occ_w(e->occ);
    if (nkp > 1)
      for (double& x : occ_w) x *= e->kset[0].weight;
    if ((rc = dput(c, &e->d_g2w, e->wb.kpg2)) || (rc = dput(c, &e->d_gxw, e->wb.kpgx)) ||
        (rc = dput(c, &e->d_idxv, e->vb.idx)) || (rc = dput(c, &e->d_g2iv, g2iv)) ||
        (rc = dput(c, &e->d_giv, giv)) || (rc = dput(c, &e->d_gxv, e->vb.gx)) ||
        (rc = dput(c, &e->d_vps, vps)) || (rc = dput(c, &e->d_dvps, dvps)) ||
        (rc = dput(c, &e->d_rhops, rhops)) || (rc = dput(c, &e->d_twnl, twnl)) ||
        (rc = dput(c, &e->d_b, bm)) || (rc = dput(c, &e->d_sp_first, spf)) ||
        (rc = dput(c, &e->d_atom_sp, atom_sp)) || (rc = dput(c, &e->d_occ, occ_w)) ||
        (rc = dput(c, &e->d_m_atom, m_atom)) || (rc = dput(c, &e->d_m_twnl, m_twnl)) ||
        (rc = dput(c, &e->d_m_l, m_l)) || (rc = dput(c, &e->d_m_ipr, m_ipr)) ||
        (rc = dput(c, &e->d_m_npr, m_npr)) || (rc = dput(c, &e->d_m_stride, m_stride)) ||
        (rc = dput(c, &e->d_m_dbase, m_dbase)) || (rc = dput(c, &e->d_m_wt, m_wt)) ||
        (rc = dput(c, &e->d_dmat, dmat)) || (rc = dput(c, &e->d_a_m0, a_m0)) ||
        (rc = dput(c, &e->d_a_npr, a_npr)) || (rc = dput(c, &e->d_a_stride, a_stride)) ||
        (rc = dput(c, &e->d_dtwnl, dtwnl)) || (rc = dput(c, &e->d_dt_off, dt_off)) ||
        (rc = dput(c, &e->d_lfgrp, lfgrp)) || (rc = dalloc(c, &e->d_vtemp, (size_t)2 * nsp * ngv)) ||
        (rc = dput(c, &e->d_sp_npr, sp_npr)) || (rc = dput(c, &e->d_fac2, fac2)) ||
        (rc = dalloc(c, &e->d_tau, 3 * e->nat)) ||
        (rc = dalloc(c, &e->d_sftab, (size_t)e->nat * 3 * krange * 2)) ||
        (rc = dalloc(c, &e->d_sfac, (size_t)2 * nsp * ngv)) ||
        (rc = dalloc(c, &e->d_vion, (size_t)2 * ngv)) || (rc = dalloc(c, &e->d_dvion, (size_t)2 * ngv)) ||
        (rc = dalloc(c, &e->d_rhopst, (size_t)2 * ngv)) || (rc = dalloc(c, &e->d_rhor, N)) ||
        (rc = dalloc(c, &e->d_vxc, N)) || (rc = dalloc(c, &e->d_vr, N)) ||
        (rc = dalloc(c, &e->d_rhog, (size_t)2 * ngv)) || (rc = dalloc(c, &e->d_rhoelg, (size_t)2 * ngv)) ||
        (rc = dalloc(c, &e->d_rhogt, (size_t)2 * ngv)) || (rc = dalloc(c, &e->d_vlocal, (size_t)2 * ngv)) ||
        (rc = dalloc(c, &e->d_fesr, 3 * e->nat)) || (rc = dalloc(c, &e->d_fion, 3 * e->nat)) ||
        (rc = dalloc(c, &e->d_scal, S_COUNT)) || (rc = dalloc(c, &e->d_partial, e->partial_doubles)) ||
        (rc = dalloc(c, &e->d_anl, (size_t)2 * e->M * 2 * ngwm)) ||
        (rc = dalloc(c, &e->d_fnl, (size_t)2 * e->M * nst)) ||
        (rc = dalloc(c, &e->d_fout, (size_t)e->M * nst)) ||
        (rc = dalloc(c, &e->d_work, e->work_doubles)) ||
        (e->core && ((rc = dput(c, &e->d_rhoc, rhoc)) || (rc = dput(c, &e->d_drhoc, drhoc)) ||
                     (rc = dalloc(c, &e->d_rhocg, (size_t)2 * ngv)) ||
                     (rc = dalloc(c, &e->d_rhocr, N)) ||
                     (rc = dalloc(c, &e->d_vxcg, (size_t)2 * ngv)))) ||
        (e->cplx && ((rc = dput(c, &e->d_fold, fold)) || (rc = dput(c, &e->d_unfold, unfold)) ||
                     (rc = dput(c, &e->d_a_m0h, a_m0h)) ||
                     (rc = dput(c, &e->d_occh, occ_half(occ_w))) ||
                     (rc = dalloc(c, &e->d_cpair, (size_t)4 * e->ub.size * nst * nkp)) ||
                     (rc = dalloc(c, &e->d_hpair, (size_t)4 * e->ub.size * nst * nkp)))) ||
        (nkp > 1 && (rc = dalloc(c, &e->d_kacc, (size_t)16 * nkp)))) {
      ef_free(e);
      delete e;
      return rc;
    }
    e->o_kin = e->d_scal + S_KIN;
    e->o_enl = e->d_scal + S_ENL;
    e->o_nls = e->d_scal + S_NLS;
    // the further k-points of a set: their tables (k + G, twnl, the twnl
    // offsets of the projector columns, fold maps, weighted occupations)
    for (int ik = 1; ik < nkp; ik++) {
      qbg_ef::KPoint& k = e->kset[ik];
      const int ngk = k.wb.size;
      std::swap(e->wb, k.wb);
      std::vector<double> tk, dtk;
      ef_twnl(e, tk, dtk);
      std::swap(e->wb, k.wb);
      std::vector<int32_t> mt;
      std::vector<int64_t> dto(nsp);
      for (int is = 0, tw = 0; is < nsp; is++) {
        const int npr = e->sp[is].nlm;
        dto[is] = (int64_t)6 * tw;
        for (int ipr = 0; ipr < npr; ipr++)
          for (int ia = 0; ia < e->na[is]; ia++) mt.push_back(tw + ipr * ngk);
        tw += npr * ngk;
      }
      const size_t mh = mt.size();
      for (size_t m = 0; m < mh; m++) mt.push_back(mt[m]);   // the i anl half
      std::vector<double> ow(e->occ);
      for (double& x : ow) x *= k.weight;
      if ((rc = dput(c, &k.d_g2w, k.wb.kpg2)) || (rc = dput(c, &k.d_gxw, k.wb.kpgx)) ||
          (rc = dput(c, &k.d_twnl, tk)) || (rc = dput(c, &k.d_dtwnl, dtk)) ||
          (rc = dput(c, &k.d_m_twnl, mt)) || (rc = dput(c, &k.d_dt_off, dto)) ||
          (rc = dput(c, &k.d_fold, kfold[ik])) || (rc = dput(c, &k.d_unfold, kunfold[ik])) ||
          (rc = dput(c, &k.d_occ, ow)) || (rc = dput(c, &k.d_occh, occ_half(ow)))) {
        ef_free(e);
        delete e;
        return rc;
      }
    }
    // density factors and result staging
    {
      std::vector<double> dfac = density_factors(e);
      e->nres = S_COUNT + 4 + 3 * e->nat;
      std::vector<const double*> segp = {e->d_scal, e->d_rhoelg, e->d_vion, e->d_fion};
      std::vector<int> segn = {S_COUNT, 2, 2, 3 * e->nat};
      if ((rc = dput(c, &e->d_dfac, dfac)) || (rc = dput(c, &e->d_seg_ptr, segp)) ||
          (rc = dput(c, &e->d_seg_n, segn))) {
        ef_free(e);
        delete e;
        return rc;
      }
      if (hipHostMalloc((void**)&e->h_res, sizeof(double) * e->nres, hipHostMallocDefault) !=
          hipSuccess) {
        e->h_res = nullptr;
        ef_free(e);
        delete e;
        return c->fail(QBG_E_NOMEM, "pinned result buffer");
      }
    }
    if ((rc = ef_atoms_moved(e))) {
      ef_free(e);
      delete e;
      return rc;
    }
    *out = e;
    return QBG_OK;
  }
}

extern "C" {

int qbg_ef_create(qbg_ctx* c, const double* cell9, const double* refcell9, double ecut, int nsp,
                  const qbg_species_desc* spd, const int* na, const double* tau, int nst,
                  const double* occ, qbg_ef** out) {
  return qbg::guarded(c, [&]() -> int {
    return ef_create(c, cell9, refcell9, ecut, nullptr, nsp, spd, na, tau, nst, occ, out);
  });
}

int qbg_ef_create_k(qbg_ctx* c, const double* cell9, const double* refcell9, double ecut,
                    const double* kpoint3, int nsp, const qbg_species_desc* spd, const int* na,
                    const double* tau, int nst, const double* occ, qbg_ef** out) {
  return qbg::guarded(c, [&]() -> int {
    if (!kpoint3) return QBG_E_INVAL;
    return ef_create(c, cell9, refcell9, ecut, kpoint3, nsp, spd, na, tau, nst, occ, out);
  });
}

int qbg_ef_create_kset(qbg_ctx* c, const double* cell9, const double* refcell9, double ecut,
                       int nkp, const double* kpoints3n, const double* weights, int nsp,
                       const qbg_species_desc* spd, const int* na, const double* tau, int nst,
                       const double* occ, qbg_ef** out) {
  return qbg::guarded(c, [&]() -> int {
    if (!kpoints3n || !weights || nkp < 1) return QBG_E_INVAL;
    if (nkp == 1)
      return ef_create(c, cell9, refcell9, ecut, kpoints3n, nsp, spd, na, tau, nst, occ, out);
    return ef_create(c, cell9, refcell9, ecut, kpoints3n, nsp, spd, na, tau, nst, occ, out, nkp,
                     weights);
  });
}

int qbg_ef_destroy(qbg_ef* e) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e) return QBG_E_INVAL;
    QBG_HIP(e->ctx, hipSetDevice(e->ctx->device));
    (void)hipStreamSynchronize(e->ctx->stream);
    ef_free(e);
    delete e;
    return QBG_OK;
  });
}

// out: ngw, ngv, np0, np1, np2, nvec_w, nvec_v, nat, M, nst
int qbg_ef_sizes(const qbg_ef* e, int* out) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !out) return QBG_E_INVAL;
    const int v[10] = {e->wb.size, e->vb.size, e->np[0], e->np[1], e->np[2], e->wm.nvec,
                       e->vm.nvec, e->nat, e->M, e->nst};
    std::memcpy(out, v, sizeof(v));
    return QBG_OK;
  });
}

// host copies of a basis (which 0: wavefunctions, 1: density, 2: the
// wavefunctions' k + G): idx[3*ng], g2[ng], gx[3*ng] (each may be null), and
// the maps (real bases only)
int qbg_ef_basis(const qbg_ef* e, int which, int32_t* idx, double* g2, double* gx, int32_t* ip,
                 int32_t* im, int32_t* zvec_to_val) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || which < 0 || which > 2) return QBG_E_INVAL;
    const host::Basis& b = which == 1 ? e->vb : e->wb;
    const host::Maps& m = which == 1 ? e->vm : e->wm;
    if (which != 1 && e->cplx && (ip || im || zvec_to_val))
      return e->ctx->fail(QBG_E_INVAL, "basis: the complex (k != 0) basis has no band-grid maps");
    const bool kq = which == 2;
    if (idx) std::memcpy(idx, b.idx.data(), sizeof(int32_t) * b.idx.size());
    if (g2) std::memcpy(g2, (kq ? b.kpg2 : b.g2).data(), sizeof(double) * b.size);
    if (gx) std::memcpy(gx, (kq ? b.kpgx : b.gx).data(), sizeof(double) * b.gx.size());
    if (ip) std::memcpy(ip, m.ip.data(), sizeof(int32_t) * m.ip.size());
    if (im) std::memcpy(im, m.im.data(), sizeof(int32_t) * m.im.size());
    if (zvec_to_val) std::memcpy(zvec_to_val, m.zvec_to_val.data(), sizeof(int32_t) * m.nvec);
    return QBG_OK;
  });
}

// k-point ik of a k-point set (ik = 0 for a single k-point): its basis size
// and, where the pointers are given, idx[3 ngw], |k+G|^2, k+G[3 ngw]
int qbg_ef_kset_basis(const qbg_ef* e, int ik, int* ngw, int32_t* idx, double* kpg2,
                      double* kpgx) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || ik < 0 || ik >= e->nkp) return QBG_E_INVAL;
    const host::Basis& b = ik == 0 ? e->wb : e->kset[ik].wb;
    if (ngw) *ngw = b.size;
    if (idx) std::memcpy(idx, b.idx.data(), sizeof(int32_t) * b.idx.size());
    if (kpg2) std::memcpy(kpg2, b.kpg2.data(), sizeof(double) * b.size);
    if (kpgx) std::memcpy(kpgx, b.kpgx.data(), sizeof(double) * b.kpgx.size());
    return QBG_OK;
  });
}

int qbg_ef_set_positions(qbg_ef* e, const double* tau) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !tau) return QBG_E_INVAL;
    QBG_HIP(e->ctx, hipSetDevice(e->ctx->device));
    e->tau.assign(tau, tau + 3 * e->nat);
    return ef_atoms_moved(e);
  });
}

// Wavefunction::resize to a new cell over the same reference cell, then
// EnergyFunctional::cell_moved (src/EnergyFunctional.cpp:1200-1245):
// basis vectors of both bases (membership and grid unchanged), the local
// tables, update_twnl, and atoms_moved at the current Cartesian positions
int qbg_ef_cell_moved(qbg_ef* e, const double* cell9) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !cell9) return QBG_E_INVAL;
    qbg_ctx* c = e->ctx;
    if (c->capturing) return c->fail(QBG_E_STATE, "cell_moved during graph capture");
    host::Cell cell;
    cell.set(cell9);
    if (!(cell.volume > 0.0)) return c->fail(QBG_E_INVAL, "cell_moved: singular cell");
    // without a reference cell the basis follows the cell (defcell = cell,
    // src/Basis.cpp:314-325); a change of membership or grid needs a new
    // functional
    const host::Cell& def = e->has_ref ? e->defcell : cell;
    host::Basis wb, vb;
    // (the complex basis of a k-point, the Gamma point of a set included)
    wb.build(def, cell, e->ecut, e->cplx ? e->wb.kpoint : nullptr, e->nkp > 1);
    vb.build(def, cell, 4.0 * e->ecut);
    bool same = wb.idx == e->wb.idx && vb.idx == e->vb.idx;
    std::vector<host::Basis> kwb(e->nkp);
    for (int ik = 1; ik < e->nkp && same; ik++) {
      kwb[ik].build(def, cell, e->ecut, e->kset[ik].wb.kpoint, true);
      same = kwb[ik].idx == e->kset[ik].wb.idx;
    }
    for (int i = 0; i < 3 && same; i++) {
      int n = vb.np[i] + 2;
      while (!host::factorizable(n)) n += 2;
      same = n == e->np[i];
    }
    if (!same)
      return c->fail(QBG_E_INVAL, "cell_moved: the basis or grid would change (no reference "
                                  "cell covers the new cell); create the functional again");
    QBG_HIP(c, hipSetDevice(c->device));
    QBG_HIP(c, hipStreamSynchronize(c->stream));
    e->cell = cell;
    e->omega = cell.volume;
    e->wb = std::move(wb);
    e->vb = std::move(vb);
    std::vector<double> vps, dvps, rhops, twnl, dtwnl, rhoc, drhoc;
    ef_local_tables(e, vps, dvps, rhops, rhoc, drhoc);
    ef_twnl(e, twnl, dtwnl);
    std::vector<double> bm(9);
    for (int i = 0; i < 9; i++) bm[i] = e->cell.b[i / 3][i % 3];
    int rc;
    // the further k-points of a set: k + G and twnl of the new cell
    for (int ik = 1; ik < e->nkp; ik++) {
      qbg_ef::KPoint& k = e->kset[ik];
      k.wb = std::move(kwb[ik]);
      std::swap(e->wb, k.wb);
      std::vector<double> tk, dtk;
      ef_twnl(e, tk, dtk);
      std::swap(e->wb, k.wb);
      if ((rc = dset(c, k.d_g2w, k.wb.kpg2)) || (rc = dset(c, k.d_gxw, k.wb.kpgx)) ||
          (rc = dset(c, k.d_twnl, tk)) || (rc = dset(c, k.d_dtwnl, dtk)))
        return rc;
    }
    if ((rc = dset(c, e->d_g2w, e->wb.kpg2)) || (rc = dset(c, e->d_gxw, e->wb.kpgx)) ||
        (rc = dset(c, e->d_g2iv, e->vb.g2i)) || (rc = dset(c, e->d_giv, e->vb.gi)) ||
        (rc = dset(c, e->d_gxv, e->vb.gx)) || (rc = dset(c, e->d_vps, vps)) ||
        (rc = dset(c, e->d_dvps, dvps)) || (rc = dset(c, e->d_rhops, rhops)) ||
        (rc = dset(c, e->d_twnl, twnl)) || (rc = dset(c, e->d_dtwnl, dtwnl)) ||
        (rc = dset(c, e->d_b, bm)) || (rc = dset(c, e->d_dfac, density_factors(e))) ||
        (e->core && ((rc = dset(c, e->d_rhoc, rhoc)) || (rc = dset(c, e->d_drhoc, drhoc)))))
      return rc;
    e->have_rho = false;
    e->xres_c = nullptr;
    return ef_atoms_moved(e);
  });
}

// ChargeDensity::update_density (src/ChargeDensity.cpp:136-198):
// rhor = sum_n occ_n/omega |psi_n|^2; rhog = vft.forward(omega * rhor)
// The backward z and x transforms of the bands stay in the band grid's
// resident array when it fits (keep), for an H*psi of the same bands
// (ef_launch with reuse); the density itself is the same either way.
static int ef_update_density(qbg_ef* e, const double* c, int ldc, int prec, bool keep) {
  if (!e || !c || ldc < e->ngw_max) return QBG_E_INVAL;
  qbg_ctx* ctx = e->ctx;
  QBG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (prec < QBG_PREC_FP64 || prec > QBG_PREC_BF16X2) return ctx->fail(QBG_E_INVAL, "prec");
  fill_zero(e->d_rhor, e->N, s);
  if (e->nspin == 2) fill_zero(e->d_rhor2, e->N, s);
  const int fprec = prec == QBG_PREC_FP64 ? QBG_PREC_FP64 : QBG_PREC_FP32;
  e->xres_c = nullptr;
  int pairs = 0;
  for (int sp = 0; sp < e->nspin; sp++) {
    int b0, nb, p0;
    spin_bands(e, sp, b0, nb, p0);
    pairs = p0 + (nb + 1) / 2;
  }
  // (at k != 0 every band is one pair of the band grid)
  const bool res = keep && grid_ensure_xres(e->gw, e->cplx ? e->nst * e->nkp : pairs);
  int rc = QBG_OK;
  if (e->cplx) {
    // each complex band as the pair of its real and imaginary parts on the
    // real band basis, both with the band's occupation; one density per spin
    const int nu = e->ub.size, nst = e->nst;
    ctx->prof_begin(QBG_K_OTHER);
    for (int ik = 0; ik < e->nkp; ik++) {
      if (ik > 0) kswap(e, e->kset[ik]);
      launch_fold(c + 2L * ldc * nst * ik, ldc, nst, nu, e->d_fold, e->d_cpair + 4L * nu * nst * ik,
                  nu, s);
      if (ik > 0) kswap(e, e->kset[ik]);
    }
    ctx->prof_end();
    // a set: every k-point's pairs in one pass, each with its weighted
    // occupations (src/ChargeDensity.cpp:136-160); with two spins each
    // spin's bands of each k-point
    if (e->nkp > 1 && e->nspin == 1)
      rc = density_batches(e->gw, e->d_cpair, nu, 2 * nst * e->nkp, e->d_dfac, e->d_rhor, fprec,
                           res, 0);
    for (int ik = 0; ik < e->nkp && e->nkp > 1 && e->nspin == 2 && !rc; ik++)
      for (int sp = 0; sp < 2 && !rc; sp++) {
        int b0, nb, p0;
        spin_bands(e, sp, b0, nb, p0);
        const int q0 = ik * nst + b0;   // the first pair of these bands
        if (nb > 0)
          rc = density_batches(e->gw, e->d_cpair + 4L * nu * q0, nu, 2 * nb, e->d_dfac + 2 * q0,
                               sp == 0 ? e->d_rhor : e->d_rhor2, fprec, res, q0);
      }
    for (int sp = 0; sp < e->nspin && !rc && e->nkp == 1; sp++) {
      int b0, nb, p0;
      spin_bands(e, sp, b0, nb, p0);
      if (nb > 0)
        rc = density_batches(e->gw, e->d_cpair + 4L * nu * b0, nu, 2 * nb, e->d_dfac + 2 * b0,
                             sp == 0 ? e->d_rhor : e->d_rhor2, fprec, res, b0);
    }
  }
  // one density per spin (src/ChargeDensity.cpp:136-160)
  for (int sp = 0; sp < e->nspin && !rc && !e->cplx; sp++) {
    int b0, nb, p0;
    spin_bands(e, sp, b0, nb, p0);
    if (nb > 0)
      rc = density_batches(e->gw, c + 2L * ldc * b0, ldc, nb, e->d_dfac + 2 * p0,
                           sp == 0 ? e->d_rhor : e->d_rhor2, fprec, res, p0);
  }
  if (rc) return rc;
  if (res) {
    e->xres_c = c;
    e->xres_ldc = ldc;
    e->xres_prec = fprec;
  }
  // rhog of the total density (rhog[0] + rhog[1], src/EnergyFunctional.cpp:256-261)
  const double* rtot = e->d_rhor;
  if (e->nspin == 2) {
    add2(e->d_rhor, e->d_rhor2, e->d_rhot, e->N, s);
    rtot = e->d_rhot;
  }
  qbg_grid* g = e->gv;
  if ((rc = grid_ensure_ws(g, 1))) return rc;
  int r1, r2, r3;
  ctx->prof_begin(QBG_K_VGRID);
  r1 = ypass(g, QBG_PREC_FP64, Y_FWD_REAL, 1, rtot, nullptr, nullptr, nullptr, e->omega, s);
  ctx->prof_end();
  ctx->prof_begin(QBG_K_VGRID);
  r2 = r1 ? r1 : xfwd(g, QBG_PREC_FP64, 1, s);
  ctx->prof_end();
  ctx->prof_begin(QBG_K_VGRID);
  r3 = r2 ? r2 : zfwd(g, QBG_PREC_FP64, Z_SINGLE, OUT_STORE, e->d_rhog, 0, 0, 0, 1, s);
  ctx->prof_end();
  if (r3) return ctx->fail(r3, "update_density: density-grid transform refused");
  e->have_rho = true;
  e->have_vhxc = false;
  hipError_t er = hipGetLastError();
  if (er != hipSuccess) return ctx->fail(QBG_E_HIP, hipGetErrorString(er));
  return QBG_OK;
}

int qbg_ef_update_density(qbg_ef* e, const double* c, int ldc, int prec) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    return ef_update_density(e, c, ldc, prec, true);
  });
}

// EnergyFunctional::update_vhxc (src/EnergyFunctional.cpp:231-394), LDA
int qbg_ef_update_vhxc(qbg_ef* e, int compute_stress) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e) return QBG_E_INVAL;
    qbg_ctx* ctx = e->ctx;
    if (!e->have_rho) return ctx->fail(QBG_E_STATE, "update_vhxc before update_density");
    QBG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    ctx->prof_begin(QBG_K_OTHER);
    const double* rcore = e->core ? e->d_rhocr : nullptr;
    if (e->nspin == 2)
      launch_lsda(e->d_rhor, e->d_rhor2, rcore, e->N, e->d_vxc, e->d_vxc2, e->d_partial,
                  e->d_scal + S_LDA, s);
    else
      launch_lda(e->d_rhor, rcore, e->N, e->d_vxc, e->d_partial, e->d_scal + S_LDA, s);
    launch_vhxc_g(e->d_rhog, e->vb.size, 1.0 / e->omega, e->d_vion, e->d_dvion, e->d_rhopst,
                  e->d_g2iv, e->d_giv, e->d_gxv, e->d_rhoelg, e->d_rhogt, e->d_vlocal,
                  compute_stress, e->d_partial, e->d_scal + S_VHXC, s);
    ctx->prof_end();
    qbg_grid* g = e->gv;
    int rc = grid_ensure_ws(g, 1);
    if (rc) return rc;
    int r1, r2, r3;
    ctx->prof_begin(QBG_K_VGRID);
    r1 = zbwd(g, QBG_PREC_FP64, Z_SINGLE, e->d_vlocal, 0, 0, 0, 1, s);
    ctx->prof_end();
    ctx->prof_begin(QBG_K_VGRID);
    r2 = r1 ? r1 : xbwd(g, QBG_PREC_FP64, 1, s);
    ctx->prof_end();
    ctx->prof_begin(QBG_K_VGRID);
    // v_r[ispin] = vxc_r[ispin] + Re(vlocal) (src/EnergyFunctional.cpp:372-389)
    if (e->nspin == 2) {
      r3 = r2 ? r2 : ypass(g, QBG_PREC_FP64, Y_BWD_REAL_ADD, 1, nullptr, nullptr, nullptr,
                           e->d_vloc, 1.0, s);
      if (!r3) {
        add2(e->d_vxc, e->d_vloc, e->d_vr, e->N, s);
        add2(e->d_vxc2, e->d_vloc, e->d_vr2, e->N, s);
      }
    } else {
      r3 = r2 ? r2 : ypass(g, QBG_PREC_FP64, Y_BWD_REAL_ADD, 1, e->d_vxc, nullptr, nullptr,
                           e->d_vr, 1.0, s);
    }
    ctx->prof_end();
    if (r3) return ctx->fail(r3, "update_vhxc: density-grid transform refused");
    if (e->core) {
      // vxc_g = forward(vxc_r) (the spin average for nspin 2, by linearity
      // the transform of the averaged grid) and the core term of dxc
      // (src/EnergyFunctional.cpp:275-306)
      const double* vx = e->d_vxc;
      double scale = 1.0;
      if (e->nspin == 2) {
        add2(e->d_vxc, e->d_vxc2, e->d_vloc, e->N, s);   // d_vloc is free again here
        vx = e->d_vloc;
        scale = 0.5;
      }
      ctx->prof_begin(QBG_K_VGRID);
      int r = ypass(g, QBG_PREC_FP64, Y_FWD_REAL, 1, vx, nullptr, nullptr, nullptr, scale, s);
      if (!r) r = xfwd(g, QBG_PREC_FP64, 1, s);
      if (!r) r = zfwd(g, QBG_PREC_FP64, Z_SINGLE, OUT_STORE, e->d_vxcg, 0, 0, 0, 1, s);
      ctx->prof_end();
      if (r) return ctx->fail(r, "update_vhxc: density-grid transform refused");
      ctx->prof_begin(QBG_K_OTHER);
      launch_core_dxc(e->vb.size, e->d_vxcg, e->d_rhocg, e->d_partial, e->d_scal + S_CORE, s);
      ctx->prof_end();
    }
    e->have_vhxc = true;
    e->vhxc_stress = compute_stress != 0;
    hipError_t er = hipGetLastError();
    if (er != hipSuccess) return ctx->fail(QBG_E_HIP, hipGetErrorString(er));
    return QBG_OK;
  });
}

// EnergyFunctional::energy's nonlocal part (src/NonLocalPotential.cpp:1258-2069)
// with the projector GEMMs on FP64 or FP32 MFMA (gemm.hip)
// over the bands [b0, b0 + nst) of the functional (c and hpsi point at band
// b0; their occupations are occ[b0 + n])
static int nl_mfma(qbg_ef* e, const double* c, int ldc, int compute_hpsi, double* hpsi, int ldh,
                   int compute_forces, int compute_stress, int prec, int b0, int nst) {
  qbg_ctx* ctx = e->ctx;
  hipStream_t s = ctx->stream;
  const int ngw = e->wb.size, M = e->M;
  const double* occ = e->d_occ + b0;
  const long K = 2L * ngw;
  ctx->prof_begin(QBG_K_NL);
  const bool f32 = prec == QBG_PREC_FP32;
  float* anl32 = reinterpret_cast<float*>(e->d_anl);
  launch_anl(ngw, e->nat, K, e->d_gxw, e->d_tau, e->d_a_m0, e->d_a_npr, e->d_a_stride,
             e->d_m_twnl, e->d_m_l, e->d_twnl, e->d_anl, f32 ? 1 : 0, s);
  // k != 0: the second half of the columns, i anl (the imaginary part of
  // the complex projection anl^H c comes out of the same real TN product)
  if (e->cplx)
    launch_anl(ngw, e->nat, K, e->d_gxw, e->d_tau, e->d_a_m0h, e->d_a_npr, e->d_a_stride,
               e->d_m_twnl, e->d_m_l, e->d_twnl, e->d_anl, f32 ? 1 : 0, s);
  ctx->prof_end();
  // Gamma: fnl = 2 anl^T c - anl0 c0 (half sphere); k != 0: the full sphere
  // with neither the factor nor the G = 0 term (src/NonLocalPotential.cpp:1399-1445)
  const double alpha = e->cplx ? 1.0 : 2.0, gamma = e->cplx ? 0.0 : -1.0;
  ctx->prof_begin(QBG_K_GEMM);
  const int grc = f32 ? gemm_tn_splitk(anl32, K, c, 2L * ldc, M, nst, K, alpha, gamma, e->d_fnl,
                                       M, e->d_work, e->work_doubles, prec, s)
                      : gemm_tn_splitk(e->d_anl, K, c, 2L * ldc, M, nst, K, alpha, gamma,
                                       e->d_fnl, M, e->d_work, e->work_doubles, prec, s);
  ctx->prof_end();
  if (grc) return ctx->fail(QBG_E_INVAL, "gemm workspace too small");
  ctx->prof_begin(QBG_K_NL);
  launch_dmat_enl(e->d_fnl, M, nst, occ, e->d_m_ipr, e->d_m_npr, e->d_m_stride,
                  e->d_m_dbase, e->d_dmat, e->d_m_wt, 1.0 / e->omega, e->d_fout, e->d_partial,
                  e->o_enl, s);
  ctx->prof_end();
  if (compute_hpsi) {
    ctx->prof_begin(QBG_K_GEMM);
    if (f32) gemm_nn(anl32, K, e->d_fout, M, K, nst, M, hpsi, 2L * ldh, 1, prec, s);
    else gemm_nn(e->d_anl, K, e->d_fout, M, K, nst, M, hpsi, 2L * ldh, 1, prec, s);
    ctx->prof_end();
  }
  if (compute_forces || compute_stress) {
    // Y = c W with W[n,m] = occ_n fout[m,n] (stored in the second fnl block);
    // forces f_a,j = -4 sum_{m in a} sum_K (-i G_j anl)[K,m] Y[K,m] and the
    // nonlocal stress (src/NonLocalPotential.cpp:1786-2066) both reduce Y
    double* W = e->d_fnl + (size_t)M * nst;
    void* Y = f32 ? (void*)(anl32 + (size_t)M * K) : (void*)(e->d_anl + (size_t)M * K);
    ctx->prof_begin(QBG_K_NL);
    if (compute_forces && !e->fion_keep)
      fill_zero(e->d_fion, 3 * (size_t)e->nat, s);
    // the force kernel's factor 4 is 2 (|F|^2) x 2 (G and -G); the full
    // sphere of k != 0 takes half the occupation
    launch_force_w(e->d_fout, e->cplx ? e->d_occh + b0 : occ, M, nst, W, s);
    ctx->prof_end();
    ctx->prof_begin(QBG_K_GEMM);
    if (f32) gemm_nn(c, 2L * ldc, W, nst, K, M, nst, (float*)Y, K, 0, prec, s);
    else gemm_nn(c, 2L * ldc, W, nst, K, M, nst, (double*)Y, K, 0, prec, s);
    ctx->prof_end();
    ctx->prof_begin(QBG_K_NL);
    if (compute_forces) {
      // k != 0: the i anl half again, at its own column offset
      for (int half = 0; half < (e->cplx ? 2 : 1); half++) {
        const int32_t* am0 = half ? e->d_a_m0h : e->d_a_m0;
        if (f32)
          launch_nl_forces_y(anl32, (const float*)Y, K, ngw, e->d_gxw, am0, e->d_a_npr,
                             e->d_a_stride, e->nat, e->d_partial, 64, e->d_fion, s);
        else
          launch_nl_forces_y(e->d_anl, (const double*)Y, K, ngw, e->d_gxw, am0, e->d_a_npr,
                             e->d_a_stride, e->nat, e->d_partial, 64, e->d_fion, s);
      }
    }
    if (compute_stress) {
      if (f32)
        launch_nl_stress_y((const float*)Y, K, ngw, e->d_gxw, e->d_tau, e->nsp, e->d_sp_first,
                           e->d_sp_npr, e->d_dt_off, e->d_dtwnl, e->d_a_m0, e->d_a_stride,
                           e->d_m_l, e->d_partial, 64, e->o_nls, s,
                           e->cplx ? 2 : 1, e->Mh);
      else
        launch_nl_stress_y((const double*)Y, K, ngw, e->d_gxw, e->d_tau, e->nsp, e->d_sp_first,
                           e->d_sp_npr, e->d_dt_off, e->d_dtwnl, e->d_a_m0, e->d_a_stride,
                           e->d_m_l, e->d_partial, 64, e->o_nls, s,
                           e->cplx ? 2 : 1, e->Mh);
    }
    ctx->prof_end();
  }
  return QBG_OK;
}

// allocate the BF16-split workspace once; padding is zeroed here and never
// written afterwards
static int bf16_ensure(qbg_ef* e) {
  if (e->b_anl) return QBG_OK;
  qbg_ctx* c = e->ctx;
  if (c->capturing)
    return c->fail(QBG_E_STATE, "BF16 workspace must exist before graph capture "
                                "(run the call once uncaptured)");
  auto up = [](long x, long m) { return (x + m - 1) / m * m; };
  // 256-edge tiles of the split GEMMs when the projector and band counts
  // padded to 128 are multiples of 256 already (no extra padded work);
  // otherwise 128
  const long edge = (up(e->M, 128) % 256 == 0 && up(e->nst, 128) % 256 == 0) ? 256 : 128;
  e->Kp = up(2L * e->ngw_max, edge);   // (the largest basis of a k-point set)
  e->Mp = (int)up(e->M, edge);
  e->Np = (int)up(e->nst, edge);
  const int ntile = (e->Mp / 128) * (e->Np / 128);
  e->b_part_floats = (size_t)(1024 + ntile) * 128 * 128;
  const size_t n_anl = (size_t)e->Mp * e->Kp, n_c = (size_t)e->Np * e->Kp;
  const size_t n_f = (size_t)e->Mp * e->Np;
  float** bufs[6] = {&e->b_anl, &e->b_c32, &e->b_y32, &e->b_fout, &e->b_w32, &e->b_yh};
  const size_t sz[6] = {n_anl, n_c, n_anl, n_f, n_f, n_c};
  for (int i = 0; i < 6; i++) {
    if (hipMalloc((void**)bufs[i], sizeof(float) * sz[i]) != hipSuccess) {
      *bufs[i] = nullptr;
      return c->fail(QBG_E_NOMEM, "bf16 workspace");
    }
    QBG_HIP(c, hipMemsetAsync(*bufs[i], 0, sizeof(float) * sz[i], c->stream));
  }
  if (hipMalloc((void**)&e->b_part, sizeof(float) * e->b_part_floats) != hipSuccess) {
    e->b_part = nullptr;
    return c->fail(QBG_E_NOMEM, "bf16 workspace");
  }
  if (edge == 256) {
    int rc;
    if ((rc = dalloc(c, &e->b_fpart, (size_t)e->M * 3 * (e->Kp / 256))) ||
        (rc = dalloc(c, &e->b_fred, (size_t)e->M * 3)))
      return rc;
  }
  return QBG_OK;
}

// GEMM launch tuner: the TN projection (anl^T c) of the 2-part split timed
// with the 256-edge kernel at its default reduction split, at twice and half
// that split, and with the 128-edge kernel, on the functional's own operands
// (the output goes to d_fnl, which the real product overwrites next); the
// fastest plan by more than 3 % over the default is kept and cached per
// device and padded shape for the process.  The H*psi and force products
// keep their rule (256-edge tiles where the padding allows): their outputs
// are the caller's H*psi and the fused force sums.
namespace {
struct BfKey {
  int dev, Mp, Np;
  long Kp;
  bool operator<(const BfKey& o) const {
    return std::tie(dev, Mp, Np, Kp) < std::tie(o.dev, o.Mp, o.Np, o.Kp);
  }
};
std::map<BfKey, qbg::BfTnPlan> g_bf_cache;
std::mutex g_bf_mu;
}  // namespace

static int bf16_tune(qbg_ef* e, const double* c, int ldc, int nst) {
  qbg_ctx* ctx = e->ctx;
  if (e->bf_tuned || ctx->capturing) return QBG_OK;
  const BfKey key{ctx->device, e->Mp, e->Np, e->Kp};
  {
    std::lock_guard<std::mutex> lk(g_bf_mu);
    auto it = g_bf_cache.find(key);
    if (it != g_bf_cache.end()) {
      e->bf_tn = it->second;
      e->bf_tuned = true;
      return QBG_OK;
    }
  }
  hipStream_t s = ctx->stream;
  const int S0 = gemm_bf_tn_default_split(e->Mp, e->Np, e->Kp, 2, 256);
  BfTnPlan cand[4];
  cand[0] = BfTnPlan{256, S0};
  cand[1] = BfTnPlan{256, 2 * S0};
  cand[2] = BfTnPlan{256, std::max(1, S0 / 2)};
  cand[3] = BfTnPlan{128, 0};
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    return ctx->fail(QBG_E_HIP, "gemm tuner events");
  }
  float best_ms = 0.f;
  int best = 0, rc = QBG_OK;
  for (int i = 0; i < 4 && !rc; i++) {
    auto run = [&]() {
      return gemm_bf_tn(e->b_anl, e->Kp, e->b_c32, e->Kp, e->Mp, e->Np, e->Kp, e->M, nst, c,
                        2L * ldc, 2.0, -1.0, e->d_fnl, e->M, e->b_part, e->b_part_floats, 2, s,
                        &cand[i]);
    };
    if (run()) continue;                                   // shape not allowed
    (void)hipEventRecord(a, s);
    for (int r = 0; r < 3; r++) (void)run();
    (void)hipEventRecord(b, s);
    float ms = 0.f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
      rc = ctx->fail(QBG_E_HIP, "gemm tuner timing");
      break;
    }
    if (i == 0) best_ms = ms;
    else if (ms < 0.97f * best_ms) { best_ms = ms; best = i; }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  e->bf_tn = cand[best];
  e->bf_tuned = true;
  std::lock_guard<std::mutex> lk(g_bf_mu);
  g_bf_cache.emplace(key, e->bf_tn);
  return QBG_OK;
}

// the nonlocal part with the projector GEMMs as BF16-split products
// (gemm_bf.hip) on FP32 copies of anl, c, fout and W
// (bands [b0, b0 + nst) as nl_mfma; band columns [nst, e->nst) of the FP32
// copies are zeroed, so the padded band sums see zeros there)
// yh (or null): write the nonlocal part of H*psi there (FP32, ld Kp)
// instead of adding it to hpsi
static int nl_bf16(qbg_ef* e, const double* c, int ldc, int compute_hpsi, double* hpsi, int ldh,
                   int compute_forces, int compute_stress, int parts, bool c32_ready, int b0,
                   int nst, float* yh = nullptr) {
  qbg_ctx* ctx = e->ctx;
  hipStream_t s = ctx->stream;
  int rc = bf16_ensure(e);
  if (rc) return rc;
  const int ngw = e->wb.size, M = e->M, Mp = e->Mp, Np = e->Np;
  const double* occ = e->d_occ + b0;
  const long K = 2L * ngw, Kp = e->Kp;
  ctx->prof_begin(QBG_K_NL);
  launch_anl(ngw, e->nat, Kp, e->d_gxw, e->d_tau, e->d_a_m0, e->d_a_npr, e->d_a_stride,
             e->d_m_twnl, e->d_m_l, e->d_twnl, e->b_anl, 1, s);
  if (e->cplx)   // k != 0: the i anl columns (see nl_mfma)
    launch_anl(ngw, e->nat, Kp, e->d_gxw, e->d_tau, e->d_a_m0h, e->d_a_npr, e->d_a_stride,
               e->d_m_twnl, e->d_m_l, e->d_twnl, e->b_anl, 1, s);
  if (!c32_ready) convert_f32(c, 2L * ldc, K, nst, e->b_c32, Kp, s, e->nst);
  ctx->prof_end();
  if (parts == 2 && (rc = bf16_tune(e, c, ldc, nst))) return rc;
  ctx->prof_begin(QBG_K_GEMM);
  rc = gemm_bf_tn(e->b_anl, Kp, e->b_c32, Kp, Mp, Np, Kp, M, nst, c, 2L * ldc,
                  e->cplx ? 1.0 : 2.0, e->cplx ? 0.0 : -1.0, e->d_fnl, M, e->b_part,
                  e->b_part_floats, parts, s, parts == 2 ? &e->bf_tn : nullptr);
  ctx->prof_end();
  if (rc) return ctx->fail(QBG_E_INVAL, "bf16 gemm shape");
  ctx->prof_begin(QBG_K_NL);
  launch_dmat_enl(e->d_fnl, M, nst, occ, e->d_m_ipr, e->d_m_npr, e->d_m_stride,
                  e->d_m_dbase, e->d_dmat, e->d_m_wt, 1.0 / e->omega, e->d_fout, e->d_partial,
                  e->o_enl, s);
  const bool needy = compute_forces || compute_stress;
  // W = occ fout for the force product (half the occupation on the full
  // sphere of k != 0, see nl_mfma)
  fout_f32(e->d_fout, e->cplx ? e->d_occh + b0 : occ, M, nst, e->b_fout, Mp,
           needy ? e->b_w32 : nullptr, Np, s, e->nst);
  ctx->prof_end();
  if (compute_hpsi) {
    ctx->prof_begin(QBG_K_GEMM);
    rc = yh ? gemm_bf_nn(e->b_anl, Kp, e->b_fout, Mp, Kp, Np, Mp, K, nst, yh, Kp, 0, parts, s)
            : gemm_bf_nn(e->b_anl, Kp, e->b_fout, Mp, Kp, Np, Mp, K, nst, hpsi, 2L * ldh, 1,
                         parts, s);
    ctx->prof_end();
    if (rc) return ctx->fail(QBG_E_INVAL, "bf16 gemm shape");
  }
  // forces alone on 256-tile shapes: the force reduction runs in the product's
  // epilogue (Y is not stored)
  if (compute_forces && !compute_stress && parts == 2 && e->b_fpart) {
    ctx->prof_begin(QBG_K_GEMM);
    rc = gemm_bf_nn_forces(e->b_c32, Kp, e->b_w32, Np, Kp, Mp, Np, M, e->b_anl, Kp, e->d_gxw,
                           ngw, e->b_fpart, s);
    ctx->prof_end();
    if (rc == 0) {
      ctx->prof_begin(QBG_K_NL);
      if (!e->fion_keep) fill_zero(e->d_fion, 3 * (size_t)e->nat, s);
      for (int half = 0; half < (e->cplx ? 2 : 1); half++)
        launch_nl_forces_tiles(e->b_fpart, M, (int)(Kp / 256), e->b_fred,
                               half ? e->d_a_m0h : e->d_a_m0, e->d_a_npr, e->d_a_stride, e->nat,
                               e->d_fion, s);
      ctx->prof_end();
      return QBG_OK;
    }
  }
  if (needy) {
    ctx->prof_begin(QBG_K_GEMM);
    rc = gemm_bf_nn(e->b_c32, Kp, e->b_w32, Np, Kp, Mp, Np, Kp, Mp, e->b_y32, Kp, 0, parts, s);
    ctx->prof_end();
    if (rc) return ctx->fail(QBG_E_INVAL, "bf16 gemm shape");
    ctx->prof_begin(QBG_K_NL);
    if (compute_forces) {
      if (!e->fion_keep) fill_zero(e->d_fion, 3 * (size_t)e->nat, s);
      for (int half = 0; half < (e->cplx ? 2 : 1); half++)
        launch_nl_forces_y(e->b_anl, (const float*)e->b_y32, Kp, ngw, e->d_gxw,
                           half ? e->d_a_m0h : e->d_a_m0, e->d_a_npr, e->d_a_stride, e->nat,
                           e->d_partial, 64, e->d_fion, s);
    }
    if (compute_stress)
      launch_nl_stress_y((const float*)e->b_y32, Kp, ngw, e->d_gxw, e->d_tau, e->nsp,
                         e->d_sp_first, e->d_sp_npr, e->d_dt_off, e->d_dtwnl, e->d_a_m0,
                         e->d_a_stride, e->d_m_l, e->d_partial, 64, e->o_nls, s,
                         e->cplx ? 2 : 1, e->Mh);
    ctx->prof_end();
  }
  return QBG_OK;
}

// EnergyFunctional::energy (src/EnergyFunctional.cpp:397-1044): nspin 1,
// Gamma, LDA.  energies (host, 14): ekin eps enl ehart ehart_e ehart_ep
// ehart_p exc dxc esr eself ecoul etotal total_charge; fion (host, 3*nat)
// species-major; sigma (host, 7x6: total, ekin, eps, enl, ehart, exc, esr).
// reuse: the caller guarantees that c holds what the last update_density
// read; H*psi then starts from the transforms update_density kept
static int ef_launch(qbg_ef* e, const double* c, int ldc, int compute_hpsi, double* hpsi,
                     int ldh, int compute_forces, int compute_stress, int prec,
                     bool reuse = false) {
  if (!e || !c || ldc < e->ngw_max) return QBG_E_INVAL;
  qbg_ctx* ctx = e->ctx;
  if (!e->have_vhxc) return ctx->fail(QBG_E_STATE, "energy before update_vhxc");
  if (compute_hpsi && (!hpsi || ldh < e->ngw_max)) return QBG_E_INVAL;
  if (compute_stress && !e->vhxc_stress)
    return ctx->fail(QBG_E_STATE, "stress requested but update_vhxc ran without stress");
  if (prec < QBG_PREC_FP64 || prec > QBG_PREC_BF16X2) return ctx->fail(QBG_E_INVAL, "energy: prec");
  const int fprec = prec == QBG_PREC_FP64 ? QBG_PREC_FP64 : QBG_PREC_FP32;   // band FFTs
  QBG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int ngw = e->wb.size, nst = e->nst, M = e->M;
  // kinetic energy and hpsi = 0.5 kpg2 c (this store replaces the clear of dwf);
  // in the split modes the same pass writes the FP32 copy of c for the GEMMs
  const bool split = prec == QBG_PREC_BF16X3 || prec == QBG_PREC_BF16X2;
  if (split) {
    const int rc = bf16_ensure(e);
    if (rc) return rc;
  }
  // in the split modes the forward z pass completes H*psi: it stores
  // 0.5 kpg2 c + the nonlocal part (FP32, from the NN product) + the local
  // part, so neither the kinetic pass nor the NN product touches H*psi
  const bool fin = split && compute_hpsi && hpsi != c && !e->cplx && zfwd_fin_ok(e->gw);
  // the kinetic and nonlocal parts of each k-point (one pass for one
  // k-point; a set sums each k-point's weighted parts, src/EnergyFunctional.cpp:446-543, 703-716)
  const int nk = e->nkp;
  if (nk > 1) fill_zero(e->d_kacc, (size_t)16 * nk, s);
  int nrc = QBG_OK;
  for (int ik = 0; ik < nk && !nrc; ik++) {
    if (ik > 0) kswap(e, e->kset[ik]);
    if (nk > 1) {
      e->o_kin = e->d_kacc + 16 * ik;
      e->o_enl = e->d_kacc + 16 * ik + 8;
      e->o_nls = e->d_kacc + 16 * ik + 9;
      e->fion_keep = ik > 0;
    }
    const int ngk = e->wb.size;
    const double* ck = c + 2L * ldc * nst * ik;
    double* hk = compute_hpsi ? hpsi + 2L * ldh * nst * ik : hpsi;
    ctx->prof_begin(QBG_K_OTHER);
    // a smaller basis than the set's largest: its FP32 copy's padding rows
    // (written by a larger k-point) are cleared for the split products
    if (split && nk > 1 && 2L * ngk < e->Kp)
      launch_zero_rows_f32(e->b_c32, e->Kp, 2L * ngk, e->Kp, nst, s);
    // (the half sphere's G and -G times 1/2 is occ |c|^2 |G|^2; the full
    // sphere of k != 0 takes occ/2 |c|^2 |k+G|^2)
    launch_kinetic(ck, ldc, nst, e->cplx ? e->d_occh : e->d_occ, ngk, e->d_g2w, e->d_gxw, hk,
                   ldh, compute_hpsi && !fin, e->d_partial, e->o_kin, s,
                   split ? e->b_c32 : nullptr, split ? e->Kp : 0);
    ctx->prof_end();
    // nonlocal: anl, fnl = 2 anl^T c - anl0 c0^T (FP64 partial sums), the D
    // contraction and enl in FP64; in FP32 mode anl and Y are FP32 arrays and
    // the GEMMs run on FP32 MFMA; in the BF16 modes on the BF16-split path
    nrc = (prec == QBG_PREC_BF16X3 || prec == QBG_PREC_BF16X2)
              ? nl_bf16(e, ck, ldc, compute_hpsi, hk, ldh, compute_forces, compute_stress,
                        prec == QBG_PREC_BF16X3 ? 3 : 2, true, 0, nst,
                        fin ? e->b_yh : nullptr)
              : nl_mfma(e, ck, ldc, compute_hpsi, hk, ldh, compute_forces, compute_stress,
                        prec, 0, nst);
    if (ik > 0) kswap(e, e->kset[ik]);
  }
  if (nk > 1) {
    e->o_kin = e->d_scal + S_KIN;
    e->o_enl = e->d_scal + S_ENL;
    e->o_nls = e->d_scal + S_NLS;
    e->fion_keep = false;
    launch_kacc_sum(e->d_kacc, nk, e->d_scal + S_KIN, e->d_scal + S_ENL, e->d_scal + S_NLS, s);
  }
  if (nrc) return nrc;
  if (compute_stress) {
    ctx->prof_begin(QBG_K_OTHER);
    launch_hart_ps_stress(e->d_rhogt, e->d_sfac, e->d_rhops, e->d_g2iv, e->d_gxv, e->vb.size,
                          e->nsp, e->d_fac2, e->core ? e->d_vxcg : nullptr, e->d_drhoc,
                          e->d_partial, e->d_scal + S_HPS, s);
    ctx->prof_end();
  }
  // local potential: hpsi += FFT(v_r * FFT^-1 c); at k != 0 on the folded
  // pairs, unfolded into hpsi
  if (compute_hpsi && e->cplx) {
    const int nu = e->ub.size;
    // the x-pass output of the folded bands that update_density kept, or a
    // fold and full transforms here
    const bool kept = reuse && e->xres_c == c && e->xres_ldc == ldc && e->xres_prec == fprec;
    ctx->prof_begin(QBG_K_OTHER);
    for (int ik = 0; ik < e->nkp && !kept; ik++) {
      if (ik > 0) kswap(e, e->kset[ik]);
      launch_fold(c + 2L * ldc * nst * ik, ldc, nst, nu, e->d_fold, e->d_cpair + 4L * nu * nst * ik,
                  nu, s);
      if (ik > 0) kswap(e, e->kset[ik]);
    }
    fill_zero(e->d_hpair, (size_t)4 * nu * nst * e->nkp, s);
    ctx->prof_end();
    if (e->nkp > 1) {
      // every k-point's pairs through one batched pass; with two spins each
      // spin's bands of each k-point with that spin's potential
      for (int ik = 0; ik < (e->nspin == 1 ? 1 : e->nkp); ik++)
        for (int sp = 0; sp < e->nspin; sp++) {
          int b0, nb, p0;
          spin_bands(e, sp, b0, nb, p0);
          const int q0 = e->nspin == 1 ? 0 : ik * nst + b0;
          const int nbb = e->nspin == 1 ? nst * e->nkp : nb;
          if (nbb == 0) continue;
          const double* v = sp == 0 ? e->d_vr : e->d_vr2;
          const int rc = kept ? hpsi_resident(e->gw, v, 2 * nbb, e->d_hpair + 4L * nu * q0, nu,
                                              fprec, q0, nullptr)
                              : rs_mul_add_bands(e->gw, v, e->d_cpair + 4L * nu * q0, nu, 2 * nbb,
                                                 e->d_hpair + 4L * nu * q0, nu, fprec, nullptr);
          if (rc) return rc;
        }
      ctx->prof_begin(QBG_K_OTHER);
      for (int ik = 0; ik < e->nkp; ik++) {
        if (ik > 0) kswap(e, e->kset[ik]);
        launch_unfold_add(e->d_hpair + 4L * nu * nst * ik, nu, nst, e->wb.size, e->d_unfold,
                          hpsi + 2L * ldh * nst * ik, ldh, s);
        if (ik > 0) kswap(e, e->kset[ik]);
      }
      ctx->prof_end();
      if (kept) e->xres_c = nullptr;
    } else {
      // each band with its spin's potential (src/EnergyFunctional.cpp:831)
      for (int sp = 0; sp < e->nspin; sp++) {
        int b0, nb, p0;
        spin_bands(e, sp, b0, nb, p0);
        if (nb == 0) continue;
        const double* v = sp == 0 ? e->d_vr : e->d_vr2;
        double* hb = e->d_hpair + 4L * nu * b0;
        const int rc = kept ? hpsi_resident(e->gw, v, 2 * nb, hb, nu, fprec, b0, nullptr)
                            : rs_mul_add_bands(e->gw, v, e->d_cpair + 4L * nu * b0, nu, 2 * nb, hb,
                                               nu, fprec, nullptr);
        if (rc) return rc;
      }
      if (kept) e->xres_c = nullptr;
      ctx->prof_begin(QBG_K_OTHER);
      launch_unfold_add(e->d_hpair, nu, nst, ngw, e->d_unfold, hpsi, ldh, s);
      ctx->prof_end();
    }
  } else if (compute_hpsi) {
    // with the potential of each band's spin (src/EnergyFunctional.cpp:831)
    const bool kept = reuse && e->xres_c == c && e->xres_ldc == ldc && e->xres_prec == fprec;
    for (int sp = 0; sp < e->nspin; sp++) {
      int b0, nb, p0;
      spin_bands(e, sp, b0, nb, p0);
      if (nb == 0) continue;
      const double* v = sp == 0 ? e->d_vr : e->d_vr2;
      ZFin zf;
      if (fin) {
        zf.c = c + 2L * ldc * b0;
        zf.cpst = 2L * ldc;
        zf.coff2 = ldc;
        zf.g2 = e->d_g2w;
        zf.y = e->b_yh + e->Kp * b0;
        zf.ldy = e->Kp;
      }
      const int rc =
          kept ? hpsi_resident(e->gw, v, nb, hpsi + 2L * ldh * b0, ldh, fprec, p0,
                               fin ? &zf : nullptr)
               : rs_mul_add_bands(e->gw, v, c + 2L * ldc * b0, ldc, nb, hpsi + 2L * ldh * b0,
                                  ldh, fprec, fin ? &zf : nullptr);
      if (rc) return rc;
    }
    if (kept) e->xres_c = nullptr;   // transformed in place: consumed
  }
  if (compute_forces) {
    ctx->prof_begin(QBG_K_OTHER);
    const int nchunk = 64;
    launch_local_forces(e->d_idxv, e->vb.size, e->d_sftab, e->krange, e->kmin, e->d_lfgrp,
                        e->nlfgrp, e->nsp, e->d_vps, e->d_rhops, e->d_g2iv, e->d_gxv, e->d_rhogt,
                        e->d_rhoelg, e->core ? e->d_vxcg : nullptr, e->d_rhoc, e->omega,
                        e->d_fesr, e->nat, e->d_vtemp, e->d_partial, nchunk, e->d_fion, s);
    ctx->prof_end();
  }
  // what this evaluation computed travels with its results, so a replayed
  // graph is read correctly after other evaluations ran in between
  fill_value(e->d_scal + S_FLAGS, 1, compute_forces ? 1.0 : 0.0, s);
  fill_value(e->d_scal + S_FLAGS + 1, 1, compute_stress ? 1.0 : 0.0, s);
  // stage the scalars, G=0 values and forces in pinned host memory
  gather_segments(e->d_seg_ptr, e->d_seg_n, 4, e->h_res, s);
  e->last_forces = compute_forces;
  e->last_stress = compute_stress;
  e->launched = 1;
  hipError_t er = hipGetLastError();
  if (er != hipSuccess) return ctx->fail(QBG_E_HIP, hipGetErrorString(er));
  return QBG_OK;
}

// wait for the last launched evaluation and assemble its results
static int ef_collect(qbg_ef* e, double* energies, double* fion, double* sigma) {
  qbg_ctx* ctx = e->ctx;
  if (ctx->capturing) return ctx->fail(QBG_E_STATE, "results during graph capture");
  if (!e->launched) return ctx->fail(QBG_E_STATE, "no evaluation launched");
  QBG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double* sc = e->h_res;
  const int compute_forces = sc[S_FLAGS] != 0.0, compute_stress = sc[S_FLAGS + 1] != 0.0;
  const double* g0 = e->h_res + S_COUNT;   // rhoelg[0] (2), vion[0] (2)
  if (compute_forces && fion) std::memcpy(fion, e->h_res + S_COUNT + 4, sizeof(double) * 3 * e->nat);
  const double omega = e->omega, fpi = 4.0 * M_PI, Nd = (double)e->N;
  const double exc = sc[S_LDA + 0] * omega / Nd;
  // the core term of dxc, nspin sum_G w Re(conj(vxc_g) rhocore_g)
  // (src/EnergyFunctional.cpp:290-305)
  const double dxc = sc[S_LDA + 1] * omega / Nd + (e->core ? e->nspin * sc[S_CORE] : 0.0);
  const double charge = sc[S_LDA + 2] * omega / Nd;
  const double eps = (2.0 * sc[S_VHXC + 0] - (g0[0] * g0[2] + g0[1] * g0[3])) * omega;
  const double ehart = omega * fpi * sc[S_VHXC + 1];
  const double ehart_e = omega * fpi * sc[S_VHXC + 2];
  const double ehart_ep = omega * fpi * sc[S_VHXC + 3];
  const double ehart_p = omega * fpi * sc[S_VHXC + 4];
  const double ekin = sc[S_KIN], enl = sc[S_ENL];
  const double ecoul = ehart + e->esr - e->eself;
  const double etotal = ekin + eps + enl + ecoul + exc;
  const double en[14] = {ekin, eps, enl, ehart, ehart_e, ehart_ep, ehart_p, exc, dxc,
                         e->esr, e->eself, ecoul, etotal, charge};
  std::memcpy(energies, en, sizeof(en));
  if (sigma) {
    // rows: total, ekin, eps, enl, ehart, exc, esr; columns xx yy zz xy yz xz
    // (src/EnergyFunctional.cpp:560-683, :932-936; NonLocalPotential.cpp:2057-2066)
    std::memset(sigma, 0, sizeof(double) * 42);
    if (compute_stress) {
      const double oi = 1.0 / omega;
      for (int i = 0; i < 6; i++) {
        const bool d = i < 3;
        double* sg = sigma;
        sg[6 + i] = sc[S_KIN + 1 + i] * oi;
        sg[12 + i] = sc[S_VHXC + 11 + i] + (d ? eps * oi : 0.0);
        sg[18 + i] = (d ? enl : 0.0) * oi + sc[S_NLS + i] * oi;
        sg[24 + i] = (d ? ehart * oi : 0.0) - 2.0 * fpi * (sc[S_VHXC + 5 + i] + sc[S_HPS + i]);
        sg[30 + i] = d ? -sc[S_LDA + 1] / Nd : 0.0;
        sg[36 + i] = e->sigma_esr[i];
        sg[i] = sg[6 + i] + sg[12 + i] + sg[18 + i] + sg[24 + i] + sg[30 + i] + sg[36 + i];
      }
    }
  }
  return QBG_OK;
}


int qbg_ef_energy(qbg_ef* e, const double* c, int ldc, int compute_hpsi, double* hpsi, int ldh,
                  int compute_forces, int compute_stress, int prec, double* energies,
                  double* fion, double* sigma) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !c || !energies) return QBG_E_INVAL;
    int rc = ef_launch(e, c, ldc, compute_hpsi, hpsi, ldh, compute_forces, compute_stress, prec);
    if (rc) return rc;
    return ef_collect(e, energies, fion, sigma);
  });
}

int qbg_ef_evaluate_async(qbg_ef* e, const double* c, int ldc, double* hpsi, int ldh, int flags,
                          int prec) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !c) return QBG_E_INVAL;
    const int hp = (flags & QBG_EF_HPSI) != 0, fo = (flags & QBG_EF_FORCES) != 0,
              st = (flags & QBG_EF_STRESS) != 0;
    int rc = ef_update_density(e, c, ldc, prec, hp != 0);
    if (!rc) rc = qbg_ef_update_vhxc(e, st);
    if (!rc) rc = ef_launch(e, c, ldc, hp, hpsi, ldh, fo, st, prec, true);
    return rc;
  });
}

// EnergyFunctional::energy alone, without host synchronization (after
// update_density / update_vhxc of the same iteration); results as above
int qbg_ef_energy_async(qbg_ef* e, const double* c, int ldc, double* hpsi, int ldh, int flags,
                        int prec) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !c) return QBG_E_INVAL;
    return ef_launch(e, c, ldc, (flags & QBG_EF_HPSI) != 0, hpsi, ldh, (flags & QBG_EF_FORCES) != 0,
                     (flags & QBG_EF_STRESS) != 0, prec, (flags & QBG_EF_SAME_C) != 0);
  });
}

// NonLocalPotential of one spin's SlaterDet of the functional
// (src/NonLocalPotential.h:61-68; constructed by EnergyFunctional,
// src/EnergyFunctional.cpp:113-118)
struct qbg_nl {
  qbg_ef* ef = nullptr;
  qbg_ctx* ctx = nullptr;
  int ispin = 0;
};

int qbg_nl_create(qbg_ef* e, int ispin, qbg_nl** out) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !out) return QBG_E_INVAL;
    *out = nullptr;
    if (ispin < 0 || ispin >= e->nspin) return e->ctx->fail(QBG_E_INVAL, "nl: ispin");
    if (e->nkp > 1) return e->ctx->fail(QBG_E_INVAL, "nl: one k-point per NonLocalPotential (a set runs inside qbg_ef_* only)");
    qbg_nl* nl = new qbg_nl;
    nl->ef = e;
    nl->ctx = e->ctx;
    nl->ispin = ispin;
    *out = nl;
    return QBG_OK;
  });
}

int qbg_nl_destroy(qbg_nl* nl) {
  if (!nl) return QBG_E_INVAL;
  delete nl;
  return QBG_OK;
}

// NonLocalPotential::update_twnl (src/NonLocalPotential.cpp:235-1255): the
// projector tables twnl and dtwnl of the functional's current basis
int qbg_nl_update_twnl(qbg_nl* nl) {
  return qbg::guarded((nl ? nl->ctx : nullptr), [&]() -> int {
    if (!nl) return QBG_E_INVAL;
    qbg_ef* e = nl->ef;
    qbg_ctx* c = e->ctx;
    if (c->capturing) return c->fail(QBG_E_STATE, "update_twnl during graph capture");
    QBG_HIP(c, hipSetDevice(c->device));
    QBG_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<double> twnl, dtwnl;
    ef_twnl(e, twnl, dtwnl);
    int rc;
    if ((rc = dset(c, e->d_twnl, twnl)) || (rc = dset(c, e->d_dtwnl, dtwnl))) return rc;
    return QBG_OK;
  });
}

// NonLocalPotential::energy (src/NonLocalPotential.cpp:1258-2069) for the
// bands of the handle's spin: returns enl in *enl; fion (host, 3 nat,
// species-major) and sigma_enl (host, 6: xx yy zz xy yz xz) are zeroed on
// entry and hold the nonlocal forces and stress when asked for; H*psi_NL is
// added into dsd (device, ld ldd complex).  c points at the spin's first band.
int qbg_nl_energy(qbg_nl* nl, const double* c, int ldc, int compute_hpsi, double* dsd, int ldd,
                  int compute_forces, double* fion, int compute_stress, double* sigma_enl,
                  int prec, double* enl) {
  return qbg::guarded((nl ? nl->ctx : nullptr), [&]() -> int {
    if (!nl || !c || !enl) return QBG_E_INVAL;
    qbg_ef* e = nl->ef;
    qbg_ctx* ctx = e->ctx;
    if (ctx->capturing) return ctx->fail(QBG_E_STATE, "nl energy during graph capture");
    if (ldc < e->wb.size || (compute_hpsi && (!dsd || ldd < e->wb.size)))
      return ctx->fail(QBG_E_INVAL, "nl energy: leading dimension");
    if ((compute_forces && !fion) || (compute_stress && !sigma_enl))
      return ctx->fail(QBG_E_INVAL, "nl energy: missing output");
    if (prec < QBG_PREC_FP64 || prec > QBG_PREC_BF16X2) return ctx->fail(QBG_E_INVAL, "prec");
    QBG_HIP(ctx, hipSetDevice(ctx->device));
    int b0, nb, p0;
    spin_bands(e, nl->ispin, b0, nb, p0);
    if (fion) std::memset(fion, 0, sizeof(double) * 3 * e->nat);
    if (sigma_enl) std::memset(sigma_enl, 0, sizeof(double) * 6);
    *enl = 0.0;
    if (nb == 0) return QBG_OK;
    const int rc = (prec == QBG_PREC_BF16X3 || prec == QBG_PREC_BF16X2)
                       ? nl_bf16(e, c, ldc, compute_hpsi, dsd, ldd, compute_forces,
                                 compute_stress, prec == QBG_PREC_BF16X3 ? 3 : 2, false, b0, nb)
                       : nl_mfma(e, c, ldc, compute_hpsi, dsd, ldd, compute_forces,
                                 compute_stress, prec, b0, nb);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    double sc[7];
    QBG_HIP(ctx, hipMemcpyAsync(sc, e->d_scal + S_ENL, sizeof(double) * 7, hipMemcpyDeviceToHost, s));
    if (compute_forces)
      QBG_HIP(ctx, hipMemcpyAsync(fion, e->d_fion, sizeof(double) * 3 * e->nat,
                                  hipMemcpyDeviceToHost, s));
    QBG_HIP(ctx, hipStreamSynchronize(s));
    *enl = sc[0];
    if (compute_stress) {
      // src/NonLocalPotential.cpp:2057-2066
      const double oi = 1.0 / e->omega;
      for (int i = 0; i < 6; i++) sigma_enl[i] = ((i < 3 ? sc[0] : 0.0) + sc[1 + i]) * oi;
    }
    return QBG_OK;
  });
}

int qbg_ef_results(qbg_ef* e, double* energies, double* fion, double* sigma) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !energies) return QBG_E_INVAL;
    return ef_collect(e, energies, fion, sigma);
  });
}

// device arrays for tests: 0 rhor (N), 1 v_r (N), 2 vxc_r (N), 3 rhog
// (2*ngv), 4 vion_local_g, 5 rhopst, 6 sfac (2*nsp*ngv), 7 fion_esr (3 nat)
int qbg_ef_get(qbg_ef* e, int what, double* host) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !host) return QBG_E_INVAL;
    const int ngv = e->vb.size;
    const double* src = nullptr;
    size_t n = 0;
    switch (what) {
      case 0: src = e->d_rhor; n = e->N; break;
      case 1: src = e->d_vr; n = e->N; break;
      case 2: src = e->d_vxc; n = e->N; break;
      case 3: src = e->d_rhog; n = 2 * (size_t)ngv; break;
      case 4: src = e->d_vion; n = 2 * (size_t)ngv; break;
      case 5: src = e->d_rhopst; n = 2 * (size_t)ngv; break;
      case 6: src = e->d_sfac; n = 2 * (size_t)e->nsp * ngv; break;
      case 7: std::memcpy(host, e->fesr.data(), sizeof(double) * e->fesr.size()); return QBG_OK;
      case 8: src = e->d_rhor2; n = e->N; break;
      case 9: src = e->d_vr2; n = e->N; break;
      case 10: src = e->d_vxc2; n = e->N; break;
      case 11: src = e->d_rhocr; n = e->N; break;
      case 12: src = e->d_vxcg; n = 2 * (size_t)ngv; break;
      default: return QBG_E_INVAL;
    }
    if (!src && what >= 11)
      return e->ctx->fail(QBG_E_STATE, "core density arrays exist with core densities only");
    if (!src) return e->ctx->fail(QBG_E_STATE, "spin-down arrays exist with nspin 2 only");
    QBG_HIP(e->ctx, hipMemcpyAsync(host, src, sizeof(double) * n, hipMemcpyDeviceToHost,
                                   e->ctx->stream));
    QBG_HIP(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return QBG_OK;
  });
}

// FourierTransform(basis, basis.np(0), basis.np(1), basis.np(2)) of the
// wavefunction basis (src/SlaterDet.cpp:1385, src/WavefunctionHandler.cpp:411):
// the grid of the grid functions in sample files
int qbg_ef_wf_grid(qbg_ef* e, int* dims, qbg_grid** out) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !out) return QBG_E_INVAL;
    *out = nullptr;
    qbg_ctx* c = e->ctx;
    if (e->cplx) return c->fail(QBG_E_INVAL, "wf_grid: real (Gamma) basis only");
    if (!e->gwf) {
      if (c->capturing) return c->fail(QBG_E_STATE, "wf_grid during graph capture");
      qbg::host::Maps m;
      if (!m.build(e->wb, e->wb.np[0], e->wb.np[1], e->wb.np[2]))
        return c->fail(QBG_E_INVAL, "wf_grid: basis does not fit its grid");
      int rc = qbg_grid_create(c, e->wb.np[0], e->wb.np[1], e->wb.np[2], e->wb.size, m.nvec,
                               m.ip.data(), m.im.data(), m.zvec_to_val.data(), &e->gwf);
      if (rc) return rc;
    }
    if (dims)
      for (int i = 0; i < 3; i++) dims[i] = e->wb.np[i];
    *out = e->gwf;
    return QBG_OK;
  });
}

// the launch plan of the functional's TN projection product (tuned at the
// first 2-part BF16 evaluation): out[0] 1 once tuned, out[1] tile edge,
// out[2] reduction splits (0: the default rule)
int qbg_ef_gemm_plan(const qbg_ef* e, int* out) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !out) return QBG_E_INVAL;
    out[0] = e->bf_tuned ? 1 : 0;
    out[1] = e->bf_tn.edge;
    out[2] = e->bf_tn.split;
    return QBG_OK;
  });
}

// the grid of the band transforms (ChargeDensity's FourierTransform on the
// density grid, src/ChargeDensity.cpp:44-49, 98), owned by ef
int qbg_ef_band_grid(qbg_ef* e, qbg_grid** out) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !out) return QBG_E_INVAL;
    if (e->cplx) return e->ctx->fail(QBG_E_INVAL, "band_grid: real (Gamma) basis only");
    *out = e->gw;
    return e->gw ? QBG_OK : e->ctx->fail(QBG_E_STATE, "band grid not built");
  });
}

// Wavefunction::set_nspin (src/Wavefunction.cpp:365-390) with the bands
// [0, nst_up) spin up: per-spin densities and potentials (ChargeDensity
// rhor[ispin], EnergyFunctional v_r[ispin], src/EnergyFunctional.h:78-79)
int qbg_ef_set_nspin(qbg_ef* e, int nspin, int nst_up) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e) return QBG_E_INVAL;
    qbg_ctx* c = e->ctx;
    if (c->capturing) return c->fail(QBG_E_STATE, "set_nspin during graph capture");
    if (nspin == 1) nst_up = 0;
    if ((nspin != 1 && nspin != 2) || nst_up < 0 || nst_up > e->nst)
      return c->fail(QBG_E_INVAL, "set_nspin: nspin must be 1 or 2, 0 <= nst_up <= nst");
    QBG_HIP(c, hipSetDevice(c->device));
    QBG_HIP(c, hipStreamSynchronize(c->stream));
    if (nspin == 2 && !e->d_rhor2) {
      int rc;
      if ((rc = dalloc(c, &e->d_rhor2, e->N)) || (rc = dalloc(c, &e->d_vxc2, e->N)) ||
          (rc = dalloc(c, &e->d_vr2, e->N)) || (rc = dalloc(c, &e->d_vloc, e->N)) ||
          (rc = dalloc(c, &e->d_rhot, e->N)))
        return rc;
    }
    e->nspin = nspin;
    e->nst_up = nst_up;
    std::vector<double> dfac = density_factors(e);
    if (e->d_dfac) (void)hipFree(e->d_dfac);
    e->d_dfac = nullptr;
    int rc = dput(c, &e->d_dfac, dfac);
    if (rc) return rc;
    if ((rc = ef_core_density(e))) return rc;      // each spin's share of the core
    e->have_rho = e->have_vhxc = false;
    e->xres_c = nullptr;
    return QBG_OK;
  });
}

// the potential v_r used by H*psi may be replaced (tests of rs_mul_add)
int qbg_ef_set_vr(qbg_ef* e, const double* host_v) {
  return qbg::guarded((e ? e->ctx : nullptr), [&]() -> int {
    if (!e || !host_v) return QBG_E_INVAL;
    QBG_HIP(e->ctx, hipMemcpyAsync(e->d_vr, host_v, sizeof(double) * e->N, hipMemcpyHostToDevice,
                                   e->ctx->stream));
    QBG_HIP(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return QBG_OK;
  });
}

// SlaterDet::gram (src/SlaterDet.cpp:480-534), real basis:
// cleanup; S = 2 P^T P - p0 p0^T; S = L L^T; C <- C L^-T
int qbg_gram(qbg_ctx* c, double* cc, int ldc, int ngw, int nst, int prec) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !cc || ngw <= 0 || nst <= 0 || ldc < ngw) return QBG_E_INVAL;
    if (prec != QBG_PREC_FP64) return c->fail(QBG_E_INVAL, "gram: FP64 only");
    QBG_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const long K = 2L * ngw;
    const int tm = (nst + 63) / 64;
    const size_t work = (size_t)nst * nst * std::max(1, 2048 / (tm * tm) + 1);
    double *S = nullptr, *X = nullptr, *T = nullptr, *W = nullptr;
    int* info = nullptr;
    int rc = QBG_OK;
    if ((rc = dalloc(c, &S, (size_t)nst * nst)) || (rc = dalloc(c, &X, (size_t)nst * nst)) ||
        (rc = dalloc(c, &T, (size_t)K * nst)) || (rc = dalloc(c, &W, work)) ||
        (rc = dalloc(c, &info, 1))) {
      (void)hipFree(S); (void)hipFree(X); (void)hipFree(T); (void)hipFree(W); (void)hipFree(info);
      return rc;
    }
    auto release = [&]() {
      (void)hipFree(S); (void)hipFree(X); (void)hipFree(T); (void)hipFree(W); (void)hipFree(info);
    };
    c->prof_begin(QBG_K_GRAM);
    cleanup(cc, ldc, nst, s);
    if (gemm_syrk(cc, 2L * ldc, nst, K, 2.0, -1.0, S, nst, W, work, s)) {
      (void)hipStreamSynchronize(s);
      release();
      return c->fail(QBG_E_INVAL, "gram workspace");
    }
    potrf(S, nst, nst, info, s, W);
    c->prof_end();
    // the coefficients are only touched once the factorization succeeded
    // (the reference aborts in potrf, src/Matrix.cpp:2324-2327)
    int hinfo = 0;
    (void)hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t er = hipStreamSynchronize(s);
    if (er != hipSuccess || hinfo) {
      release();
      if (er != hipSuccess) return c->fail(QBG_E_HIP, hipGetErrorString(er));
      return c->fail(QBG_E_NOT_SPD, "gram: overlap matrix not positive definite");
    }
    c->prof_begin(QBG_K_GRAM);
    trtri(S, nst, nst, X, s);
    gemm_nn_triu(cc, 2L * ldc, X, nst, K, nst, T, K, s);
    (void)hipMemcpy2DAsync(cc, sizeof(double) * 2 * ldc, T, sizeof(double) * K, sizeof(double) * K,
                           nst, hipMemcpyDeviceToDevice, s);
    c->prof_end();
    er = hipStreamSynchronize(s);
    release();
    if (er != hipSuccess) return c->fail(QBG_E_HIP, hipGetErrorString(er));
    return rc;
  });
}

}  // extern "C"

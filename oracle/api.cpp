// api.cpp -- C entry points of the CPU oracle (ctypes binding for tests and
// the CPU baseline).  TEST INFRASTRUCTURE ONLY (see qbo.h).
#include "qbo.h"
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

using namespace qbo;

extern "C" {

typedef struct {
  int zval;
  int lmax;
  int nmesh;
  double deltar;
  double rcps;
  const double* vloc;   // nmesh
  const double* proj;   // (lmax+1) * nmesh
  const double* dl;     // lmax+1
  int type;             // 0 SLPP, 1 NCPP (vps, phi, llocal)
  int llocal;
  const double* vps;    // (lmax+1) * nmesh
  const double* phi;    // (lmax+1) * nmesh
  const double* nlcc;   // core density (nmesh) or null
  int nquad;            // NCPP: semilocal quadrature form when > 0
  double rquad;
} qbo_species_desc;

static std::string g_err;
const char* qbo_last_error(void) { return g_err.c_str(); }

// kpoint3 (null: k = 0) in units of the reciprocal vectors; force_complex
// builds the complex basis at k = 0 too
void* qbo_create_k(const double* cell9, const double* refcell9, double ecut_ha,
                   int nsp, const qbo_species_desc* sp, const int* na,
                   const double* tau, int nst, const double* occ, const int* grid3,
                   const double* kpoint3, int force_complex) {
  try {
    System* s = new System;
    s->cell.set(cell9);
    if (kpoint3) for (int i = 0; i < 3; i++) s->kpoint[i] = kpoint3[i];
    s->force_complex = force_complex != 0;
    if (grid3) for (int i = 0; i < 3; i++) s->grid[i] = grid3[i];
    if (refcell9) { s->has_ref = true; s->refcell.set(refcell9); }
    s->ecut = ecut_ha;
    int off = 0;
    for (int is = 0; is < nsp; is++) {
      Species q;
      q.zval = sp[is].zval; q.lmax = sp[is].lmax; q.deltar = sp[is].deltar;
      q.rcps = sp[is].rcps;
      if (sp[is].type == 1 && sp[is].nquad > 0) {
        // semilocal quadrature form: the semilocal potentials as given
        const int nm = sp[is].nmesh;
        q.nquad = sp[is].nquad; q.rquad = sp[is].rquad; q.llocal = sp[is].llocal;
        if (q.llocal < 0 || q.llocal > q.lmax || !(q.rquad > 0.0))
          throw std::runtime_error("semilocal quadrature species: llocal or rquad");
        q.vsl.resize(q.lmax + 1);
        for (int l = 0; l <= q.lmax; l++)
          q.vsl[l].assign(sp[is].vps + (size_t)l * nm, sp[is].vps + (size_t)(l + 1) * nm);
        q.vlocal = q.vsl[q.llocal];
      } else if (sp[is].type == 1) {
        ncpp_species(q, sp[is].lmax, sp[is].llocal, sp[is].nmesh, sp[is].vps, sp[is].phi);
      } else {
        q.vlocal.assign(sp[is].vloc, sp[is].vloc + sp[is].nmesh);
        q.proj.resize(q.lmax + 1);
        for (int l = 0; l <= q.lmax; l++)
          q.proj[l].assign(sp[is].proj + (size_t)l * sp[is].nmesh,
                           sp[is].proj + (size_t)(l + 1) * sp[is].nmesh);
        q.dl.assign(sp[is].dl, sp[is].dl + q.lmax + 1);
      }
      if (sp[is].nlcc) q.nlcc.assign(sp[is].nlcc, sp[is].nlcc + sp[is].nmesh);
      s->sp.push_back(q);
      s->tau.emplace_back(tau + 3 * off, tau + 3 * (off + na[is]));
      off += na[is];
    }
    s->nst = nst;
    s->occ.assign(occ, occ + nst);
    s->setup();
    return s;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void* qbo_create(const double* cell9, const double* refcell9, double ecut_ha,
                 int nsp, const qbo_species_desc* sp, const int* na,
                 const double* tau, int nst, const double* occ, const int* grid3) {
  return qbo_create_k(cell9, refcell9, ecut_ha, nsp, sp, na, tau, nst, occ, grid3, nullptr, 0);
}

// several k-points (kpoints3n: 3 per k-point, units of b; weights summing
// to 1); every k-point on the complex basis
void* qbo_create_kset(const double* cell9, const double* refcell9, double ecut_ha,
                      int nsp, const qbo_species_desc* sp, const int* na,
                      const double* tau, int nst, const double* occ, const int* grid3,
                      int nkp, const double* kpoints3n, const double* weights) {
  if (nkp < 1 || !kpoints3n || !weights) { g_err = "kset: nkp, kpoints, weights"; return nullptr; }
  if (nkp == 1) return qbo_create_k(cell9, refcell9, ecut_ha, nsp, sp, na, tau, nst, occ, grid3,
                                    kpoints3n, 1);
  // the set is attached before setup: create with k-point 0, then rebuild
  void* h = qbo_create_k(cell9, refcell9, ecut_ha, nsp, sp, na, tau, nst, occ, grid3,
                         kpoints3n, 1);
  if (!h) return nullptr;
  try {
    System* s = (System*)h;
    s->kpts.assign(kpoints3n, kpoints3n + 3 * nkp);
    s->kwts.assign(weights, weights + nkp);
    s->setup();
    return s;
  } catch (std::exception& e) {
    g_err = e.what();
    delete (System*)h;
    return nullptr;
  }
}

// k-point ik of a set (ik = 0 also for a single k-point): its basis size,
// and idx (3 ngw), |k+G|^2, k+G (3 ngw) when the pointers are given
int qbo_kset_basis(void* h, int ik, int* idx, double* kpg2, double* kpgx) {
  System* s = (System*)h;
  if (ik < 0 || (ik > 0 && ik >= (int)s->kset.size())) return -1;
  const Basis& b = ik == 0 ? s->wb : s->kset[ik].wb;
  if (idx) std::memcpy(idx, b.idx.data(), sizeof(int) * 3 * b.size);
  if (kpg2) std::memcpy(kpg2, b.kpg2.data(), sizeof(double) * b.size);
  if (kpgx) std::memcpy(kpgx, b.kpgx.data(), sizeof(double) * 3 * b.size);
  return b.size;
}

void qbo_destroy(void* h) { delete (System*)h; }

// the wavefunction basis's k + G: |k+G|^2 (ngw) and its components (3 ngw);
// returns 1 for a real basis, 0 for a complex one
int qbo_kpg(void* h, double* kpg2, double* kpgx) {
  System* s = (System*)h;
  const Basis& b = s->wb;
  if (kpg2) std::memcpy(kpg2, b.kpg2.data(), sizeof(double) * b.size);
  if (kpgx) std::memcpy(kpgx, b.kpgx.data(), sizeof(double) * 3 * b.size);
  return b.real ? 1 : 0;
}

// the Kleinman-Bylander form of an NCPP species: proj[(lkb+1)*nmesh], dl[lkb+1]
// (room for lmax+1 rows); returns lkb or -1
int qbo_ncpp_kb(int lmax, int llocal, int nmesh, double deltar, const double* vps,
                const double* phi, double* proj, double* dl) {
  try {
    Species q;
    q.deltar = deltar;
    ncpp_species(q, lmax, llocal, nmesh, vps, phi);
    for (int l = 0; l <= q.lmax; l++) {
      std::memcpy(proj + (size_t)l * nmesh, q.proj[l].data(), 8 * (size_t)nmesh);
      dl[l] = q.dl[l];
    }
    return q.lmax;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// out[0..]: ngw, ngv, np0, np1, np2, nvec_w, ntrans0_w, nvec_v, ntrans0_v,
//           nrods_w, nrods_v, ndft(species 0)
void qbo_sizes(void* h, int* out) {
  System* s = (System*)h;
  out[0] = s->wb.size; out[1] = s->vb.size;
  out[2] = s->ft.map.np0; out[3] = s->ft.map.np1; out[4] = s->ft.map.np2;
  out[5] = s->ft.map.nvec; out[6] = s->ft.map.ntrans0;
  out[7] = s->vft.map.nvec; out[8] = s->vft.map.ntrans0;
  out[9] = s->wb.nrods; out[10] = s->vb.nrods;
  out[11] = s->sp.empty() ? 0 : s->sp[0].ndft;
}

// projectors per atom of species is (one per (l, m), l <= lmax of the KB form)
int qbo_nlm(void* h, int is) { return ((System*)h)->sp[is].nlm; }
// radial projector channels of species is (lmax + 1 of the KB form)
int qbo_nop(void* h, int is) { return ((System*)h)->sp[is].nop; }

// which = 0: wavefunction basis (on the density grid), 1: density basis
void qbo_basis(void* h, int which, int* idx, double* g2, double* gx, int* ip,
               int* im, int* zvec_to_val, int* np_basis) {
  System* s = (System*)h;
  const Basis& b = which ? s->vb : s->wb;
  const FFT3& f = which ? s->vft : s->ft;
  if (idx) std::memcpy(idx, b.idx.data(), sizeof(int) * 3 * b.size);
  if (g2) std::memcpy(g2, b.g2.data(), sizeof(double) * b.size);
  if (gx) std::memcpy(gx, b.gx.data(), sizeof(double) * 3 * b.size);
  if (ip) std::memcpy(ip, f.map.ip.data(), sizeof(int) * b.size);
  if (im) std::memcpy(im, f.map.im.data(), sizeof(int) * b.size);
  if (zvec_to_val) std::memcpy(zvec_to_val, f.map.zvec_to_val.data(), sizeof(int) * f.map.nvec);
  if (np_basis) { np_basis[0] = b.np[0]; np_basis[1] = b.np[1]; np_basis[2] = b.np[2];
                  for (int i = 0; i < 3; i++) { np_basis[3 + i] = b.idxmin[i]; np_basis[6 + i] = b.idxmax[i]; } }
}

void qbo_rods(void* h, int which, int* rh, int* rk, int* rlmin, int* rsize) {
  System* s = (System*)h;
  const Basis& b = which ? s->vb : s->wb;
  for (int i = 0; i < b.nrods; i++) {
    rh[i] = b.rod_h[i]; rk[i] = b.rod_k[i]; rlmin[i] = b.rod_lmin[i]; rsize[i] = b.rod_size[i];
  }
}

// transforms on the wavefunction (which=0) or density (which=1) basis
void qbo_fft_backward(void* h, int which, const double* c, double* f) {
  System* s = (System*)h;
  (which ? s->vft : s->ft).backward((const cplx*)c, (cplx*)f);
}
void qbo_fft_backward_pair(void* h, const double* c1, const double* c2, double* f) {
  System* s = (System*)h;
  s->ft.backward((const cplx*)c1, (const cplx*)c2, (cplx*)f);
}
void qbo_fft_forward(void* h, int which, double* f, double* c) {
  System* s = (System*)h;
  (which ? s->vft : s->ft).forward((cplx*)f, (cplx*)c);
}
void qbo_fft_forward_pair(void* h, double* f, double* c1, double* c2) {
  System* s = (System*)h;
  s->ft.forward((cplx*)f, (cplx*)c1, (cplx*)c2);
}

void qbo_compute_density(void* h, const double* c, double* rho) {
  System* s = (System*)h;
  compute_density(s->ft, (const cplx*)c, s->wb.size, s->nst, s->occ.data(), 1.0,
                  s->cell.volume, rho);
}
void qbo_rs_mul_add(void* h, const double* v, const double* c, int nst, double* hpsi) {
  System* s = (System*)h;
  rs_mul_add(s->ft, v, (const cplx*)c, s->wb.size, nst, (cplx*)hpsi);
}
void qbo_compute_density_occ(void* h, const double* c, int nst, const double* occ,
                             double* rho) {
  System* s = (System*)h;
  compute_density(s->ft, (const cplx*)c, s->wb.size, nst, occ, 1.0, s->cell.volume, rho);
}
int qbo_gram(double* c, int ngw, int nst) { return gram((cplx*)c, ngw, nst); }
int qbo_gram_complex(double* c, int ngw, int nst) { return gram_complex((cplx*)c, ngw, nst); }

void qbo_update_density(void* h, const double* c) {
  ((System*)h)->update_density((const cplx*)c);
}
void qbo_update_vhxc(void* h, int stress) { ((System*)h)->update_vhxc(stress != 0); }

// arrays: 0 rhor, 1 v_r, 2 vxc_r (N doubles); 3 rhog, 4 vion_local_g,
// 5 rhopst (ngv complex); 6 vps, 7 rhops, 8 dvps of species `is` (ngv);
// 9 twnl of species is (npr*ngw); 10 sfac of species is (ngv complex);
// 11 fion_esr (3 nat); 12 dtwnl (npr*6*ngw); 13 dvion_local_g;
// 14 rhor, 15 v_r, 16 vxc_r of spin down (nspin 2; arrays 0-2 are spin up);
// 17 rhocore_r (N), 18 vxc_g (ngv complex), 19 rhocore_sp of species is (ngv)
// (core density systems only, after update_vhxc)
void qbo_get(void* h, int what, int is, double* out) {
  System* s = (System*)h;
  auto cpy = [&](const void* p, size_t bytes) { std::memcpy(out, p, bytes); };
  switch (what) {
    case 0: cpy(s->rhor.data(), 8 * s->rhor.size()); break;
    case 1: cpy(s->v_r.data(), 8 * s->v_r.size()); break;
    case 2: cpy(s->vxc_r.data(), 8 * s->vxc_r.size()); break;
    case 3: cpy(s->rhog.data(), 16 * s->rhog.size()); break;
    case 4: cpy(s->vion_local_g.data(), 16 * s->vion_local_g.size()); break;
    case 5: cpy(s->rhopst.data(), 16 * s->rhopst.size()); break;
    case 6: cpy(s->vps[is].data(), 8 * s->vps[is].size()); break;
    case 7: cpy(s->rhops[is].data(), 8 * s->rhops[is].size()); break;
    case 8: cpy(s->dvps[is].data(), 8 * s->dvps[is].size()); break;
    case 9: cpy(s->twnl[is].data(), 8 * s->twnl[is].size()); break;
    case 10: cpy(s->sfac[is].data(), 16 * s->sfac[is].size()); break;
    case 11: cpy(s->fion_esr.data(), 8 * s->fion_esr.size()); break;
    case 12: cpy(s->dtwnl[is].data(), 8 * s->dtwnl[is].size()); break;
    case 13: cpy(s->dvion_local_g.data(), 16 * s->dvion_local_g.size()); break;
    case 14: cpy(s->rhor_dn.data(), 8 * s->rhor_dn.size()); break;
    case 15: cpy(s->v_r_dn.data(), 8 * s->v_r_dn.size()); break;
    case 16: cpy(s->vxc_r_dn.data(), 8 * s->vxc_r_dn.size()); break;
    case 17: cpy(s->rhocore_r.data(), 8 * s->rhocore_r.size()); break;
    case 18: cpy(s->vxc_g.data(), 16 * s->vxc_g.size()); break;
    case 19: if (s->core) cpy(s->rhocore_sp[is].data(), 8 * s->rhocore_sp[is].size()); break;
  }
}

// nspin 2 with bands [0, nst_up) up (Wavefunction::set_nspin)
int qbo_set_nspin(void* h, int nspin, int nst_up) {
  System* s = (System*)h;
  if (nspin == 1) { s->nspin = 1; s->nst_up = 0; return 0; }
  if (nspin != 2 || nst_up < 0 || nst_up > s->nst) return -1;
  s->nspin = 2;
  s->nst_up = nst_up;
  return 0;
}

void qbo_lsda(int n, const double* rup, const double* rdn, double* e, double* vup,
              double* vdn) {
  for (int i = 0; i < n; i++) lsda(rup[i], rdn[i], e[i], vup[i], vdn[i]);
}
void qbo_set_vr(void* h, const double* v) {
  System* s = (System*)h;
  std::memcpy(s->v_r.data(), v, 8 * s->v_r.size());
}

// energies: ekin eps enl ehart ehart_e ehart_ep ehart_p exc dxc esr eself
//           ecoul etotal total_charge
// sigma: 7 x 6 (total, ekin, eps, enl, ehart, exc, esr)
int qbo_energy(void* h, const double* c, int compute_hpsi, double* hpsi,
               int compute_forces, int compute_stress, double* energies,
               double* fion, double* sigma) {
  System* s = (System*)h;
  try {
    Results r;
    s->energy((const cplx*)c, compute_hpsi != 0, (cplx*)hpsi, compute_forces != 0,
              compute_stress != 0, r);
    const double e[14] = {r.ekin, r.eps, r.enl, r.ehart, r.ehart_e, r.ehart_ep,
                          r.ehart_p, r.exc, r.dxc, r.esr, r.eself, r.ecoul,
                          r.etotal, s->total_charge};
    std::memcpy(energies, e, sizeof(e));
    if (fion && compute_forces) std::memcpy(fion, r.fion.data(), 8 * r.fion.size());
    if (sigma) {
      const double* parts[7] = {r.sigma, r.sigma_ekin, r.sigma_eps, r.sigma_enl,
                                r.sigma_ehart, r.sigma_exc, r.sigma_esr};
      for (int k = 0; k < 7; k++) std::memcpy(sigma + 6 * k, parts[k], 48);
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// species radial tables of species is at n points g: v, dv (local), and
// vnl, dvnl for each projector (l-major)
// NonLocalPotential::energy alone (enl returned; hpsi += H*psi_NL)
double qbo_nl_energy(void* h, const double* c, int compute_hpsi, double* hpsi, int forces,
                     int stress, double* fion, double* sigma6) {
  System* s = (System*)h;
  std::vector<double> f;
  double sig[6];
  const double enl = s->nl_energy((const cplx*)c, compute_hpsi != 0, (cplx*)hpsi, forces != 0,
                                  f, stress != 0, sig);
  if (fion) std::memcpy(fion, f.data(), sizeof(double) * f.size());
  if (sigma6) std::memcpy(sigma6, sig, sizeof(sig));
  return enl;
}

// core density transform of species is at n points g: rho, drho/dg, and
// its G = 0 value (the integral of the core density) as the return value
double qbo_species_core(void* h, int is, int n, const double* g, double* rho, double* drho) {
  const Species& q = ((System*)h)->sp[is];
  for (int i = 0; i < n; i++) q.drhocoreg(g[i], rho[i], drho[i]);
  return q.zcore;
}

void qbo_species_eval(void* h, int is, int n, const double* g, double* vloc,
                      double* dvloc, double* vnl, double* dvnl) {
  System* s = (System*)h;
  const Species& q = s->sp[is];
  for (int i = 0; i < n; i++) {
    q.dvlocg(g[i], vloc[i], dvloc[i]);
    for (int iop = 0; iop < q.nop; iop++)
      q.dvnlg(iop, g[i], vnl[iop * n + i], dvnl[iop * n + i]);
  }
}

void qbo_lda(int n, const double* rho, double* e, double* v) {
  for (int i = 0; i < n; i++) lda_unpolarized(rho[i], e[i], v[i]);
}
void qbo_sinft(int n, double* y) { sinft(n, y); }
void qbo_cosft1(int n, double* y) { cosft1(n, y); }
void qbo_spline(int n, const double* x, const double* y, double yp1, double ypn,
                int bl, int br, double* y2) { spline(n, x, y, yp1, ypn, bl, br, y2); }
void qbo_splintd(int n, const double* xa, const double* ya, const double* y2a,
                 double x, double* y, double* dy) { splintd(n, xa, ya, y2a, x, y, dy); }
void qbo_fft1d(double* x, int n, int sign) { fft1d((cplx*)x, n, 1, sign); }
void qbo_cell(const double* a9, double* b9, double* vol) {
  Cell c; c.set(a9);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) b9[3 * i + j] = c.b[i][j];
  *vol = c.volume;
}
void qbo_fold_in_ws(const double* a9, double* v) { Cell c; c.set(a9); c.fold_in_ws(v); }

// ------------------------------------------------------ SCF iteration pieces
double qbo_sd_dot(const double* c, const double* d, int ngw, int nst) {
  return sd_dot((const cplx*)c, (const cplx*)d, ngw, nst);
}
void qbo_prec_ekin(const double* c, const double* g2, int ngw, int nst, double* ekin) {
  prec_ekin((const cplx*)c, g2, ngw, nst, ekin);
}
double qbo_prec_diag(double ekin, double g2, double ecutprec) {
  return prec_diag(ekin, g2, ecutprec);
}
void qbo_residual(const double* c, double* d, int ngw, int nst, double* a) {
  residual((const cplx*)c, (cplx*)d, ngw, nst, a);
}
int qbo_polar(double* x, int n, double tol, int maxiter) { return polar(x, n, tol, maxiter); }
int qbo_ortho_align(double* c, const double* cl, int ngw, int nst) {
  return ortho_align((cplx*)c, (const cplx*)cl, ngw, nst);
}
int qbo_wf_update(int kind, double param, const double* g2, double* c, double* d, int ngw,
                  int nst, double* c_last, double* d_last, int* extrapolate) {
  return wf_update(kind, param, g2, (cplx*)c, (cplx*)d, ngw, nst, (cplx*)c_last, (cplx*)d_last,
                   extrapolate);
}
void* qbo_anderson_create(long m, int nmax) {
  Anderson* a = new Anderson;
  a->m = m;
  a->nmax = nmax;
  a->xs.assign((size_t)(nmax + 1) * m, 0.0);
  a->fs.assign((size_t)(nmax + 1) * m, 0.0);
  a->restart();
  return a;
}
void qbo_anderson_update(void* a, const double* x, const double* f, double* xbar, double* fbar) {
  ((Anderson*)a)->update(x, f, xbar, fbar);
}
void qbo_anderson_destroy(void* a) { delete (Anderson*)a; }
void* qbo_mixer_create(void* h, double alpha, double rc_kerker, int ndim, double rc1) {
  Mixer* m = new Mixer;
  m->init(*(System*)h, alpha, rc_kerker, ndim, rc1);
  return m;
}
void qbo_mixer_restart(void* m) { ((Mixer*)m)->restart(); }
void qbo_mixer_update(void* m, void* h) { ((Mixer*)m)->update(*(System*)h); }
void qbo_mixer_destroy(void* m) { delete (Mixer*)m; }
void qbo_update_rhor(void* h) { ((System*)h)->update_rhor(); }

}  // extern "C"

// ef.h -- the device-resident state of one system (qbg_ef): Gamma, one
// k-point off Gamma, or a k-point set,
// shared by the energy functional (ef.cpp) and the SCF-iteration steppers
// and charge mixer (scf_api.cpp).
#pragma once
#include <vector>
#include "energy.h"
#include "internal.h"
#include "setup.h"

struct qbg_ef {
  qbg_ctx* ctx = nullptr;
  qbg::host::Cell cell;
  // the cell that fixes basis membership and grid (Wavefunction's refcell,
  // src/Basis.cpp:314-325; the cell itself when none was given)
  qbg::host::Cell defcell;
  bool has_ref = false;
  double ecut = 0, omega = 0;
  std::vector<qbg::host::Species> sp;
  std::vector<int> na, sp_first;
  int nsp = 0, nat = 0, nst = 0;
  std::vector<double> tau, occ;
  qbg::host::Basis wb, vb;
  qbg::host::Maps wm, vm;
  int np[3] = {0, 0, 0};
  long N = 0;
  qbg_grid* gw = nullptr;
  // the wavefunction basis on its own grid (Basis::np), for sample I/O;
  // built on first use by qbg_ef_wf_grid
  qbg_grid* gwf = nullptr;
  qbg_grid* gv = nullptr;
  int kmin[3] = {0, 0, 0}, krange = 0;
  // device: bases and tables
  double *d_g2w = nullptr, *d_gxw = nullptr;
  int32_t* d_idxv = nullptr;
  double *d_g2iv = nullptr, *d_giv = nullptr, *d_gxv = nullptr;
  double *d_vps = nullptr, *d_dvps = nullptr, *d_rhops = nullptr, *d_twnl = nullptr;
  double *d_tau = nullptr, *d_b = nullptr, *d_sftab = nullptr, *d_sfac = nullptr;
  double *d_vion = nullptr, *d_dvion = nullptr, *d_rhopst = nullptr;
  int32_t *d_sp_first = nullptr, *d_atom_sp = nullptr;
  // local forces: atom groups (first, count, species) and vtemp per species
  int32_t* d_lfgrp = nullptr;
  int nlfgrp = 0;
  double* d_vtemp = nullptr;
  // device: state
  double *d_rhor = nullptr, *d_vxc = nullptr, *d_vr = nullptr, *d_rhog = nullptr;
  double *d_rhoelg = nullptr, *d_rhogt = nullptr, *d_vlocal = nullptr, *d_occ = nullptr;
  double *d_fesr = nullptr, *d_fion = nullptr, *d_scal = nullptr, *d_partial = nullptr;
  size_t partial_doubles = 0;
  // nonlocal
  int M = 0;
  int32_t *d_m_atom = nullptr, *d_m_twnl = nullptr, *d_m_l = nullptr, *d_m_ipr = nullptr,
          *d_m_npr = nullptr, *d_m_stride = nullptr, *d_m_dbase = nullptr;
  double *d_m_wt = nullptr, *d_dmat = nullptr;
  int32_t *d_a_m0 = nullptr, *d_a_npr = nullptr, *d_a_stride = nullptr;
  // stress: dtwnl[s] at dt_off[s], layout [ipr][ij][ig]; npr per species;
  // 2 (rcps/2)^2 per species
  double* d_dtwnl = nullptr;
  int64_t* d_dt_off = nullptr;
  int32_t* d_sp_npr = nullptr;
  double* d_fac2 = nullptr;
  double *d_anl = nullptr, *d_fnl = nullptr, *d_fout = nullptr, *d_work = nullptr;
  size_t work_doubles = 0;
  // BF16-split GEMM workspace (allocated on first use): FP32 arrays padded
  // to Kp rows (2 ngw up to a multiple of 128), Mp projector and Np band
  // columns (multiples of 128), zero padding
  int Mp = 0, Np = 0;
  long Kp = 0;
  float *b_anl = nullptr, *b_c32 = nullptr, *b_y32 = nullptr, *b_fout = nullptr,
        *b_w32 = nullptr, *b_part = nullptr;
  // the nonlocal part of H*psi (Kp x Np, FP32) when the forward z pass
  // completes H*psi (qbg::ZFin)
  float* b_yh = nullptr;
  size_t b_part_floats = 0;
  // launch plan of the TN projections, timed at the first 2-part BF16 use
  qbg::BfTnPlan bf_tn;
  bool bf_tuned = false;
  // fused force product (256-tile shapes): per-tile partials (M x 3 x Kp/256)
  // and their per-projector sums
  double *b_fpart = nullptr, *b_fred = nullptr;
  // per-band density factors occ/omega (2 * pairs) and the result staging:
  // d_scal, rhoelg[0], vion[0], fion gathered by a kernel into pinned host
  // memory, so an evaluation needs no host synchronization until collected
  double* d_dfac = nullptr;
  double* h_res = nullptr;
  const double** d_seg_ptr = nullptr;
  int* d_seg_n = nullptr;
  int nres = 0;
  int last_forces = 0, last_stress = 0, launched = 0;
  // host results of atoms_moved
  double esr = 0, sigma_esr[6] = {0, 0, 0, 0, 0, 0}, eself = 0;
  std::vector<double> fesr;
  bool have_rho = false, have_vhxc = false, vhxc_stress = false;
  // nspin 2 (qbg_ef_set_nspin): bands [0, nst_up) spin up, [nst_up, nst)
  // spin down; d_rhor, d_vxc, d_vr are then the spin-up arrays and these the
  // spin-down ones, d_vloc = Re(vlocal) on the grid, d_rhot the total density
  int nspin = 1, nst_up = 0;
  double *d_rhor2 = nullptr, *d_vxc2 = nullptr, *d_vr2 = nullptr, *d_vloc = nullptr,
         *d_rhot = nullptr;
  // non-linear core correction (any species with a core density): per
  // species rhocoreg/omega and drhocoreg gi/(4 pi omega) [is][ig], the core
  // density in G (each spin's share) and on the grid, vxc in G (spin
  // average), written by update_vhxc for the forces and stress
  bool core = false;
  double *d_rhoc = nullptr, *d_drhoc = nullptr, *d_rhocg = nullptr, *d_rhocr = nullptr,
         *d_vxcg = nullptr;
  // the bands (pointer, leading dimension, band-FFT precision) whose x-pass
  // output the last update_density left in the band grid's resident array;
  // xres_c = nullptr when there is none (consumed by an H*psi or not kept)
  const double* xres_c = nullptr;
  int xres_ldc = 0, xres_prec = -1;
  // one k-point off Gamma (qbg_ef_create_k): wb is then the complex basis
  // |k+G| < sqrt(2 ecut) and d_g2w, d_gxw hold |k+G|^2 and k+G; the band
  // transforms run on ub, a real basis holding every G of wb and its -G, on
  // which each complex band travels as two real bands (the real and the
  // imaginary part of psi(r)): fold[2p], fold[2p+1] = the wb index of G_p
  // and of -G_p (or -1), unfold[j] = p (G_j is ub vector p) or -1-p (-G_j
  // is); cpair/hpair: the folded bands and their H*psi (ub.size x 2 nst);
  // occh = occ/2 (the full sphere's kinetic and force sums); the nonlocal
  // projectors are doubled, M = 2 Mh: columns Mh.. are i times columns 0..
  // (phase (-i)^(l+3)), a_m0h = a_m0 + Mh
  bool cplx = false;
  // a k-point set (qbg_ef_create_kset, nkp > 1, every k-point complex):
  // k-point ik's bands are the columns [ik nst, (ik+1) nst) of c and H*psi;
  // wb and the per-k-point tables (d_g2w, d_gxw, d_twnl, d_dtwnl, d_occ,
  // d_occh, d_m_twnl, d_dt_off, d_fold, d_unfold; occupations times the
  // weight) hold k-point 0 between calls and kswap exchanges them with
  // kset[ik]; kacc: each k-point's kinetic (8), enl (1) and nonlocal stress
  // (6) sums, added into the scalars after the loop; o_kin, o_enl, o_nls:
  // where the kinetic and nonlocal passes put those sums (the scalars for
  // one k-point); fion_keep: the nonlocal force pass adds to fion
  struct KPoint {
    qbg::host::Basis wb;
    double weight = 1.0;
    double *d_g2w = nullptr, *d_gxw = nullptr, *d_twnl = nullptr, *d_dtwnl = nullptr,
           *d_occ = nullptr, *d_occh = nullptr;
    int32_t *d_m_twnl = nullptr, *d_fold = nullptr, *d_unfold = nullptr;
    int64_t* d_dt_off = nullptr;
  };
  std::vector<KPoint> kset;
  int nkp = 1, ngw_max = 0;
  double* d_kacc = nullptr;
  double *o_kin = nullptr, *o_enl = nullptr, *o_nls = nullptr;
  bool fion_keep = false;
  qbg::host::Basis ub;
  int Mh = 0;
  int32_t *d_fold = nullptr, *d_unfold = nullptr, *d_a_m0h = nullptr;
  double *d_cpair = nullptr, *d_hpair = nullptr, *d_occh = nullptr;
  std::vector<void*> allocs;
};

// setup.h -- host-side setup of one system (Gamma or k-points): the work Qbox does
// on the host once per geometry (Basis, BasisMapping, Species SLPP tables,
// NonLocalPotential::update_twnl, the esr lattice sums).  Everything here
// runs at creation / atoms_moved / cell_moved time, never per evaluation.
#pragma once
#include <vector>
#include <cstdint>

namespace qbg {
namespace host {

struct Cell {
  double a[3][3], b[3][3], volume;
  double an[13][3], an2h[13];
  void set(const double* a9);                   // src/UnitCell.cpp:25-124
  void fold_in_ws(double v[3]) const;           // src/UnitCell.cpp:155-186
};

bool factorizable(int n);                       // src/Basis.cpp:105-133

// Basis on one task (src/Basis.cpp:255-690): the real (half-sphere) basis
// at k = 0, or the complex (full-sphere) basis |k + G| < sqrt(2 ecut) of a
// k-point k != 0 (kpoint in units of the reciprocal vectors, :237, 446-481)
struct Basis {
  int size = 0, nrods = 0;
  int idxmin[3], idxmax[3], np[3];
  bool real = true;
  double kpoint[3] = {0.0, 0.0, 0.0};
  std::vector<int> rod_h, rod_k, rod_lmin, rod_size;
  std::vector<int32_t> idx;                     // 3*size
  std::vector<double> g, g2, gi, g2i, gx;       // gx[j*size + i]
  std::vector<double> kpg, kpg2, kpgi, kpgx;    // k + G (equal to g at k = 0)
  // force_complex: the complex branch at k = 0 too (a k-point set's Gamma point)
  void build(const Cell& defcell, const Cell& cell, double ecut, const double* kpoint = nullptr,
             bool force_complex = false);
};

// BasisMapping maps for one task, real basis (src/BasisMapping.cpp:38-228)
struct Maps {
  int nvec = 0;
  std::vector<int32_t> ip, im, zvec_to_val;
  bool build(const Basis& b, int np0, int np1, int np2);
};

// SLPP species with one projector per l (src/Species.cpp:510-1053)
struct Species {
  int zval = 0, lmax = 0, nmesh = 0;
  double deltar = 0, rcps = 0;
  std::vector<double> vloc, proj, dl;           // inputs
  std::vector<double> nlcc;                     // core density on the mesh, or empty
  // semilocal quadrature form (norm-conserving, nquad > 0): semilocal
  // potentials vsl[(lmax+1)*nmesh], local channel llocal; projectors
  // 4 pi j_l(G r_q) r_q Y_lm at r_q = (q+1) rquad/(nquad+1) with weights
  // (v_l(r_q) - v_loc(r_q)) rquad/(nquad+1) (src/NonLocalPotential.cpp:104-229)
  int nquad = 0, llocal = -1;
  double rquad = 0;
  std::vector<double> vsl, rq;
  std::vector<std::vector<double>> wql;         // [l][q]
  bool quad() const { return nquad > 0; }
  // nlm: the projector count (one per (l, m), or per (l, m, q))
  int ndft = 0, nlm = 0;
  std::vector<double> gspl, vlocg, vlocg2;
  std::vector<std::vector<double>> vnlg, vnlg2;
  std::vector<double> nlccg, nlccg2;            // core density transform, spline
  double zcore = 0;
  bool has_nlcc() const { return !nlcc.empty(); }
  bool init();
  void dvlocg(double g, double& v, double& dv) const;
  void dvnlg(int l, double g, double& v, double& dv) const;
  // core density transform and its derivative (src/Species.cpp:1194-1220)
  void drhocoreg(double g, double& rho, double& drho) const;
  // radial function 4 pi j_l(G r_q) r_q of a quadrature projector and its G
  // derivative (src/NonLocalPotential.cpp:317-420, 560-580, 838-855)
  void quad_radial(int l, int q, double g, double& v, double& dv) const;
  // projector table: for projector ipr its l, real-harmonic index l^2 + m,
  // quadrature point (-1 for a channel projector) and D-matrix diagonal
  // (one projector per (l, m) in harmonic order; in the quadrature form
  // ipr = q + nquad (ilm + m) over the nonlocal channels, :196-229)
  void projectors(std::vector<int>& l, std::vector<int>& h, std::vector<int>& q,
                  std::vector<double>& d) const;
  double rhopsg(double g) const;
  double eself() const;
};

// Kleinman-Bylander form of a norm-conserving species given by semilocal
// potentials vps[l] and pseudo-wavefunctions phi[l] (src/Species.cpp:71-507,
// nquad = 0): local potential vps[llocal]; for each nonlocal l the projector
// (vps[l] - vps[llocal]) phi[l] with weight 1 / <phi_l|vps_l - vps_llocal|phi_l>
// (Simpson on the 40-bohr transform mesh).  Output in the one-projector-per-l
// form of Species (lmax_kb = largest nonlocal l, channels without a projector
// zero); false if a nonlocal channel has l > 1 or a weight is undefined.
bool ncpp_to_kb(int lmax, int llocal, int nmesh, double deltar, const double* vps,
                const double* phi, std::vector<double>& vloc, std::vector<double>& proj,
                std::vector<double>& dl, int& lmax_kb);

}  // namespace host
}  // namespace qbg

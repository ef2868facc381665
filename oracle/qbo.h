// qbo.h -- CPU oracle for the Qbox H*psi / energy path.
//
// TEST INFRASTRUCTURE ONLY.  This directory holds a plain C++ (FP64, OpenMP)
// restatement of the reference algorithms (Qbox rel1_78_5, src/*.cpp).  It is
// the checker the HIP library is compared against and the CPU baseline timed
// beside it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it; the product library never links or calls it.
//
// Every function cites the reference file:line it restates.  Pinning of each
// part (golden values, closed forms, invariants, the reference's own spline /
// LDA / UnitCell sources compiled under oracle/_ref) lives in tests/.
#pragma once
#include <complex>
#include <vector>
#include <cstdint>

namespace qbo {

typedef std::complex<double> cplx;

// ---------------------------------------------------------------- UnitCell
// src/UnitCell.cpp:25-124 (set), :155-186 (fold_in_ws)
struct Cell {
  double a[3][3];      // a[i] = lattice vector i
  double b[3][3];      // b[i] = 2 pi (A^-T) row i
  double volume;
  double an[13][3], an2h[13];
  void set(const double* a9);
  void fold_in_ws(double v[3]) const;
};

// ------------------------------------------------------------------- Basis
// src/Basis.cpp:255-690 restated for one task: the real basis at k = 0 and
// the complex (full-sphere) basis at k != 0 (real_ = (k == 0), :237).
struct Basis {
  Cell cell;
  double ecut = 0.0;            // Hartree
  double kpoint[3] = {0.0, 0.0, 0.0};   // in units of the reciprocal vectors b
  bool real = true;
  int idxmin[3], idxmax[3], np[3];
  int size = 0, nrods = 0;
  std::vector<int> rod_h, rod_k, rod_lmin, rod_size, rod_first;
  std::vector<int> idx;         // idx[3*i+j]
  std::vector<double> g, g2, gi, g2i, gx;   // gx[j*size+i]
  std::vector<double> kpg, kpg2, kpgi, kpgx;   // k + G (equal to g at k = 0)
  // defcell: cell that selects the vectors (refcell if given, else cell).
  // kpoint (null: k = 0); force_complex builds the complex branch at k = 0
  // as well (the same formulas with k = 0; used to pin the complex path
  // against the real one)
  void resize(const Cell& c, const Cell* refcell, double ecut_ha,
              const double* kpoint = nullptr, bool force_complex = false);
  void update_g();
};
bool factorizable(int n);       // src/Basis.cpp:105-133

// ------------------------------------------------------------ BasisMapping
// src/BasisMapping.cpp:38-228 (single task; real basis: ip and im, two
// columns per rod but one for rod (0,0); complex basis: ip only, one column
// per rod)
struct Mapping {
  int np0, np1, np2, nvec;
  int ntrans0;                  // src/FourierTransform.cpp:104
  std::vector<int> ip, im, zvec_to_val;
  void build(const Basis& b, int n0, int n1, int n2);
};

// ---------------------------------------------------------- 3D FFT (sparse)
// FourierTransform semantics, src/FourierTransform.cpp:147-827.
struct FFT3 {
  const Basis* basis;
  Mapping map;
  int np012() const { return map.np0 * map.np1 * map.np2; }
  void init(const Basis& b, int n0, int n1, int n2);
  // backward: f(r) = sum_G c(G) e^{+iGr}, unscaled (real basis: G and -G
  // from c(G) and conj(c(G)); complex basis: the G of the basis only)
  void backward(const cplx* c, cplx* f) const;                       // :175
  void backward(const cplx* c1, const cplx* c2, cplx* f) const;      // :187
  // forward: c(G) = 1/N sum_r f(r) e^{-iGr}; overwrites f
  void forward(cplx* f, cplx* c) const;                              // :147
  void forward(cplx* f, cplx* c1, cplx* c2) const;                   // :161
  void bwd_grid(std::vector<cplx>& zvec, cplx* f) const;
  void fwd_grid(cplx* f, std::vector<cplx>& zvec) const;
};

// 1D complex FFT of length n, sign = +1 (synthesis) or -1 (analysis),
// strided in/out; no scaling.  Own mixed-radix implementation.
void fft1d(cplx* x, int n, int stride, int sign);

// ---------------------------------------------------------------- Species
// SLPP species, src/Species.cpp:510-1053, spline.cpp, sinft.cpp.
struct Species {
  int zval; double deltar; double rcps; int lmax;
  std::vector<double> vlocal;                  // input radial local potential
  std::vector<std::vector<double>> proj;       // proj[l][r]
  std::vector<double> dl;                      // D_l (one projector per l)
  std::vector<double> nlcc;                    // core density on the mesh (NLCC), or empty
  // semilocal quadrature form (norm-conserving, nquad > 0): the semilocal
  // potentials vsl[l] on the mesh and the local channel; projectors
  // 4 pi j_l(G r_q) r_q at r_q = (q+1) h, h = rquad/(nquad+1), weights
  // (v_l(r_q) - v_loc(r_q)) h (src/NonLocalPotential.cpp:104-229, 317-960)
  int nquad = 0, llocal = -1;
  double rquad = 0.0;
  std::vector<std::vector<double>> vsl;
  std::vector<double> rq;                      // quadrature points
  std::vector<std::vector<double>> wql;        // [l][q] projector weights
  bool quad() const { return nquad > 0; }
  void quad_radial(int l, int q, double g, double& v, double& dv) const;
  // derived
  int ndft = 0, nop = 0, nlm = 0;
  std::vector<double> rps, gspl, vlocg, vlocg2;
  std::vector<std::vector<double>> vnlg, vnlg2;
  std::vector<double> nlccg, nlccg2;           // core density transform and its spline
  double zcore = 0.0;                          // integral of the core density
  bool has_nlcc() const { return !nlcc.empty(); }
  void initialize();
  void dvlocg(double g, double& v, double& dv) const;     // :1145-1156
  void dvnlg(int iop, double g, double& v, double& dv) const; // :1172-1185
  double rhopsg(double g) const;                            // :1187-1191
  void drhocoreg(double g, double& rho, double& drho) const; // :1194-1220
  double eself() const;                                     // Species.h:133
  double dmatrix(int ipr, int jpr) const;                   // :1450-1464
};
void sinft(int n, double* y);                  // src/sinft.cpp:50-105 (restated)
void cosft1(int n, double* y);                 // src/sinft.cpp:108-174 (restated)
void spline(int n, const double* x, const double* y, double yp1, double ypn,
            int bcnat_left, int bcnat_right, double* y2);   // spline.cpp:44-108
void splint(int n, const double* xa, const double* ya, const double* y2a,
            double x, double* y);                          // spline.cpp:110-136
void splintd(int n, const double* xa, const double* ya, const double* y2a,
             double x, double* y, double* dy);             // spline.cpp:138-168

// Kleinman-Bylander species from semilocal potentials and pseudo-wavefunctions
// (src/Species.cpp:71-330); fills vlocal, proj, dl, lmax of q (deltar set)
void ncpp_species(Species& q, int lmax, int llocal, int nmesh, const double* vps,
                  const double* phi);

// LDA (Perdew-Zunger/Ceperley-Alder), src/LDAFunctional.cpp:92-151
void lda_unpolarized(double rh, double& ee, double& vv);
// fully polarized fit (src/LDAFunctional.cpp:153-213) and the spin
// interpolation (:42-89)
void lda_ferro(double rh, double& ee, double& vv);
void lsda(double rup, double rdn, double& ee, double& vup, double& vdn);

// --------------------------------------------------------- whole evaluation
struct Results {
  double ekin, eps, enl, ehart, ehart_e, ehart_ep, ehart_p, exc, dxc, esr,
         eself, ecoul, etotal;
  std::vector<double> fion;          // 3*nat, species-major
  double sigma[6];
  double sigma_ekin[6], sigma_eps[6], sigma_enl[6], sigma_ehart[6],
         sigma_exc[6], sigma_esr[6];
};

struct System {
  // inputs
  Cell cell; bool has_ref = false; Cell refcell;
  // density grid forced by the caller (0: ChargeDensity's rule); the basis
  // must fit (n >= the density basis's own grid on every axis)
  int grid[3] = {0, 0, 0};
  double ecut = 0.0;                  // Hartree
  std::vector<Species> sp;
  std::vector<std::vector<double>> tau;   // tau[is][3*ia]
  int nst = 0;
  std::vector<double> occ;
  // one k-point (in units of b, weight 1): k != 0 (or force_complex) makes
  // the wavefunction basis complex
  double kpoint[3] = {0.0, 0.0, 0.0};
  bool force_complex = false;
  // several k-points (Wavefunction::add_kpoint, src/Wavefunction.cpp:400-456):
  // each holds its basis, transform and projector tables; the weights sum
  // to 1; the coefficients of k-point ik follow those of ik - 1 (nst x ngw_k
  // each).  Empty: the one k-point of wb and ft.  Between calls the members
  // wb, ft, twnl, dtwnl, lproj hold k-point 0; a loop over the k-points
  // swaps each one in (swap_k) and back.
  struct KState {
    Basis wb;
    FFT3 ft;
    std::vector<std::vector<double>> twnl, dtwnl;
    std::vector<std::vector<int>> lproj;
    double weight = 1.0;
  };
  std::vector<KState> kset;
  std::vector<double> kpts, kwts;      // set before setup(): 3 per k-point, weights
  void swap_k(KState& k);
  size_t k_offset(int ik) const;       // complex elements before k-point ik's coefficients
  // nspin 2: bands [0, nst_up) spin up, [nst_up, nst) spin down
  // (Wavefunction nst_[ispin], src/Wavefunction.cpp:326-336)
  int nspin = 1, nst_up = 0;
  // wavefunction and density bases and grids
  Basis wb, vb;
  FFT3 ft;      // wf basis on the density grid (ChargeDensity ft_[0])
  FFT3 vft;     // density basis on the density grid
  // setup tables (cell_moved / atoms_moved)
  std::vector<std::vector<double>> vps, dvps, rhops;     // [is][ngv]
  std::vector<std::vector<cplx>> sfac;                   // [is][ngv]
  std::vector<cplx> vion_local_g, dvion_local_g, rhopst;
  std::vector<std::vector<double>> twnl, dtwnl;          // [is][ipr*ngw+ig]
  std::vector<std::vector<int>> lproj;
  std::vector<double> fion_esr; double esr = 0.0; double sigma_esr[6];
  // non-linear core correction (any species with a core density):
  // rhocore_sp[is][ig] = rhocoreg(g)/omega, rhocore_g = sum_is sfac
  // rhocoreg / nspin, rhocore_r its real-space grid (each spin's share),
  // vxc_g the transform of vxc_r (spin average) (src/EnergyFunctional.cpp:
  // 155-163, 275-306, 1064-1083, 1224-1228)
  bool core = false;
  std::vector<std::vector<double>> rhocore_sp;           // [is][ngv]
  std::vector<cplx> rhocore_g, vxc_g;
  std::vector<double> rhocore_r;
  double eself_total = 0.0;
  // state
  std::vector<double> rhor;           // density, real space (spin up if nspin 2)
  std::vector<cplx> rhog;             // total density G
  std::vector<double> v_r, vxc_r;     // (spin up if nspin 2)
  std::vector<double> rhor_dn, v_r_dn, vxc_r_dn;   // spin down (nspin 2)
  std::vector<cplx> rhoelg, rhogt;
  double exc = 0, dxc = 0, eps = 0, ehart = 0, ehart_e = 0, ehart_ep = 0,
         ehart_p = 0, total_charge = 0, sigma_exc[6];

  void setup();                         // EnergyFunctional ctor: bases, tables
  void cell_moved();                    // src/EnergyFunctional.cpp:1200-1245
  void atoms_moved();                   // :1047-1197
  void update_twnl();                   // src/NonLocalPotential.cpp:235-1255
  void update_density(const cplx* c);   // src/ChargeDensity.cpp:136-198
  void update_rhor();                   // src/ChargeDensity.cpp:201-221
  // one spin's density in G (the forward transform of update_density) and
  // back: rhor = Re(backward(rhog)) / omega (src/ChargeDensity.cpp:201-221)
  void rhog_of(const double* r, cplx* g);
  void rhor_of(const cplx* g, double* r);
  void core_density();                  // src/EnergyFunctional.cpp:1064-1083
  void update_vhxc(bool compute_stress);// src/EnergyFunctional.cpp:231-394
  double energy(const cplx* c, bool compute_hpsi, cplx* hpsi,
                bool compute_forces, bool compute_stress, Results& r);  // :397
  double nl_energy(const cplx* c, bool compute_hpsi, cplx* hpsi,
                   bool compute_forces, std::vector<double>& fion_enl,
                   bool compute_stress, double* sigma_enl);   // NLP:1258-2069
  double nl_energy_complex(const cplx* c, bool compute_hpsi, cplx* hpsi,
                           bool compute_forces, std::vector<double>& fion_enl,
                           bool compute_stress, double* sigma_enl);
};

// SlaterDet pieces (src/SlaterDet.cpp)
void compute_density(const FFT3& ft, const cplx* c, int ngw, int nst,
                     const double* occ, double weight, double omega, double* rho);
void rs_mul_add(const FFT3& ft, const double* v, const cplx* c, int ngw,
                int nst, cplx* hpsi);
int gram(cplx* c, int ngw, int nst);   // returns 0 or the failing pivot+1
int gram_complex(cplx* c, int ngw, int nst);   // complex basis (SlaterDet.cpp:523-532)
void cleanup(cplx* c, int ngw, int nst);

// SCF iteration pieces (stepper.cpp)
double sd_dot(const cplx* c, const cplx* d, int ngw, int nst);
void prec_ekin(const cplx* c, const double* g2, int ngw, int nst, double* ekin);
double prec_diag(double ekin_n, double g2, double ecutprec);
void residual(const cplx* c, cplx* d, int ngw, int nst, double* a_out);
int polar(double* x, int n, double tol, int maxiter);
int ortho_align(cplx* c, const cplx* cl, int ngw, int nst);
int wf_update(int kind, double param, const double* g2, cplx* c, cplx* d, int ngw, int nst,
              cplx* c_last, cplx* d_last, int* extrapolate);

struct Anderson {
  long m = 0;
  int nmax = 0, n_ = -1, k_ = -1;
  std::vector<double> xs, fs;        // (nmax+1) x m ring buffers
  void restart();
  void update(const double* x, const double* f, double* xbar, double* fbar);
};

struct Mixer {
  double alpha = 0.5;
  bool anderson = true;
  int itscf = 0;
  std::vector<double> wkerker, wls;
  std::vector<cplx> rhog_in;
  Anderson am;
  void init(const System& s, double alpha, double rc_kerker, int ndim, double rc1);
  void restart();
  void update(System& s);            // after update_density
};

}  // namespace qbo

/* qbg.h -- C ABI of the MI355X H*psi / energy library (libqbg.so).
 *
 * The library replaces, for a single-rank (1x1 Context) Qbox run at the
 * Gamma point (and at k-points off Gamma, qbg_ef_create_k and
 * qbg_ef_create_kset), the hot path inside EnergyFunctional::energy: the band FFTs
 * with their G-sphere scatter/gather, the local potential multiply, the
 * density, the Hartree/XC/local terms, the Kleinman-Bylander projections
 * and SlaterDet::gram.  All state stays on one GPU; arrays passed to the
 * functions below are DEVICE pointers unless noted "host".
 *
 * Conventions (all as in the reference):
 *   - complex coefficients are interleaved doubles (re, im), band-major:
 *     c[2*(ig + ldc*n) + {0,1}]  (src/SlaterDet.cpp:84-86, ComplexMatrix);
 *   - real-space index i + np0*(j + np1*k) (src/FourierTransform.h:133);
 *   - backward = unscaled synthesis e^{+iGr}; forward = analysis e^{-iGr}
 *     scaled by 1/(np0*np1*np2); forward may overwrite its input
 *     (src/FourierTransform.h:105-121).
 * Every function returns QBG_OK (0) or a negative QBG_E_* code; the text of
 * the last error of a context is qbg_last_error(ctx).  No function aborts or
 * throws across the ABI.  A handle is used by one host thread at a time;
 * calls are ordered on the context's stream.
 */
#ifndef QBG_H
#define QBG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QBG_OK 0
#define QBG_E_INVAL (-1)   /* bad argument, size or map */
#define QBG_E_NOMEM (-2)   /* device allocation failed */
#define QBG_E_HIP (-3)     /* HIP runtime error */
#define QBG_E_NOT_SPD (-4) /* overlap matrix not positive definite */
#define QBG_E_STATE (-5)   /* call out of order (e.g. energy before vhxc) */

/* precision of the FFT passes and projector GEMMs.  In every mode the
 * Cholesky, the orthogonalization update and the energy, force and stress
 * reductions run in FP64. */
#define QBG_PREC_FP64 0
#define QBG_PREC_FP32 1    /* FP32 FFTs; projector GEMMs on FP32 MFMA */
#define QBG_PREC_BF16X3 2  /* FP32 FFTs; projector GEMMs with FP32 operands split
                              into 3 bfloat16 parts (6 part products, FP32
                              accumulation: FP32-level accuracy) */
#define QBG_PREC_BF16X2 3  /* as BF16X3 with 2 parts (3 part products, about
                              16 significant bits) */

typedef struct qbg_ctx qbg_ctx;
typedef struct qbg_grid qbg_grid;
typedef struct qbg_ef qbg_ef;
typedef struct qbg_graph qbg_graph;
typedef struct qbg_wfstep qbg_wfstep;
typedef struct qbg_mixer qbg_mixer;

/* ---------------------------------------------------------------- context */
/* Replaces the MPI/BLACS context of a 1x1 run (src/Context.cpp:60-75). */
int qbg_init(int device, qbg_ctx** ctx);
/* grids and energy functionals created on ctx must be destroyed first */
int qbg_finalize(qbg_ctx* ctx);
const char* qbg_last_error(const qbg_ctx* ctx);
int qbg_sync(qbg_ctx* ctx);
/* device memory, ordered on the context stream */
int qbg_malloc(qbg_ctx* ctx, size_t bytes, void** dptr);
int qbg_free(qbg_ctx* ctx, void* dptr);
int qbg_memcpy_h2d(qbg_ctx* ctx, void* dst, const void* host_src, size_t bytes);
int qbg_memcpy_d2h(qbg_ctx* ctx, void* host_dst, const void* src, size_t bytes);
int qbg_memset(qbg_ctx* ctx, void* dst, int value, size_t bytes);
/* elapsed milliseconds between the last qbg_timer_start/stop pair on the
 * context stream (HIP events) */
int qbg_timer_start(qbg_ctx* ctx);
int qbg_timer_stop(qbg_ctx* ctx, float* ms);

/* Per-kernel timing with HIP events on the context stream (the timers of
 * src/FourierTransform.h:139-141 and src/EnergyFunctional.cpp:265-791).
 * While enabled, each launch of a kernel kind below is bracketed by events;
 * qbg_prof_read returns the summed milliseconds and the launch count. */
#define QBG_K_ZBWD 0      /* z pass backward (+ G-sphere pack) */
#define QBG_K_XBWD 1      /* x pass backward (+ column scatter) */
#define QBG_K_YVMUL 2     /* y pass: bwd * v(r) * fwd */
#define QBG_K_YDENS 3     /* y pass: bwd + density accumulation */
#define QBG_K_XFWD 4      /* x pass forward (+ column gather) */
#define QBG_K_ZFWD 5      /* z pass forward (+ unpack / accumulate) */
#define QBG_K_YOTHER 6    /* y pass, full-grid transforms */
#define QBG_K_NL 7        /* projector generation, D contraction, force sums */
#define QBG_K_GRAM 8      /* overlap, Cholesky, triangular update */
#define QBG_K_OTHER 9     /* reductions, pointwise, structure factor */
#define QBG_K_VGRID 10    /* single transforms on the density basis */
#define QBG_K_GEMM 11     /* one projector GEMM (anl^T c, anl fout or c W):
                             2 * nproj * 2ngw * nst flops each */
#define QBG_K_WFSTEP 12   /* wavefunction stepper update incl. orthogonalization */
#define QBG_K_MIX 13      /* charge mixing incl. rhor from the mixed rhog */
#define QBG_K_COUNT 14
/* A plain device-to-device streaming copy of `bytes` (a multiple of 16):
 * the reference kernel of known traffic for calibrating HBM byte
 * counters.  No counterpart in the reference. */
int qbg_stream_copy(qbg_ctx* ctx, void* dst, const void* src, size_t bytes);
int qbg_prof_enable(qbg_ctx* ctx, int on);
int qbg_prof_reset(qbg_ctx* ctx);
int qbg_prof_read(qbg_ctx* ctx, int kind, double* total_ms, int* count);

/* ------------------------------------------- FourierTransform + mapping */
/* FourierTransform(const Basis&, np0, np1, np2) with BasisMapping's maps
 * (src/FourierTransform.cpp:97-114, src/BasisMapping.cpp:38-228).  The
 * caller passes Qbox's own host maps ip[ng], im[ng] (offsets into the
 * nvec*np2 column array) and zvec_to_val[nvec] (x + np0*y of each column),
 * so index maps are bit-exact by construction; maps that do not have
 * Qbox's real-basis structure (rod (0,0) on column 0, rod r on columns
 * 2r-1 and 2r, G vectors of a rod contiguous) are rejected. */
int qbg_grid_create(qbg_ctx* ctx, int np0, int np1, int np2, int ng, int nvec,
                    const int32_t* ip, const int32_t* im,
                    const int32_t* zvec_to_val, qbg_grid** out);
int qbg_grid_destroy(qbg_grid* g);
/* FourierTransform's size accessors np0(), np1(), np2() and the basis size
 * (src/FourierTransform.h:123-137), plus the kernels that serve the grid:
 * out[0..2] np0 np1 np2, [3] ng, [4] nvec, [5] ntrans0, [6] 1 if the
 * compiled four-step / prime-factor kernels run every axis (0: the generic
 * Stockham kernels), [7..12] the split n1 x n2 of axes 0, 1, 2 (0 if
 * generic).  out has room for 13 ints. */
int qbg_grid_info(const qbg_grid* g, int* out);
/* Launch tuning of the batched transforms, in place of FourierTransform's
 * FFTW_MEASURE planning (src/FourierTransform.cpp:45-54, 830-1062).  At the
 * first batched call of a grid in a precision (compute_density, rs_mul_add,
 * the energy functional's passes), outside graph capture, the library times
 * every compiled line split of np0 (x pass) and np1 (y pass) and the y
 * pass's band pairs per block (8, 16, 32) on the grid's own workspace, keeps
 * the fastest, and caches the choice per device, grid shape and precision
 * for the process.  qbg_grid_tuning reports it: out[0] 1 once tuned for prec
 * (FP64, or any of the FP32-FFT modes), out[1..2] the x split n1 x n2,
 * out[3..4] the y split, out[5] the pairs per block (6 ints).
 * qbg_grid_plan fixes the choice instead (FFTW_ESTIMATE-like): x_n1, y_n1
 * name a compiled split by its n1 (0: the default split), yppb the pairs per
 * block (0: the default, 16); QBG_E_INVAL if no such split is compiled. */
int qbg_grid_tuning(const qbg_grid* g, int prec, int* out);
int qbg_grid_plan(qbg_grid* g, int prec, int x_n1, int y_n1, int yppb);
/* the n1 of every compiled split n1 x n2 of line length n, default first,
 * at most cap of them written to n1 (host); returns their number (0 if n
 * has no compiled kernels: the generic ones serve it) */
int qbg_fft_splits(int n, int* n1, int cap);

/* FourierTransform::backward(c, f) and backward(c1, c2, f)
 * (src/FourierTransform.cpp:175-199): f has np012 complex entries. */
int qbg_fft_backward(qbg_grid* g, const double* c, double* f, int prec);
int qbg_fft_backward_pair(qbg_grid* g, const double* c1, const double* c2,
                          double* f, int prec);
/* FourierTransform::forward(f, c) and forward(f, c1, c2)
 * (src/FourierTransform.cpp:147-173); f is overwritten. */
int qbg_fft_forward(qbg_grid* g, double* f, double* c, int prec);
int qbg_fft_forward_pair(qbg_grid* g, double* f, double* c1, double* c2,
                         int prec);

/* SlaterDet::compute_density(ft, weight, rho) (src/SlaterDet.cpp:225-298):
 * rho[np012] += sum_n weight*occ[n]/omega |psi_n(r)|^2 over nst bands of c
 * (leading dimension ldc complex); occ is a host array. */
int qbg_compute_density(qbg_grid* g, const double* c, int ldc, int nst,
                        const double* occ, double weight, double omega,
                        double* rho, int prec);
/* SlaterDet::rs_mul_add(ft, v, sdp) (src/SlaterDet.cpp:410-477):
 * hpsi[:, n] += FFT_fwd( v(r) * FFT_bwd(c[:, n]) ) for nst bands. */
int qbg_rs_mul_add(qbg_grid* g, const double* v, const double* c, int ldc,
                   int nst, double* hpsi, int ldh, int prec);

/* ------------------------------------------------ SlaterDet::gram */
/* SlaterDet::gram() for the real basis (src/SlaterDet.cpp:480-534):
 * cleanup (Im c(G=0) = 0), S = 2 Re(C^H C) - c0 c0^T, S = L L^T (lower
 * Cholesky), C <- C L^-T, in place on the device array c (ngw x nst,
 * leading dimension ldc).  QBG_E_NOT_SPD if S is not positive definite, in
 * which case c is left as cleaned up but otherwise unchanged. */
int qbg_gram(qbg_ctx* ctx, double* c, int ldc, int ngw, int nst, int prec);
/* SlaterDet::gram() for the complex basis of a k-point (src/SlaterDet.cpp:
 * 523-532): S = C^H C (herk), S = L L^H (lower Cholesky), C <- C L^-H, in
 * place on the device array c (ngw x nst complex, leading dimension ldc),
 * FP64, nst <= 2048.  QBG_E_NOT_SPD if S is not positive definite, in which
 * case c is unchanged.  Synchronizes the context stream. */
int qbg_gram_complex(qbg_ctx* ctx, double* c, int ldc, int ngw, int nst);

/* qbg_gram without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as qbg_gram (equal to rounding: the overlap's
 * split-K partition may differ); the overlap, its inverse factor and the
 * split-K partials live in one fixed device array per GPU of
 * qbg_gram_noalloc_bytes() bytes, held until the process ends, and C is
 * updated in place.  nst above QBG_GRAM_NOALLOC_MAX_NST is refused with
 * QBG_E_INVAL.  Every call on a GPU shares that GPU's array: calls that
 * overlap in time on one GPU (other streams or threads) race, so the caller
 * serializes them. */
#define QBG_GRAM_NOALLOC_MAX_NST 1024
#define QBG_GRAM_NOALLOC_SPLITS 6
int qbg_gram_noalloc(qbg_ctx* ctx, double* c, int ldc, int ngw, int nst, int prec);
size_t qbg_gram_noalloc_bytes(void);

/* ---------------------------------------------------- EnergyFunctional */
/* A species as Qbox holds it, on the radial mesh i*deltar (host arrays);
 * rcps is the pseudocharge radius (1.5 in Qbox, src/AtomSet.cpp:43-44).
 *   type QBG_SPECIES_SLPP: norm_conserving_semilocal_pseudopotential
 *     (src/SpeciesHandler.cpp:200-241, Species::initialize_slpp) with one
 *     projector per l = 0..lmax (lmax <= 3): local potential vloc[nmesh],
 *     projectors proj[(lmax+1)*nmesh], D_l in dl[lmax+1];
 *   type QBG_SPECIES_NCPP: norm_conserving_pseudopotential in the
 *     Kleinman-Bylander form (nquad = 0; Species::initialize_ncpp,
 *     src/Species.cpp:71-507): semilocal potentials vps[(lmax+1)*nmesh]
 *     and pseudo-wavefunctions phi[(lmax+1)*nmesh] (the row of the local
 *     channel llocal may hold anything); lmax <= 3, every nonlocal channel
 *     l != llocal must have l <= 2.  With nquad > 0 the semilocal
 *     quadrature form instead (src/NonLocalPotential.cpp:104-229, 317-960):
 *     phi is not used; each nonlocal channel l <= 2 gets (2l+1) nquad
 *     projectors 4 pi j_l(G r_q) r_q Y_lm at r_q = (q+1) rquad/(nquad+1),
 *     weighted by (v_l(r_q) - v_llocal(r_q)) rquad/(nquad+1).
 *   nlcc (either type, may be NULL): the core density <core_density> on the
 *     same mesh (nmesh values) for the non-linear core correction
 *     (Species::initialize_nlcc, src/Species.cpp:1056-1100): the
 *     exchange-correlation functional then sees the valence plus the core
 *     density, with the matching dxc, force and stress terms
 *     (src/EnergyFunctional.cpp:275-306, 650-659, 855-856, 1064-1083).
 * Fields after dl were added later; a zeroed tail selects SLPP without a
 * core density. */
#define QBG_SPECIES_SLPP 0
#define QBG_SPECIES_NCPP 1
typedef struct {
  int zval;
  int lmax;
  int nmesh;
  double deltar;
  double rcps;
  const double* vloc;
  const double* proj;
  const double* dl;
  int type;
  int llocal;
  const double* vps;
  const double* phi;
  const double* nlcc;
  int nquad;
  double rquad;
} qbg_species_desc;

/* EnergyFunctional(Sample&, ChargeDensity&) for one Gamma-point system
 * (src/EnergyFunctional.cpp:46-197, src/ChargeDensity.cpp:34-108): builds
 * the wavefunction basis (ecut, Hartree) and the density basis (4 ecut) on
 * the density grid, the species tables and projectors (cell_moved), the
 * structure factor on the device and the esr sums (atoms_moved).  cell9 is
 * a0,a1,a2 (bohr); refcell9 (may be NULL) is Wavefunction's reference cell,
 * which then fixes the basis membership and the grid sizes
 * (src/Basis.cpp:314-325); tau[3*nat] is species-major; occ[nst].  Host
 * arrays. */
int qbg_ef_create(qbg_ctx* ctx, const double* cell9, const double* refcell9, double ecut, int nsp,
                  const qbg_species_desc* species, const int* na,
                  const double* tau, int nst, const double* occ, qbg_ef** out);
/* The same for one k-point off Gamma, weight 1 (the Wavefunction k-point
 * of src/Wavefunction.cpp:400-456 with src/Basis.cpp:237, 446-481):
 * kpoint3 in units of the reciprocal vectors b0,b1,b2 (0,0,0 gives the
 * Gamma functional of qbg_ef_create).  The wavefunction basis is then the
 * full sphere |k+G| < sqrt(2 ecut) of the reference's complex Basis, in its
 * order (qbg_ef_basis which 0 and 2), and the coefficients are complex on
 * it; update_density, update_vhxc and the evaluation (energies, H*psi,
 * forces) follow src/EnergyFunctional.cpp:397-1044 and
 * src/NonLocalPotential.cpp:1429-2066 for the complex basis (the stress
 * included), in every precision, with nspin 1 or 2.  Not built at k != 0
 * (QBG_E_INVAL): the wavefunction-grid entry points and the SCF steppers;
 * the NonLocalPotential entry points (qbg_nl_*) take one k-point, not a
 * set. */
int qbg_ef_create_k(qbg_ctx* ctx, const double* cell9, const double* refcell9, double ecut,
                    const double* kpoint3, int nsp, const qbg_species_desc* species,
                    const int* na, const double* tau, int nst, const double* occ, qbg_ef** out);
/* A Wavefunction with nkp k-points (src/Wavefunction.cpp:400-456; nkp = 1
 * is qbg_ef_create_k): kpoints3n[3*nkp] in units of b0,b1,b2, weights[nkp]
 * (summing to 1, Wavefunction::weight).  Every k-point takes the complex
 * basis (the Gamma point of a set too).  k-point ik's bands are the columns
 * [ik*nst, (ik+1)*nst) of c and H*psi, with one leading dimension of at
 * least the largest basis (qbg_ef_kset_basis); the density sums the
 * k-points with their weights, and so do ekin, enl, the forces and the
 * stress (src/EnergyFunctional.cpp:446-543, 703-716; ChargeDensity.cpp:
 * 136-160).  Sets run in every precision, with nspin 1 or 2.  Rows of a k-point's H*psi columns past its basis size are
 * not written. */
int qbg_ef_create_kset(qbg_ctx* ctx, const double* cell9, const double* refcell9, double ecut,
                       int nkp, const double* kpoints3n, const double* weights, int nsp,
                       const qbg_species_desc* species, const int* na, const double* tau,
                       int nst, const double* occ, qbg_ef** out);
/* k-point ik of a functional's k-points (0 for one k-point): the basis size
 * ngw and, for non-null pointers, idx[3*ngw], |k+G|^2[ngw], k+G[3*ngw]. */
int qbg_ef_kset_basis(const qbg_ef* ef, int ik, int* ngw, int32_t* idx, double* kpg2,
                      double* kpgx);
int qbg_ef_destroy(qbg_ef* ef);
/* out[10]: ngw, ngv, np0, np1, np2, nvec_w, nvec_v, nat, nproj, nst */
int qbg_ef_sizes(const qbg_ef* ef, int* out);
/* host copies of the wavefunction (which=0) or density (which=1) basis:
 * idx[3*ng], g2[ng], gx[3*ng] (x block, y block, z block), maps ip[ng],
 * im[ng], zvec_to_val[nvec]; any pointer may be null.  which=2: the
 * wavefunction basis with g2 = |k+G|^2 and gx = k+G (Basis::kpg2, kpgx;
 * equal to which=0 at Gamma).  A k != 0 basis has no maps (they must be
 * null). */
int qbg_ef_basis(const qbg_ef* ef, int which, int32_t* idx, double* g2,
                 double* gx, int32_t* ip, int32_t* im, int32_t* zvec_to_val);
/* EnergyFunctional::atoms_moved() with new positions (host tau[3*nat]) */
int qbg_ef_set_positions(qbg_ef* ef, const double* tau);
/* Wavefunction::resize(cell, refcell, ecut) followed by
 * EnergyFunctional::cell_moved() (src/EnergyFunctional.cpp:1200-1245): the
 * basis vectors of both bases for the new cell cell9 (host, a0 a1 a2), with
 * the membership and grid fixed by the reference cell given at creation; the
 * local tables vps, dvps, rhops; NonLocalPotential::update_twnl; then
 * atoms_moved at the unchanged Cartesian positions (move them afterwards
 * with qbg_ef_set_positions if they scale with the cell).  The density must
 * be updated again before update_vhxc.  Without a reference cell the basis
 * follows the cell, as in the reference: QBG_E_INVAL (nothing changed) if
 * the new cell would change the basis membership or the grid; create the
 * functional again then. */
int qbg_ef_cell_moved(qbg_ef* ef, const double* cell9);
/* ChargeDensity::update_density() from device coefficients c */
int qbg_ef_update_density(qbg_ef* ef, const double* c, int ldc, int prec);
/* EnergyFunctional::update_vhxc(compute_stress) */
int qbg_ef_update_vhxc(qbg_ef* ef, int compute_stress);
/* EnergyFunctional::energy(compute_hpsi, dwf, compute_forces, fion,
 * compute_stress, sigma): hpsi (device, overwritten with H*psi when
 * compute_hpsi); energies (host, 14 doubles: ekin eps enl ehart ehart_e
 * ehart_ep ehart_p exc dxc esr eself ecoul etotal total_charge); fion (host,
 * 3*nat); sigma (host, 7x6: total, ekin, eps, enl, ehart, exc, esr; columns
 * xx yy zz xy yz xz).  compute_stress needs update_vhxc(compute_stress=1). */
int qbg_ef_energy(qbg_ef* ef, const double* c, int ldc, int compute_hpsi,
                  double* hpsi, int ldh, int compute_forces, int compute_stress,
                  int prec, double* energies, double* fion, double* sigma);
/* Spin polarization: Wavefunction::set_nspin (src/Wavefunction.cpp:365-390)
 * with the coefficient columns [0, nst_up) spin up and [nst_up, nst) spin
 * down (Qbox keeps one SlaterDet per spin; here they are column ranges of one
 * array), occupations as given at creation (at most 1 per band with two
 * spins).  update_density forms one density per spin (ChargeDensity
 * rhor[ispin]), update_vhxc the spin-polarized LDA and v_r[ispin]
 * (src/EnergyFunctional.cpp:248-389, src/LDAFunctional.cpp:42-89), energy
 * applies each band's spin potential.  nspin 1 restores the unpolarized
 * functional.  With nspin 2 each spin takes its own stepper
 * (qbg_wfstep_create_bands); the mixer mixes both spins' densities.
 * qbg_ef_get: what 0-2 are the spin-up arrays, 8 rhor, 9 v_r, 10 vxc_r of
 * spin down. */
int qbg_ef_set_nspin(qbg_ef* ef, int nspin, int nst_up);

/* NonLocalPotential(atoms, sd) of the SlaterDet of spin ispin (bands
 * [0, nst_up) or [nst_up, nst) of the functional; all bands with nspin 1),
 * as EnergyFunctional constructs it (src/EnergyFunctional.cpp:113-118,
 * src/NonLocalPotential.h:61-68).  The handle uses the functional's
 * positions (atoms_moved), projector tables and workspace; it is not
 * reentrant with the functional's own calls (one host thread, one stream).
 * qbg_nl_update_twnl: NonLocalPotential::update_twnl()
 * (src/NonLocalPotential.cpp:235-1255) for the functional's current basis
 * (qbg_ef_cell_moved already does it).
 * qbg_nl_energy: NonLocalPotential::energy(compute_hpsi, dsd,
 * compute_forces, fion, compute_stress, sigma_enl)
 * (src/NonLocalPotential.cpp:1258-2069): c (device, ld ldc complex) holds
 * the spin's bands; H*psi_NL is ADDED into dsd (device, ld ldd); fion (host,
 * 3 nat, species-major) and sigma_enl (host, 6: xx yy zz xy yz xz) are
 * zeroed, then filled when asked for; *enl receives the nonlocal energy.
 * Synchronous (returns after the results are on the host). */
typedef struct qbg_nl qbg_nl;
int qbg_nl_create(qbg_ef* ef, int ispin, qbg_nl** out);
int qbg_nl_destroy(qbg_nl* nl);
int qbg_nl_update_twnl(qbg_nl* nl);
int qbg_nl_energy(qbg_nl* nl, const double* c, int ldc, int compute_hpsi, double* dsd, int ldd,
                  int compute_forces, double* fion, int compute_stress, double* sigma_enl,
                  int prec, double* enl);

/* The wavefunction basis on its own grid Basis::np (the FourierTransform of
 * SlaterDet::print, src/SlaterDet.cpp:1385, and of the sample reader,
 * src/WavefunctionHandler.cpp:411): the grid of the <grid_function> elements
 * of a sample file.  dims (host, 3, may be null) receives nx, ny, nz; the
 * grid belongs to ef (do not destroy it) and takes the qbg_fft_* calls. */
int qbg_ef_wf_grid(qbg_ef* ef, int* dims, qbg_grid** out);
/* The grid the band transforms of update_density and H*psi run on: the
 * FourierTransform of the wavefunction basis on the density grid
 * (src/ChargeDensity.cpp:44-49, 98; src/EnergyFunctional.cpp:831).  It
 * belongs to ef (do not destroy it); qbg_grid_tuning reports its launch
 * choice, qbg_grid_plan fixes one. */
int qbg_ef_band_grid(qbg_ef* ef, qbg_grid** out);
/* The GEMM half of the launch tuner (the grid's tuner covers the FFT
 * passes): at the first 2-part BF16 evaluation, outside graph capture, the
 * projection product anl^T c is timed with the 256-edge tiles at the default
 * reduction split, at twice and half of it, and with the 128-edge tiles, on
 * the functional's own operands; the fastest plan is kept (cached per device
 * and padded shape).  out (host, 3 ints): [0] 1 once tuned, [1] tile edge,
 * [2] reduction splits (0: the default rule).  No counterpart in the
 * reference (its dgemm calls take no plan). */
int qbg_ef_gemm_plan(const qbg_ef* ef, int* out);

/* One SCF-iteration evaluation without host synchronization, so that it can
 * be captured in a graph: update_density -> update_vhxc -> energy
 * (src/BOSampleStepper.cpp:1207-1217) with the flags below; the results are
 * staged on the device and returned by qbg_ef_results, which waits for the
 * stream.  Workspaces are allocated by the first (uncaptured) call. */
#define QBG_EF_HPSI 1
#define QBG_EF_FORCES 2
#define QBG_EF_STRESS 4
int qbg_ef_evaluate_async(qbg_ef* ef, const double* c, int ldc, double* hpsi, int ldh,
                          int flags, int prec);
/* the energy step alone (EnergyFunctional::energy after update_density and
 * update_vhxc), without host synchronization; results as below.  With
 * QBG_EF_SAME_C in flags the caller guarantees that c is the array the last
 * qbg_ef_update_density read and that it has not changed since: the H*psi
 * pass then starts from the backward transforms update_density kept
 * (resident in device memory when they fit in half of the free memory)
 * instead of repeating them; the result is the same.  Without the flag, or
 * when nothing was kept, the transforms are recomputed.  qbg_ef_evaluate_async
 * always reuses them (one call, same c). */
#define QBG_EF_SAME_C 8
int qbg_ef_energy_async(qbg_ef* ef, const double* c, int ldc, double* hpsi, int ldh,
                        int flags, int prec);
/* energies (14), fion (3*nat) and sigma (7x6) of the last launched
 * evaluation, as qbg_ef_energy returns them */
int qbg_ef_results(qbg_ef* ef, double* energies, double* fion, double* sigma);

/* ---------------------------------------------------------------- graphs */
/* Capture every call made on ctx between qbg_graph_begin and qbg_graph_end
 * into one HIP graph (stream capture); qbg_graph_launch replays it on the
 * context stream.  Calls that synchronize, allocate or copy from the host
 * fail with QBG_E_STATE while capturing.  No counterpart in the reference:
 * this is the per-SCF-iteration launch graph of the evaluation path. */
/* Order two contexts' streams (same device): work enqueued on `after` from
 * now on waits for the work enqueued on `before` so far.  While `before`
 * captures a graph, `after`'s stream joins the capture, and the calls made
 * on `after` become a parallel branch of the graph; join the branch back
 * with qbg_stream_join(after, before) before qbg_graph_end(before).  A graph
 * launched on one context and holding other contexts' branches is complete
 * when the launching context's stream is (qbg_sync it before reading the
 * branches' results).  No counterpart in the reference (one sequence of work
 * per MPI task); it serves several independent systems in one graph. */
int qbg_stream_join(qbg_ctx* before, qbg_ctx* after);
int qbg_graph_begin(qbg_ctx* ctx);
int qbg_graph_end(qbg_ctx* ctx, qbg_graph** out);
int qbg_graph_launch(qbg_graph* graph);
int qbg_graph_destroy(qbg_graph* graph);

/* ------------------------------------------------ SCF iteration updates */
/* The wavefunction steppers of BOSampleStepper's SCF loop
 * (src/BOSampleStepper.cpp:237-256, 936-970), all on the device for the
 * bands of an energy functional:
 *   QBG_WF_SD   SDWavefunctionStepper (src/SDWavefunctionStepper.cpp:36-51):
 *               c -= param * dc, then gram; param = dt2bye;
 *   QBG_WF_PSD  PSDWavefunctionStepper (src/PSDWavefunctionStepper.cpp:41-119):
 *               residual dc - c (c^T dc), preconditioned step, gram;
 *   QBG_WF_PSDA PSDAWavefunctionStepper (src/PSDAWavefunctionStepper.cpp:45-247):
 *               preconditioned residual with Anderson extrapolation, then
 *               SlaterDet::ortho_align against the previous bands
 *               (src/SlaterDet.cpp:768-876, polar factor src/Matrix.cpp:2446-2517).
 * For PSD and PSDA, param is ecutprec in Hartree (0: the adaptive
 * Preconditioner of src/Preconditioner.cpp:48-143; the reference's
 * "ecutprec" command stores 0.5 * Ry value). */
#define QBG_WF_SD 0
#define QBG_WF_PSD 1
#define QBG_WF_PSDA 2
int qbg_wfstep_create(qbg_ef* ef, int kind, double param, qbg_wfstep** out);
/* A stepper over a band set of nst bands (1 <= nst <= the functional's
 * bands): one per spin with nspin 2, as Qbox steps each spin's SlaterDet
 * (src/BOSampleStepper.cpp:939-944, wf_stepper->update over the spins);
 * qbg_wfstep_update then takes that spin's columns. */
int qbg_wfstep_create_bands(qbg_ef* ef, int kind, double param, int nst, qbg_wfstep** out);
int qbg_wfstep_destroy(qbg_wfstep* st);
/* WavefunctionStepper::preprocess (PSDA: the next update does not extrapolate) */
int qbg_wfstep_preprocess(qbg_wfstep* st);
/* WavefunctionStepper::update(dwf): c (ldc) holds the bands, dc (ldd) H*c on
 * entry and the stepper's direction on exit.  Also forms
 * eigenvalue_sum = SlaterDet::dot(c, dc) of the entry values
 * (src/BOSampleStepper.cpp:945-951).  Enqueued without host synchronization
 * (capturable); results come from qbg_wfstep_results. */
int qbg_wfstep_update(qbg_wfstep* st, double* c, int ldc, double* dc, int ldd);
/* waits for the stream; out[4]: eigenvalue_sum, PSDA theta, polar iterations
 * of ortho_align, Cholesky pivot of the failing orthogonalization (0 = ok).
 * Returns QBG_E_NOT_SPD when the orthogonalization failed (the bands are
 * then left as the step produced them, not orthonormalized). */
int qbg_wfstep_results(qbg_wfstep* st, double* out);

/* Charge mixing of BOSampleStepper (src/BOSampleStepper.cpp:319-406, 810-879)
 * on the density of an energy functional: Kerker weights g2/(g2 + q0^2),
 * q0 = 2 pi / rc_kerker (none if rc_kerker <= 0), least-squares row weights
 * sqrt(g2/(g2 + q1^2)), q1 = 2 pi / rc1, and the AndersonMixer
 * (src/AndersonMixer.cpp) over ndim previous vectors (ndim = 0: simple
 * mixing; ndim <= 8).  qbg_mixer_update follows qbg_ef_update_density: the
 * first call after create/restart memorizes rhog, later calls mix and
 * recompute rhor from the mixed rhog (ChargeDensity::update_rhor,
 * src/ChargeDensity.cpp:201-221).  nspin 2 (the functional's at creation):
 * the vector mixed holds rhog of spin up then spin down, each the transform
 * of that spin's rhor (src/BOSampleStepper.cpp:318-327); after mixing each
 * rhor is the inverse transform of its mixed rhog and the total rhog their
 * sum.  An update after the functional's nspin changed is refused.
 * Capturable; the iteration count is kept on the device. */
int qbg_mixer_create(qbg_ef* ef, double alpha, double rc_kerker, int ndim, double rc1,
                     qbg_mixer** out);
int qbg_mixer_destroy(qbg_mixer* mx);
int qbg_mixer_restart(qbg_mixer* mx);
int qbg_mixer_update(qbg_mixer* mx);

/* DoubleMatrix::polar (src/Matrix.cpp:2446-2517): x (device, n x n, column
 * major) is replaced by its orthogonal polar factor; iterations stop when
 * ||I - x^T x||_F <= tol or after maxiter (<= 64).  Synchronizes; *iters
 * (host, may be null) receives the iteration count.  QBG_E_NOT_SPD if a
 * Higham step met a singular matrix. */
int qbg_polar(qbg_ctx* ctx, double* x, int n, double tol, int maxiter, int* iters);

/* host copy of a device array: 0 rhor, 1 v_r, 2 vxc_r, 3 rhog,
 * 4 vion_local_g, 5 rhopst, 6 sfac, 7 fion_esr; with nspin 2, 0-2 are the
 * spin-up arrays and 8 rhor, 9 v_r, 10 vxc_r the spin-down ones; with core
 * densities 11 the core density on the grid (each spin's share, ChargeDensity
 * rhocore_r) and 12 vxc_g (after update_vhxc) */
int qbg_ef_get(qbg_ef* ef, int what, double* host_out);
/* replace v_r (host array of np012 doubles) */
int qbg_ef_set_vr(qbg_ef* ef, const double* host_v);

#ifdef __cplusplus
}
#endif
#endif /* QBG_H */

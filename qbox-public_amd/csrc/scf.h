// scf.h -- launchers of the SCF-iteration kernels (scf.hip): wavefunction
// steppers with their preconditioner, the orthogonalizations that follow
// them, and charge mixing.  Every launcher only enqueues work on the given
// stream (no allocation, copy or synchronization), so a whole iteration can
// be captured into a HIP graph.  Flags and counters live in device memory.
#pragma once
#include <hip/hip_runtime.h>

namespace qbg {

// per band n of c (real view, ld ldc2 doubles) and d:
//   ekin[n] (Preconditioner::update) and dotp[n] = 2 c_n.d_n - Re c_n(0) Re d_n(0);
// part holds band_partials(nst) doubles
int band_partials(int nst);
void launch_band_sums(const double* c, long ldc2, const double* d, long ldd2, const double* g2,
                      int ngw, int nst, double* part, double* ekin, double* dotp,
                      hipStream_t s);
// out[0] = sum_n dotp[n] in band order
void launch_sum(const double* v, int n, double* out, hipStream_t s);
// c -= diag(n, g) d  (PSDWavefunctionStepper::update)
void launch_psd_update(double* c, long ldc2, const double* d, long ldd2, const double* g2,
                       const double* ekin, int ngw, int nst, double ecutprec, hipStream_t s);
// c -= alpha d  (SDWavefunctionStepper::update)
void launch_sd_update(double* c, long ldc2, const double* d, long ldd2, int ngw, int nst,
                      double alpha, hipStream_t s);
// PSDA: d = -diag d; partial sums of a = sum w f (f - f_last), b = sum w (f - f_last)^2
// (w = 2, 1 for G = 0) over blocks; then theta; then the update of c, c_last, d_last
int psda_partials(int ngw, int nst);
void launch_psda_prec(double* d, long ldd2, const double* dl, const double* g2,
                      const double* ekin, int ngw, int nst, double ecutprec, double* partial,
                      hipStream_t s);
void launch_psda_theta(const double* partial, int nblk, int* extrapolate, double* theta,
                       hipStream_t s);
void launch_psda_update(double* c, long ldc2, const double* d, long ldd2, double* cl, double* dl,
                        const double* theta, int ngw, int nst, hipStream_t s);
// *p = v
void launch_set_int(int* p, int v, hipStream_t s);
// dst (columns of `rows` doubles, ld lddst) = src (ld ldsrc) if *info == 0;
// rows and both leading dimensions even
void launch_copy_cols_if(double* dst, long lddst, const double* src, long ldsrc, long rows,
                         int cols, const int* info, hipStream_t s);

// DoubleMatrix::polar pieces on an n x n column-major matrix.
// flags: [0] done, [1] skip Bjorck-Bowie step, [2] skip Higham step,
//        [3] iterations, [4] singular, [5] Cholesky status of a Higham step
void launch_polar_init(int* flags, hipStream_t s);
// q = q_in + I (q_in = -x^T x), its Frobenius norm; sets flags for this step
void launch_polar_norm(double* q, int n, double tol, int* flags, double* qnrm, hipStream_t s);
// t = I + a q (elementwise on the n x n matrix), skipped when *skip
void launch_ident_axpy(double* t, const double* q, int n, double a, const int* skip,
                       hipStream_t s);
// dst = src (n x n), skipped when *skip
void launch_copy_if_not(double* dst, const double* src, long count, const int* skip,
                        hipStream_t s);
// Higham step helpers (x^-T = x (x^T x)^-1 through the blocked Cholesky):
// dst = transpose of src (n x n), x = (x + z) / 2, flags[4] = 1 if *info > 0;
// the first two skipped when *skip
constexpr int POLAR_NMAX = 4096;
void launch_transpose_if(double* dst, const double* src, int n, const int* skip, hipStream_t s);
void launch_higham_update(double* x, const double* z, int n, const int* skip, hipStream_t s);
void launch_polar_singular(const int* info, int* flags, hipStream_t s);

// out[4] = eigenvalue sum, theta, polar iterations, orthogonalization status
void launch_wf_results(const double* scal, const int* flags, const int* info, double* out,
                       hipStream_t s);

// charge mixing.  st: [0] itscf, [1] n (vectors beyond the current one),
// [2] k (ring slot of the current vector), [3] mixed this call
constexpr int MIX_NMAX = 8;
void launch_mix_restart(int* st, hipStream_t s);
void launch_mix_begin(int* st, int nmax, int anderson, hipStream_t s);
void launch_mix_stage(double* rhog, double* rhog_in, const double* wls, const double* wk,
                      double* xs, double* fs, const int* st, int ng, int nw, double alpha,
                      int anderson, hipStream_t s);
int mix_partials();
void launch_mix_dots(const double* fs, const int* st, int nmax, long m, double* partial,
                     hipStream_t s);
void launch_mix_solve(const double* partial, const int* st, int anderson, double* theta,
                      hipStream_t s);
void launch_mix_apply(double* rhog, double* rhog_in, const double* wls, const double* wk,
                      const double* xs, const double* fs, const int* st, const double* theta,
                      int nmax, int ng, int nw, double alpha, int anderson, hipStream_t s);
void launch_mix_select(double* rhor, const double* rhor_new, long n, const int* st,
                       hipStream_t s);

}  // namespace qbg

// energy.h -- launchers of energy.hip and gemm.hip (stream-ordered).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace qbg {

void reduce_partials(const double* partial, int nblocks, int nv, double* out, hipStream_t s);

// returns the number of partial blocks used (partial needs nb*8 doubles)
int launch_kinetic(const double* c, long ldc, int nst, const double* occ, int ngw,
                   const double* g2, const double* gx, double* hpsi, long ldh, int store,
                   double* partial, double* out8, hipStream_t s,
                   float* c32 = nullptr, long ldk = 0);
int launch_lda(const double* rho, const double* rc, long n, double* vxc, double* partial,
               double* out4, hipStream_t s);
// spin-polarized LDA: vup, vdn and the same four partial sums as launch_lda
int launch_lsda(const double* rup, const double* rdn, const double* rc, long n, double* vup,
                double* vdn, double* partial, double* out4, hipStream_t s);
// core density in G of the current structure factor, and the core term of dxc
void launch_core_g(int ngv, int nsp, const double* sfac, const double* rhoc, double spin_fac,
                   double* rhocg, hipStream_t s);
void launch_core_dxc(int ngv, const double* vxcg, const double* rhocg, double* partial,
                     double* out1, hipStream_t s);
void launch_lda_points(const double* rho, int n, double* e, double* v, hipStream_t s);
int launch_vhxc_g(const double* rhog, int ngv, double omega_inv, const double* vion,
                  const double* dvion, const double* rhopst, const double* g2i, const double* gi,
                  const double* gx, double* rhoelg, double* rhogt, double* vlocal, int stress,
                  double* partial, double* out18, hipStream_t s);
void launch_sf_tables(const double* tau, int nat, const double* b, const int* kmin, int krange,
                      double* tab, hipStream_t s);
void launch_sfac(const int32_t* idx, int ngv, const double* tab, int krange, const int* kmin,
                 int nsp, const int32_t* sp_first, const double* vps, const double* dvps,
                 const double* rhops, double* sfac, double* vion, double* dvion, double* rhopst,
                 hipStream_t s);
// local forces; grp[3*g + {0,1,2}] = first atom, atoms (<= the group size),
// species of group g; vt holds 2*nsp*ngv doubles
int local_forces_group_size(int krange);
void launch_local_forces(const int32_t* idx, int ngv, const double* tab, int krange,
                         const int* kmin, const int32_t* grp, int ngrp, int nsp,
                         const double* vps, const double* rhops, const double* g2i,
                         const double* gx, const double* rhogt, const double* rhoelg,
                         const double* vxcg, const double* rhoc, double omega, const double* fesr, int nat, double* vt, double* partial,
                         int nchunk, double* fion, hipStream_t s);
// anl columns of every atom's projectors (FP64 if prec == 0, else FP32)
void launch_anl(int ngw, int nat, long lda, const double* gx, const double* tau,
                const int32_t* a_m0, const int32_t* a_npr, const int32_t* a_stride,
                const int32_t* m_twnl, const int32_t* m_l, const double* twnl, void* A, int prec,
                hipStream_t s);
int launch_dmat_enl(const double* F, int M, int nst, const double* occ, const int32_t* m_ipr,
                    const int32_t* m_npr, const int32_t* m_stride, const int32_t* m_dbase,
                    const double* dmat, const double* m_wt, double omega_inv, double* fout,
                    double* partial, double* out1, hipStream_t s);
void launch_nl_forces(const double* fout, const double* dF, int M, int nst, const double* occ,
                      const int32_t* a_m0, const int32_t* a_npr, const int32_t* a_stride,
                      int nat, double* fion, hipStream_t s);

void launch_force_w(const double* fout, const double* occ, int M, int nst, double* W,
                    hipStream_t s);
// complex (k != 0) bands on a real band grid (kpoint.hip): each band n of c
// (on S) to the pair (c1, c2) at columns 2n, 2n+1 of out (on U's half
// sphere), and the pair transform's output back, added into hpsi
void launch_fold(const double* c, long ldc, int nst, int ngwu, const int32_t* fold, double* out,
                 long ldu, hipStream_t s);
void launch_unfold_add(const double* H, long ldu, int nst, int ngw, const int32_t* unfold,
                       double* hpsi, long ldh, hipStream_t s);
// rows [r0, r1) of ncols columns (leading dimension ld) of a set to zero
void launch_zero_rows_f32(float* a, long ld, long r0, long r1, int ncols, hipStream_t s);
// a k-point set's scalars: kin[0..7], enl[0], nls[0..5] = the sums over the
// k-points of kacc[16 ik + {0..7, 8, 9..14}]
void launch_kacc_sum(const double* kacc, int nkp, double* kin, double* enl, double* nls,
                     hipStream_t s);
template <typename T>
void launch_nl_forces_y(const T* A, const T* Y, long lda, int ngw, const double* gx,
                        const int32_t* a_m0, const int32_t* a_npr, const int32_t* a_stride,
                        int nat, double* partial, int nchunk, double* fion, hipStream_t s);

template <typename T>
void launch_nl_stress_y(const T* Y, long lda, int ngw, const double* gx, const double* tau,
                        int nsp, const int32_t* sp_first, const int32_t* sp_npr,
                        const int64_t* dt_off, const double* dtwnl, const int32_t* a_m0,
                        const int32_t* a_stride, const int32_t* m_l, double* partial,
                        int nchunk, double* out6, hipStream_t s, int halves = 1, long mhalf = 0);
void launch_hart_ps_stress(const double* rhogt, const double* sfac, const double* rhops,
                           const double* g2i, const double* gx, int ngv, int nsp,
                           const double* fac2, const double* vxcg, const double* drhoc,
                           double* partial, double* out6, hipStream_t s);

// gemm.hip.  prec (QBG_PREC_FP64 / QBG_PREC_FP32) selects the MFMA compute
// type; operands of either type are converted as they are staged.
// TN split-K: out[m + n*ldo] = alpha sum_k A[k + m*lda] B[k + n*ldb]
//                              + gamma A[m*lda] B[n*ldb]   (partials in FP64)
// skip: a device flag; when it reads nonzero the launches do nothing
int gemm_tn_splitk(const double* A, long lda, const double* B, long ldb, int M, int N, long K,
                   double alpha, double gamma, double* out, long ldo, double* work,
                   size_t work_doubles, int prec, hipStream_t s, const int* skip = nullptr);
int gemm_tn_splitk(const float* A, long lda, const double* B, long ldb, int M, int N, long K,
                   double alpha, double gamma, double* out, long ldo, double* work,
                   size_t work_doubles, int prec, hipStream_t s);
// symmetric TN (FP64): out = alpha A^T A + gamma a0 a0^T, only the lower
// tiles formed, the upper part mirrored
int gemm_syrk(const double* A, long lda, int N, long K, double alpha, double gamma, double* out,
              long ldo, double* work, size_t work_doubles, hipStream_t s,
              const int* skip = nullptr);
// NN (FP64) split over the reduction for small outputs (partials in work):
// C[k + n*ldc] = sum_m A[k + m*lda] F[m + n*ldf].  Returns -1 on a bad shape.
int gemm_nn_splitk(const double* A, long lda, const double* F, long ldf, long K, int N, int M,
                   double* C, long ldc, double* work, size_t work_doubles, hipStream_t s,
                   const int* skip = nullptr);
// NN (FP64) with F upper triangular (F[m + n*ldf] = 0 for m > n), N x N:
// C[k + n*ldc] = sum_{m <= n} A[k + m*lda] F[m + n*ldf]
// output columns (p) of one tile of the NN product: a gemm_nn call with N <= this
// runs one tile per group of rows, which reads its rows before it writes them
int gemm_nn_tile_p();
void gemm_nn_triu(const double* A, long lda, const double* F, long ldf, long K, int N, double* C,
                  long ldc, hipStream_t s);
// NN: C[k + n*ldc] (+)= sum_m A[k + m*lda] F[m + n*ldf]  (K rows, N cols, M sum)
void gemm_nn(const double* A, long lda, const double* F, long ldf, long K, int N, int M,
             double* C, long ldc, int accumulate, int prec, hipStream_t s,
             const int* skip = nullptr);
void gemm_nn(const float* A, long lda, const double* F, long ldf, long K, int N, int M,
             double* C, long ldc, int accumulate, int prec, hipStream_t s);
void gemm_nn(const double* A, long lda, const double* F, long ldf, long K, int N, int M,
             float* C, long ldc, int accumulate, int prec, hipStream_t s);
// gemm_bf.hip: FP32 operands split into `parts` (2 or 3) bfloat16 parts on
// the BF16 MFMA path; extents padded (Mp, Np, Kp multiples of 128, Kp of 32,
// leading dimensions of 4) with zero padding.  Returns -1 on a bad shape.
// TN: out[m + n*ldo] = alpha sum_k A[k + m*lda] B[k + n*ldb] + gamma A[m*lda] B64[n*ldb64]
// launch plan of the TN product: edge 256 (the 256-tile kernel where the
// padded extents allow it) or 128; split = reduction splits (0: enough to
// fill the device once)
struct BfTnPlan {
  int edge = 256;
  int split = 0;
};
int gemm_bf_tn(const float* A, long lda, const float* B, long ldb, int Mp, int Np, long Kp,
               int M, int N, const double* B64, long ldb64, double alpha, double gamma,
               double* out, long ldo, float* work, size_t work_floats, int parts,
               hipStream_t s, const BfTnPlan* plan = nullptr);
// the reduction splits the default plan of gemm_bf_tn uses for this shape
int gemm_bf_tn_default_split(int Mp, int Np, long Kp, int parts, int edge);
// NN: C[k + n*ldc] (+)= sum_m A[k + m*lda] F[m + n*ldf], k < K, n < N stored
// force product with the fused force epilogue (256-tile shapes, 2-part split);
// -1 when the shape does not allow it
int gemm_bf_nn_forces(const float* A, long lda, const float* F, long ldf, long Kp, int Np,
                      int Mp, int N, const float* anl, long lda_anl, const double* gx, int ngw,
                      double* fpart, hipStream_t s);
// fion[3a+j] -= 4 sum over the atom's projectors and the ntq tiles of fpart
void launch_nl_forces_tiles(const double* fpart, int M, int ntq, double* red,
                            const int32_t* a_m0, const int32_t* a_npr,
                            const int32_t* a_stride, int nat, double* fion, hipStream_t s);
int gemm_bf_nn(const float* A, long lda, const float* F, long ldf, long Kp, int Np, int Mp,
               long K, int N, double* C, long ldc, int accumulate, int parts, hipStream_t s);
int gemm_bf_nn(const float* A, long lda, const float* F, long ldf, long Kp, int Np, int Mp,
               long K, int N, float* C, long ldc, int accumulate, int parts, hipStream_t s);
// c (FP64, real view ld ldc2) -> c32[k + n*ldk]; band columns [nst, nzero)
// set to zero
void convert_f32(const double* c, long ldc2, long K, int nst, float* c32, long ldk,
                 hipStream_t s, int nzero = 0);
// fout32[m + n*ldf] = fout[m + n*M]; w32[n + m*ldw] = occ_n fout[m + n*M] (if w32);
// band columns [nst, nzero) set to zero
void fout_f32(const double* fout, const double* occ, int M, int nst, float* f32, long ldf,
              float* w32, long ldw, hipStream_t s, int nzero = 0);

// lower Cholesky in place; info[0] = 0 or the failing pivot + 1.  With a
// workspace of 64 * n doubles (ws) the factorization runs blocked over the
// whole device, otherwise in one work-group
void potrf(double* S, int n, int lds, int* info, hipStream_t s, double* ws = nullptr);
// the blocked factorization for any n (ws: 64 * n doubles); info = -1 and
// nothing done when *skip
void potrf_device(double* S, int n, int lds, int* info, double* ws, hipStream_t s,
                  const int* skip);
void trtri(const double* L, int n, int ldl, double* X, hipStream_t s, const int* skip = nullptr);
void cleanup(double* c, long ldc, int nst, hipStream_t s);

}  // namespace qbg

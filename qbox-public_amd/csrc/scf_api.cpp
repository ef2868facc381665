// scf_api.cpp -- C ABI of the SCF-iteration updates on the device: the
// wavefunction steppers of BOSampleStepper (SD, PSD, PSDA with their
// orthogonalizations), the charge mixer, and DoubleMatrix::polar.  Each
// update enqueues a fixed sequence of kernels on the context stream and
// allocates nothing, so an SCF iteration (density, mixing, vhxc, H*psi,
// stepper) can be captured as one graph.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>
#include "ef.h"
#include "energy.h"
#include "internal.h"
#include "scf.h"

using namespace qbg;

namespace {

struct Allocs {
  std::vector<void*> p;
  template <typename T>
  int get(qbg_ctx* c, T** out, size_t n) {
    *out = nullptr;
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess)
      return c->fail(QBG_E_NOMEM, "hipMalloc failed (SCF workspace)");
    p.push_back(q);
    *out = (T*)q;
    return QBG_OK;
  }
  void release() {
    for (void* q : p) (void)hipFree(q);
    p.clear();
  }
};

// workspaces of DoubleMatrix::polar on an n x n matrix
struct PolarWs {
  double *Q = nullptr, *T = nullptr, *QT = nullptr, *XP = nullptr, *aug = nullptr;
  int* flags = nullptr;
  double* qnrm = nullptr;
  int alloc(qbg_ctx* c, Allocs& a, int n) {
    const size_t n2 = (size_t)n * n;
    int rc;
    if ((rc = a.get(c, &Q, n2)) || (rc = a.get(c, &T, n2)) || (rc = a.get(c, &QT, n2)) ||
        (rc = a.get(c, &XP, n2)) || (rc = a.get(c, &aug, 2 * n2)) || (rc = a.get(c, &flags, 8)) ||
        (rc = a.get(c, &qnrm, 1)))
      return rc;
    return QBG_OK;
  }
};

// DoubleMatrix::polar (src/Matrix.cpp:2446-2517) as a fixed launch sequence:
// maxiter rounds, each skipped on the device once converged
int polar_enqueue(const PolarWs& w, double* Y, int n, double tol, int maxiter, double* W,
                  size_t wsize, hipStream_t s) {
  launch_polar_init(w.flags, s);
  const long n2 = (long)n * n;
  for (int it = 0; it < maxiter; it++) {
    // q = I - Y^T Y
    if (gemm_syrk(Y, n, n, n, -1.0, 0.0, w.Q, n, W, wsize, s, w.flags + 0)) return QBG_E_INVAL;
    launch_polar_norm(w.Q, n, tol, w.flags, w.qnrm, s);
    // Bjorck-Bowie: Y <- Y (I + q/2 (I + 3q/4))
    launch_ident_axpy(w.T, w.Q, n, 0.75, w.flags + 1, s);
    if (gemm_nn_splitk(w.Q, n, w.T, n, n, n, n, w.QT, n, W, wsize, s, w.flags + 1))
      return QBG_E_INVAL;
    launch_ident_axpy(w.T, w.QT, n, 0.5, w.flags + 1, s);
    if (gemm_nn_splitk(Y, n, w.T, n, n, n, n, w.XP, n, W, wsize, s, w.flags + 1))
      return QBG_E_INVAL;
    launch_copy_if_not(Y, w.XP, n2, w.flags + 1, s);
    // Higham: Y <- (Y + Y^-T) / 2, Y^-T = Y (Y^T Y)^-1 with Y^T Y = I - q
    // (reusing the Bjorck-Bowie buffers, idle in a Higham step)
    double* M = w.T;
    double* Linv = w.QT;
    double* G = w.XP;
    double* P = w.aug;
    double* Z = w.aug + n2;
    launch_ident_axpy(M, w.Q, n, -1.0, w.flags + 2, s);
    potrf_device(M, n, n, w.flags + 5, W, s, w.flags + 2);
    trtri(M, n, n, Linv, s, w.flags + 2);
    launch_transpose_if(G, Linv, n, w.flags + 2, s);
    if (gemm_syrk(G, n, n, n, 1.0, 0.0, P, n, W, wsize, s, w.flags + 2)) return QBG_E_INVAL;
    if (gemm_nn_splitk(Y, n, P, n, n, n, n, Z, n, W, wsize, s, w.flags + 2))
      return QBG_E_INVAL;
    launch_higham_update(Y, Z, n, w.flags + 2, s);
    launch_polar_singular(w.flags + 5, w.flags, s);
  }
  return QBG_OK;
}

// split-K partials of the n x n products: up to 64 splits for n <= 1024 (the
// symmetric overlap at 512 bands has ten 128-tiles and fills 256 CUs twice
// at 51 splits), 32 above
size_t gemm_work(int nst) {
  const size_t n2 = (size_t)nst * nst;
  return std::max<size_t>((size_t)1 << 20, (nst <= 1024 ? 64 : 32) * n2);
}

}  // namespace

struct qbg_wfstep {
  qbg_ef* ef = nullptr;
  qbg_ctx* ctx = nullptr;
  int kind = 0;
  double param = 0.0;
  int ngw = 0, nst = 0;
  bool bands = false;     // created for a band set (qbg_wfstep_create_bands)
  // nst x nst matrices: residual A, overlap S / Cholesky L, X = L^-T,
  // B = c^T c_last, Y = L^-1 B (polar), Z = L^-T Q
  double *A = nullptr, *S = nullptr, *X = nullptr, *B = nullptr, *Y = nullptr, *Z = nullptr;
  double *Tc = nullptr, *W = nullptr;    // 2 ngw x nst product; split-K partials
  size_t wsize = 0;
  double *cl = nullptr, *dl = nullptr;   // PSDA: previous bands and directions
  double *ekin = nullptr, *dotp = nullptr, *partial = nullptr, *scal = nullptr;
  double* bpart = nullptr;               // per-band partial sums
  int *info = nullptr, *extrap = nullptr;
  int npart = 0;
  PolarWs pw;
  double* h_out = nullptr;               // pinned: eigsum, theta, iterations, status
  Allocs mem;
};

struct qbg_mixer {
  qbg_ef* ef = nullptr;
  qbg_ctx* ctx = nullptr;
  double alpha = 0.5;
  int nmax = 0, anderson = 0, ng = 0, nsp = 1;
  double *wk = nullptr, *wls = nullptr, *rhog_in = nullptr, *xs = nullptr, *fs = nullptr;
  double *partial = nullptr, *theta = nullptr, *rhor_new = nullptr;
  double *rhog_sp = nullptr, *rhog_tot = nullptr;   // nspin 2: rhog per spin, their sum
  int* st = nullptr;
  Allocs mem;
};

namespace {

// SlaterDet::gram (src/SlaterDet.cpp:480-534) without host synchronization:
// the coefficients are replaced only if the Cholesky succeeded (*info == 0)
int gram_enqueue(qbg_wfstep* w, double* c, int ldc, hipStream_t s) {
  const long K = 2L * w->ngw, ldc2 = 2L * ldc;
  const int n = w->nst;
  cleanup(c, ldc, n, s);
  if (gemm_syrk(c, ldc2, n, K, 2.0, -1.0, w->S, n, w->W, w->wsize, s)) return QBG_E_INVAL;
  potrf(w->S, n, n, w->info, s, w->W);
  trtri(w->S, n, n, w->X, s);
  gemm_nn_triu(c, ldc2, w->X, n, K, n, w->Tc, K, s);
  launch_copy_cols_if(c, ldc2, w->Tc, K, K, n, w->info, s);
  return QBG_OK;
}

// SlaterDet::ortho_align (src/SlaterDet.cpp:768-876): c <- c L^-T Q with
// L L^T = c^T c and Q the polar factor of L^-1 c^T c_last
int ortho_align_enqueue(qbg_wfstep* w, double* c, int ldc, hipStream_t s) {
  const long K = 2L * w->ngw, ldc2 = 2L * ldc;
  const int n = w->nst;
  cleanup(c, ldc, n, s);
  if (gemm_syrk(c, ldc2, n, K, 2.0, -1.0, w->S, n, w->W, w->wsize, s)) return QBG_E_INVAL;
  potrf(w->S, n, n, w->info, s, w->W);
  trtri(w->S, n, n, w->X, s);                       // X = L^-T (column major)
  if (gemm_tn_splitk(c, ldc2, w->cl, K, n, n, K, 2.0, -1.0, w->B, n, w->W, w->wsize,
                     QBG_PREC_FP64, s))
    return QBG_E_INVAL;
  // Y = X^T B = L^-1 B
  if (gemm_tn_splitk(w->X, n, w->B, n, n, n, n, 1.0, 0.0, w->Y, n, w->W, w->wsize,
                     QBG_PREC_FP64, s))
    return QBG_E_INVAL;
  int rc = polar_enqueue(w->pw, w->Y, n, 1.e-6, 20, w->W, w->wsize, s);
  if (rc) return rc;
  if (gemm_nn_splitk(w->X, n, w->Y, n, n, n, n, w->Z, n, w->W, w->wsize, s))   // Z = L^-T Q
    return QBG_E_INVAL;
  gemm_nn(c, ldc2, w->Z, n, K, n, n, w->Tc, K, 0, QBG_PREC_FP64, s);  // c Z
  launch_copy_cols_if(c, ldc2, w->Tc, K, K, n, w->info, s);
  return QBG_OK;
}

}  // namespace

extern "C" {

// nst = 0: all bands of ef (nspin 1); nst > 0: a band set of that size, as
// one spin's SlaterDet (the caller passes that spin's columns to update)
static int wfstep_create(qbg_ef* ef, int kind, double param, int nst, qbg_wfstep** out) {
  if (!out) return QBG_E_INVAL;
  *out = nullptr;
  if (!ef) return QBG_E_INVAL;
  qbg_ctx* c = ef->ctx;
  if (kind < QBG_WF_SD || kind > QBG_WF_PSDA) return c->fail(QBG_E_INVAL, "wfstep: unknown kind");
  if (ef->cplx) return c->fail(QBG_E_INVAL, "wfstep: real (Gamma) basis only");
  if (nst == 0 && ef->nspin != 1)
    return c->fail(QBG_E_INVAL, "wfstep: with nspin 2 create one stepper per spin "
                                "(qbg_wfstep_create_bands)");
  if (nst < 0 || nst > ef->nst) return c->fail(QBG_E_INVAL, "wfstep: band count");
  if (!(param >= 0.0)) return c->fail(QBG_E_INVAL, "wfstep: negative parameter");
  if (c->capturing) return c->fail(QBG_E_STATE, "wfstep_create during graph capture");
  if ((nst ? nst : ef->nst) > POLAR_NMAX)
    return c->fail(QBG_E_INVAL, "wfstep: more than 4096 bands");
  QBG_HIP(c, hipSetDevice(c->device));
  qbg_wfstep* w = new (std::nothrow) qbg_wfstep;
  if (!w) return QBG_E_NOMEM;
  w->ef = ef;
  w->ctx = c;
  w->kind = kind;
  w->param = param;
  w->ngw = ef->wb.size;
  w->nst = nst ? nst : ef->nst;
  w->bands = nst > 0;
  const size_t n2 = (size_t)w->nst * w->nst, K = 2 * (size_t)w->ngw;
  w->wsize = gemm_work(w->nst);
  w->npart = psda_partials(w->ngw, w->nst);
  Allocs& a = w->mem;
  int rc;
  if ((rc = a.get(c, &w->S, n2)) || (rc = a.get(c, &w->X, n2)) || (rc = a.get(c, &w->Tc, K * w->nst)) ||
      (rc = a.get(c, &w->W, w->wsize)) || (rc = a.get(c, &w->ekin, w->nst)) ||
      (rc = a.get(c, &w->dotp, w->nst)) || (rc = a.get(c, &w->scal, 4)) ||
      (rc = a.get(c, &w->bpart, band_partials(w->nst))) ||
      (rc = a.get(c, &w->info, 1)) || (rc = a.get(c, &w->extrap, 1)) ||
      (kind != QBG_WF_SD && (rc = a.get(c, &w->A, n2))) ||
      (kind == QBG_WF_PSDA &&
       ((rc = a.get(c, &w->B, n2)) || (rc = a.get(c, &w->Y, n2)) || (rc = a.get(c, &w->Z, n2)) ||
        (rc = a.get(c, &w->cl, K * w->nst)) || (rc = a.get(c, &w->dl, K * w->nst)) ||
        (rc = a.get(c, &w->partial, 2 * (size_t)w->npart)) || (rc = w->pw.alloc(c, a, w->nst))))) {
    a.release();
    delete w;
    return rc;
  }
  if (hipHostMalloc((void**)&w->h_out, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    a.release();
    delete w;
    return c->fail(QBG_E_NOMEM, "hipHostMalloc failed (wfstep results)");
  }
  for (int i = 0; i < 4; i++) w->h_out[i] = 0.0;
  hipStream_t s = c->stream;
  fill_zero(w->scal, 4, s);
  launch_set_int(w->info, 0, s);
  launch_set_int(w->extrap, 0, s);
  if (w->cl) {
    fill_zero(w->cl, K * w->nst, s);
    fill_zero(w->dl, K * w->nst, s);
    launch_polar_init(w->pw.flags, s);
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    (void)hipHostFree(w->h_out);
    a.release();
    delete w;
    return c->fail(QBG_E_HIP, "wfstep_create: initialization failed");
  }
  *out = w;
  return QBG_OK;
}

int qbg_wfstep_create(qbg_ef* ef, int kind, double param, qbg_wfstep** out) {
  return qbg::guarded((ef ? ef->ctx : nullptr), [&]() -> int {
    return wfstep_create(ef, kind, param, 0, out);
  });
}

int qbg_wfstep_create_bands(qbg_ef* ef, int kind, double param, int nst, qbg_wfstep** out) {
  return qbg::guarded((ef ? ef->ctx : nullptr), [&]() -> int {
    if (nst <= 0) {
      if (out) *out = nullptr;
      return ef ? ef->ctx->fail(QBG_E_INVAL, "wfstep: band count") : QBG_E_INVAL;
    }
    return wfstep_create(ef, kind, param, nst, out);
  });
}

int qbg_wfstep_destroy(qbg_wfstep* w) {
  return qbg::guarded((w ? w->ctx : nullptr), [&]() -> int {
    if (!w) return QBG_E_INVAL;
    QBG_HIP(w->ctx, hipSetDevice(w->ctx->device));
    (void)hipStreamSynchronize(w->ctx->stream);
    if (w->h_out) (void)hipHostFree(w->h_out);
    w->mem.release();
    delete w;
    return QBG_OK;
  });
}

int qbg_wfstep_preprocess(qbg_wfstep* w) {
  return qbg::guarded((w ? w->ctx : nullptr), [&]() -> int {
    if (!w) return QBG_E_INVAL;
    // a kernel rather than a memset node, so the call can be captured
    launch_set_int(w->extrap, 0, w->ctx->stream);
    return QBG_OK;
  });
}

int qbg_wfstep_update(qbg_wfstep* w, double* c, int ldc, double* dc, int ldd) {
  return qbg::guarded((w ? w->ctx : nullptr), [&]() -> int {
    if (!w || !c || !dc) return QBG_E_INVAL;
    qbg_ctx* ctx = w->ctx;
    if (ldc < w->ngw || ldd < w->ngw) return ctx->fail(QBG_E_INVAL, "wfstep_update: leading dimension");
    if (!w->bands && w->ef->nspin != 1)
      return ctx->fail(QBG_E_STATE, "wfstep_update: the functional was switched to nspin 2");
    QBG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int n = w->nst, ngw = w->ngw;
    const long K = 2L * ngw, ldc2 = 2L * ldc, ldd2 = 2L * ldd;
    int rc = QBG_OK;
    ctx->prof_begin(QBG_K_WFSTEP);
    launch_band_sums(c, ldc2, dc, ldd2, w->ef->d_g2w, ngw, n, w->bpart, w->ekin, w->dotp, s);
    launch_sum(w->dotp, n, w->scal, s);
    if (w->kind == QBG_WF_SD) {
      launch_sd_update(c, ldc2, dc, ldd2, ngw, n, w->param, s);
      rc = gram_enqueue(w, c, ldc, s);
    } else {
      // residual: A = 2 c^T dc - c0 dc0^T, dc <- dc - c A (stored as -A)
      if (gemm_tn_splitk(c, ldc2, dc, ldd2, n, n, K, -2.0, 1.0, w->A, n, w->W, w->wsize,
                         QBG_PREC_FP64, s))
        rc = QBG_E_INVAL;
      else
        gemm_nn(c, ldc2, w->A, n, K, n, n, dc, ldd2, 1, QBG_PREC_FP64, s);
      if (!rc && w->kind == QBG_WF_PSD) {
        launch_psd_update(c, ldc2, dc, ldd2, w->ef->d_g2w, w->ekin, ngw, n, w->param, s);
        rc = gram_enqueue(w, c, ldc, s);
      } else if (!rc) {
        launch_psda_prec(dc, ldd2, w->dl, w->ef->d_g2w, w->ekin, ngw, n, w->param, w->partial, s);
        launch_psda_theta(w->partial, w->npart, w->extrap, w->scal + 1, s);
        launch_psda_update(c, ldc2, dc, ldd2, w->cl, w->dl, w->scal + 1, ngw, n, s);
        rc = ortho_align_enqueue(w, c, ldc, s);
      }
    }
    if (!rc) launch_wf_results(w->scal, w->kind == QBG_WF_PSDA ? w->pw.flags : nullptr, w->info,
                               w->h_out, s);
    ctx->prof_end();
    if (rc) return ctx->fail(rc, "wfstep_update: workspace too small");
    hipError_t er = hipGetLastError();
    if (er != hipSuccess) return ctx->fail(QBG_E_HIP, hipGetErrorString(er));
    return QBG_OK;
  });
}

int qbg_wfstep_results(qbg_wfstep* w, double* out) {
  return qbg::guarded((w ? w->ctx : nullptr), [&]() -> int {
    if (!w || !out) return QBG_E_INVAL;
    qbg_ctx* ctx = w->ctx;
    if (ctx->capturing) return ctx->fail(QBG_E_STATE, "wfstep_results during graph capture");
    QBG_HIP(ctx, hipSetDevice(ctx->device));
    hipError_t er = hipStreamSynchronize(ctx->stream);
    if (er != hipSuccess) return ctx->fail(QBG_E_HIP, hipGetErrorString(er));
    for (int i = 0; i < 4; i++) out[i] = w->h_out[i];
    if (w->h_out[3] != 0.0)
      return ctx->fail(QBG_E_NOT_SPD, "wfstep: orthogonalization failed (overlap not positive definite)");
    return QBG_OK;
  });
}

// ----------------------------------------------------------------- mixer
int qbg_mixer_create(qbg_ef* ef, double alpha, double rc_kerker, int ndim, double rc1,
                     qbg_mixer** out) {
  return qbg::guarded((ef ? ef->ctx : nullptr), [&]() -> int {
    if (!out) return QBG_E_INVAL;
    *out = nullptr;
    if (!ef) return QBG_E_INVAL;
    qbg_ctx* c = ef->ctx;
    if (ndim < 0 || ndim > MIX_NMAX) return c->fail(QBG_E_INVAL, "mixer: ndim must be 0..8");
    if (!(alpha >= 0.0) || !(rc1 >= 0.0)) return c->fail(QBG_E_INVAL, "mixer: bad parameters");
    if (c->capturing) return c->fail(QBG_E_STATE, "mixer_create during graph capture");
    QBG_HIP(c, hipSetDevice(c->device));
    qbg_mixer* m = new (std::nothrow) qbg_mixer;
    if (!m) return QBG_E_NOMEM;
    m->ef = ef;
    m->ctx = c;
    m->alpha = alpha;
    m->nmax = ndim;
    m->anderson = ndim > 0 ? 1 : 0;
    m->ng = ef->vb.size;
    m->nsp = ef->nspin;
    const int ng = m->ng, nsp = m->nsp;
    // Kerker and least-squares row weights (src/BOSampleStepper.cpp:341-406)
    std::vector<double> wk(ng, 1.0), wls(ng, 1.0);
    const std::vector<double>& g2 = ef->vb.g2;
    if (rc1 != 0.0) {
      const double q1 = 2.0 * M_PI / rc1;
      for (int i = 0; i < ng; i++) wls[i] = g2[i] != 0.0 ? std::sqrt(g2[i] / (g2[i] + q1 * q1)) : 1.0;
    }
    if (rc_kerker > 0.0) {
      const double q0 = 2 * M_PI / rc_kerker, q02 = q0 * q0;
      for (int i = 0; i < ng; i++) wk[i] = g2[i] != 0.0 ? g2[i] / (g2[i] + q02) : 1.0;
    }
    Allocs& a = m->mem;
    // the mixed vector holds the densities of both spins (src/BOSampleStepper.cpp:318-327)
    const size_t mlen = 2 * (size_t)nsp * ng;
    int rc;
    if ((rc = a.get(c, &m->wk, ng)) || (rc = a.get(c, &m->wls, ng)) ||
        (rc = a.get(c, &m->rhog_in, mlen)) || (rc = a.get(c, &m->rhor_new, nsp * ef->N)) ||
        (nsp == 2 && ((rc = a.get(c, &m->rhog_sp, mlen)) ||
                      (rc = a.get(c, &m->rhog_tot, 2 * (size_t)ng)))) ||
        (rc = a.get(c, &m->st, 4)) || (rc = a.get(c, &m->theta, MIX_NMAX)) ||
        (m->anderson && ((rc = a.get(c, &m->xs, (m->nmax + 1) * mlen)) ||
                         (rc = a.get(c, &m->fs, (m->nmax + 1) * mlen)) ||
                         (rc = a.get(c, &m->partial, mix_partials()))))) {
      a.release();
      delete m;
      return rc;
    }
    hipStream_t s = c->stream;
    bool ok = hipMemcpyAsync(m->wk, wk.data(), 8 * ng, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(m->wls, wls.data(), 8 * ng, hipMemcpyHostToDevice, s) == hipSuccess;
    fill_zero(m->rhog_in, mlen, s);
    fill_zero(m->rhor_new, nsp * ef->N, s);
    fill_zero(m->theta, MIX_NMAX, s);
    if (m->anderson) {
      fill_zero(m->xs, (m->nmax + 1) * mlen, s);
      fill_zero(m->fs, (m->nmax + 1) * mlen, s);
    }
    launch_mix_restart(m->st, s);
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) {
      a.release();
      delete m;
      return c->fail(QBG_E_HIP, "mixer_create: initialization failed");
    }
    *out = m;
    return QBG_OK;
  });
}

int qbg_mixer_destroy(qbg_mixer* m) {
  return qbg::guarded((m ? m->ctx : nullptr), [&]() -> int {
    if (!m) return QBG_E_INVAL;
    QBG_HIP(m->ctx, hipSetDevice(m->ctx->device));
    (void)hipStreamSynchronize(m->ctx->stream);
    m->mem.release();
    delete m;
    return QBG_OK;
  });
}

int qbg_mixer_restart(qbg_mixer* m) {
  return qbg::guarded((m ? m->ctx : nullptr), [&]() -> int {
    if (!m) return QBG_E_INVAL;
    launch_mix_restart(m->st, m->ctx->stream);
    return QBG_OK;
  });
}

// BOSampleStepper's mixing step (src/BOSampleStepper.cpp:810-879) and
// ChargeDensity::update_rhor (src/ChargeDensity.cpp:201-221).  nspin 2: the
// vector mixed is rhog of the up spin followed by rhog of the down spin (each
// the forward transform of that spin's rhor); after mixing, each rhor is the
// inverse transform of its mixed rhog and the total rhog their sum
// (src/EnergyFunctional.cpp:256-261)
int qbg_mixer_update(qbg_mixer* m) {
  return qbg::guarded((m ? m->ctx : nullptr), [&]() -> int {
    if (!m) return QBG_E_INVAL;
    qbg_ef* e = m->ef;
    qbg_ctx* ctx = m->ctx;
    if (!e->have_rho) return ctx->fail(QBG_E_STATE, "mixer_update before update_density");
    if (e->nspin != m->nsp)
      return ctx->fail(QBG_E_STATE, "mixer_update: nspin changed since the mixer was created");
    QBG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    qbg_grid* g = e->gv;
    int rc = grid_ensure_ws(g, 1);
    if (rc) return rc;
    const int ng = m->ng, nsp = m->nsp, nv = nsp * ng;
    const long N = e->N;
    double* rhor[2] = {e->d_rhor, e->d_rhor2};
    double* rg = nsp == 2 ? m->rhog_sp : e->d_rhog;
    int r = 0;
    ctx->prof_begin(QBG_K_MIX);
    for (int sp = 0; nsp == 2 && sp < 2 && !r; sp++) {
      r = ypass(g, QBG_PREC_FP64, Y_FWD_REAL, 1, rhor[sp], nullptr, nullptr, nullptr, e->omega, s);
      r = r ? r : xfwd(g, QBG_PREC_FP64, 1, s);
      r = r ? r : zfwd(g, QBG_PREC_FP64, Z_SINGLE, OUT_STORE, rg + 2L * sp * ng, 0, 0, 0, 1, s);
    }
    if (!r) {
      launch_mix_begin(m->st, m->nmax, m->anderson, s);
      launch_mix_stage(rg, m->rhog_in, m->wls, m->wk, m->xs, m->fs, m->st, nv, ng, m->alpha,
                       m->anderson, s);
      if (m->anderson) {
        launch_mix_dots(m->fs, m->st, m->nmax, 2L * nv, m->partial, s);
        launch_mix_solve(m->partial, m->st, m->anderson, m->theta, s);
        launch_mix_apply(rg, m->rhog_in, m->wls, m->wk, m->xs, m->fs, m->st, m->theta, m->nmax,
                         nv, ng, m->alpha, m->anderson, s);
      }
    }
    // rhor = Re(backward(rhog)) / omega per spin, kept only when this call mixed
    for (int sp = 0; sp < nsp && !r; sp++) {
      r = zbwd(g, QBG_PREC_FP64, Z_SINGLE, rg + 2L * sp * ng, 0, 0, 0, 1, s);
      r = r ? r : xbwd(g, QBG_PREC_FP64, 1, s);
      r = r ? r : ypass(g, QBG_PREC_FP64, Y_BWD_REAL_ADD, 1, nullptr, nullptr, nullptr,
                        m->rhor_new + sp * N, 1.0 / e->omega, s);
      if (!r) launch_mix_select(rhor[sp], m->rhor_new + sp * N, N, m->st, s);
    }
    if (!r && nsp == 2) {
      add2(m->rhog_sp, m->rhog_sp + 2L * ng, m->rhog_tot, 2L * ng, s);
      launch_mix_select(e->d_rhog, m->rhog_tot, 2L * ng, m->st, s);
    }
    ctx->prof_end();
    if (r) return ctx->fail(r, "mixer_update: density-grid transform refused");
    e->have_vhxc = false;
    hipError_t er = hipGetLastError();
    if (er != hipSuccess) return ctx->fail(QBG_E_HIP, hipGetErrorString(er));
    return QBG_OK;
  });
}

// ----------------------------------------------------------------- polar
int qbg_polar(qbg_ctx* c, double* x, int n, double tol, int maxiter, int* iters) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !x || n <= 0 || maxiter < 0 || maxiter > 64 || !(tol >= 0.0)) return QBG_E_INVAL;
    if (n > POLAR_NMAX) return c->fail(QBG_E_INVAL, "polar: n > 4096");
    if (c->capturing) return c->fail(QBG_E_STATE, "polar allocates: not capturable");
    QBG_HIP(c, hipSetDevice(c->device));
    Allocs a;
    PolarWs w;
    double* W = nullptr;
    const size_t wsize = gemm_work(n);
    int rc = w.alloc(c, a, n);
    if (!rc) rc = a.get(c, &W, wsize);
    if (rc) { a.release(); return rc; }
    hipStream_t s = c->stream;
    rc = polar_enqueue(w, x, n, tol, maxiter, W, wsize, s);
    int flags[5] = {0, 0, 0, 0, 0};
    hipError_t er = hipMemcpyAsync(flags, w.flags, sizeof(flags), hipMemcpyDeviceToHost, s);
    if (er == hipSuccess) er = hipStreamSynchronize(s);
    a.release();
    if (rc) return c->fail(rc, "polar: workspace");
    if (er != hipSuccess) return c->fail(QBG_E_HIP, hipGetErrorString(er));
    if (iters) *iters = flags[3];
    if (flags[4]) return c->fail(QBG_E_NOT_SPD, "polar: singular matrix in a Higham step");
    return QBG_OK;
  });
}

}  // extern "C"

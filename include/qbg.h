/* qbg.h -- C ABI of the MI355X H*psi / energy library (libqbg.so).
 *
 * The library replaces, for a single-rank (1x1 Context) Qbox run at the
 * Gamma point, the hot path inside EnergyFunctional::energy: the band FFTs
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

/* precision of the FFT passes and projector GEMMs */
#define QBG_PREC_FP64 0
#define QBG_PREC_FP32 1

typedef struct qbg_ctx qbg_ctx;
typedef struct qbg_grid qbg_grid;
typedef struct qbg_ef qbg_ef;

/* ---------------------------------------------------------------- context */
/* Replaces the MPI/BLACS context of a 1x1 run (src/Context.cpp:60-75). */
int qbg_init(int device, qbg_ctx** ctx);
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
#define QBG_K_NL 7        /* projector GEMMs and their generation */
#define QBG_K_GRAM 8      /* overlap, Cholesky, triangular update */
#define QBG_K_OTHER 9     /* reductions, pointwise, structure factor */
#define QBG_K_COUNT 10
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

#ifdef __cplusplus
}
#endif
#endif /* QBG_H */

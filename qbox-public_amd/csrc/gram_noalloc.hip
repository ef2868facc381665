// gram_noalloc.hip -- SlaterDet::gram on a fixed device array (no allocation
// per call), for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The overlap S, the triangular inverse X, the split-K partials of the
// overlap product and the Cholesky status live in `scratch`, one file-scope
// device array per GPU (GRAM_NA_BYTES, held until the process ends).  Its
// device address comes from hipGetSymbolAddress, once per device with that
// device current; kernels only ever see that address.
//
// The update C <- C L^-T runs in place, one column block of at most GEMM_TP
// columns per launch from the last block to the first: block J of the result
// needs the old columns [0, j1) only, the columns left of J are still old
// when J runs, and within a launch each output tile is written after its
// whole reduction, by the one work-group that read those rows.  So no copy
// of C is needed.
//
// Calls on one GPU share the array: calls that overlap in time on one GPU
// (two streams, two threads) race and must be serialized by the caller.
#include <hip/hip_runtime.h>
#include <mutex>
#include "energy.h"
#include "internal.h"

namespace {

// S (n^2), X (n^2), split-K partials (GRAM_NA_SPLITS n^2), status word
constexpr size_t GRAM_NA_DOUBLES =
    (size_t)(2 + QBG_GRAM_NOALLOC_SPLITS) * QBG_GRAM_NOALLOC_MAX_NST * QBG_GRAM_NOALLOC_MAX_NST +
    32;
constexpr size_t GRAM_NA_BYTES = GRAM_NA_DOUBLES * sizeof(double);

__device__ __align__(256) double scratch[GRAM_NA_DOUBLES];

constexpr int MAXDEV = 64;
double* g_scratch[MAXDEV] = {nullptr};
std::mutex g_mu;

// the device address of this device's array (the device is current)
int scratch_ptr(qbg_ctx* c, double** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (c->device < 0 || c->device >= MAXDEV) return c->fail(QBG_E_INVAL, "device index");
  if (!g_scratch[c->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return c->fail(QBG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                    hipGetErrorString(e));
    g_scratch[c->device] = static_cast<double*>(p);
  }
  *out = g_scratch[c->device];
  return QBG_OK;
}

}  // namespace

extern "C" {

size_t qbg_gram_noalloc_bytes(void) { return GRAM_NA_BYTES; }

// SlaterDet::gram (src/SlaterDet.cpp:480-534) as qbg_gram, with every
// temporary in the fixed per-device array
int qbg_gram_noalloc(qbg_ctx* c, double* cc, int ldc, int ngw, int nst, int prec) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !cc || ngw <= 0 || nst <= 0 || ldc < ngw) return QBG_E_INVAL;
    if (prec != QBG_PREC_FP64) return c->fail(QBG_E_INVAL, "gram: FP64 only");
    if (nst > QBG_GRAM_NOALLOC_MAX_NST)
      return c->fail(QBG_E_INVAL, "gram_noalloc: nst " + std::to_string(nst) +
                                      " exceeds the fixed workspace's maximum of " +
                                      std::to_string(QBG_GRAM_NOALLOC_MAX_NST));
    QBG_HIP(c, hipSetDevice(c->device));
    double* base = nullptr;
    int rc = scratch_ptr(c, &base);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const long K = 2L * ngw, lda = 2L * ldc;
    const size_t nn = (size_t)nst * nst;
    double* S = base;
    double* X = S + nn;
    double* W = X + nn;
    const size_t work = (size_t)QBG_GRAM_NOALLOC_SPLITS * nn;
    int* info = reinterpret_cast<int*>(W + work);
    c->prof_begin(QBG_K_GRAM);
    qbg::cleanup(cc, ldc, nst, s);
    if (qbg::gemm_syrk(cc, lda, nst, K, 2.0, -1.0, S, nst, W, work, s))
      return c->fail(QBG_E_INVAL, "gram_noalloc: overlap workspace");
    qbg::potrf(S, nst, nst, info, s, W);
    c->prof_end();
    int hinfo = 0;
    QBG_HIP(c, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    QBG_HIP(c, hipStreamSynchronize(s));
    if (hinfo) return c->fail(QBG_E_NOT_SPD, "gram: overlap matrix not positive definite");
    c->prof_begin(QBG_K_GRAM);
    qbg::trtri(S, nst, nst, X, s);
    // X = L^-1 row-major is L^-T column-major (upper triangular):
    // C[:, j0:j1] = C[:, 0:j1] X[0:j1, j0:j1], last block first
    const int tp = qbg::gemm_nn_tile_p();
    for (int j0 = ((nst - 1) / tp) * tp; j0 >= 0; j0 -= tp) {
      const int j1 = std::min(nst, j0 + tp);
      qbg::gemm_nn(cc, lda, X + (size_t)j0 * nst, nst, K, j1 - j0, j1, cc + (size_t)j0 * lda,
                   lda, 0, QBG_PREC_FP64, s, nullptr);
    }
    c->prof_end();
    QBG_HIP(c, hipStreamSynchronize(s));
    return QBG_OK;
  });
}

}  // extern "C"

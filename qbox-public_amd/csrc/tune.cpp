// tune.cpp -- launch tuner of a grid's batched FFT passes.
//
// At the first batched transform of a grid in a precision, outside graph
// capture, the x-pass kernels of every compiled split of np0 and the y-pass
// kernels of every compiled split of np1 (with 8, 16 and 32 band pairs per
// block for the potential multiply) are timed on the grid's own workspace
// and maps, with HIP events on the context's stream, and the fastest are
// kept for that precision.  The choice is cached per device, grid shape and
// precision for the life of the process, so every grid of one shape runs the
// same kernels (results of equal inputs stay bit-identical within a
// process).  A candidate replaces the default split only if it is faster by
// more than 3 %.  Reference counterpart: FourierTransform's FFTW_MEASURE
// plans (src/FourierTransform.cpp:45-54, 830-1062), chosen by timing when
// the object is built.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include "internal.h"

namespace qbg {

namespace {

struct Choice {
  const F4Ops* fx;
  const F4Ops* fy;
  int yppb;
};
using Key = std::tuple<int, int, int, int, int, int, int, int>;
std::map<Key, Choice> g_cache;
std::mutex g_mu;

constexpr int kReps = 3;
constexpr double kMargin = 0.97;

// average milliseconds of `launch` over kReps launches after one warm-up
template <class F>
int time_ms(hipStream_t s, hipEvent_t a, hipEvent_t b, F&& launch, float* ms) {
  int rc = launch();
  if (rc) return rc;
  if (hipEventRecord(a, s) != hipSuccess) return QBG_E_HIP;
  for (int r = 0; r < kReps && !rc; r++) rc = launch();
  if (rc) return rc;
  if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
      hipEventElapsedTime(ms, a, b) != hipSuccess)
    return QBG_E_HIP;
  *ms /= kReps;
  return QBG_OK;
}

int measure(qbg_grid* g, int prec, int pi, Choice* best) {
  qbg_ctx* c = g->ctx;
  hipStream_t s = c->stream;
  const int nb = std::max(1, std::min(g->ws_pairs, g->batch));
  const F4Ops* cx[8];
  const F4Ops* cy[8];
  const int nx = f4_candidates(g->np0, cx, 8), ny = f4_candidates(g->np1, cy, 8);
  best->fx = nx ? cx[0] : nullptr;
  best->fy = ny ? cy[0] : nullptr;
  best->yppb = g->yppb;
  const int ppbs[3] = {16, 8, 32};   // the default first: the others must beat it
  const int nppb = g->yppb_fixed ? 1 : 3;
  if (nx <= 1 && ny <= 1 && nppb == 1) return QBG_OK;
  // scratch for the passes that read a potential or accumulate a density;
  // the workspace is cleared so that no pass reads uninitialized values
  const size_t N = (size_t)g->np0 * g->np1 * g->np2;
  double *v = nullptr, *rho = nullptr, *fac = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = QBG_OK;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    if (v) (void)hipFree(v);
    if (rho) (void)hipFree(rho);
    if (fac) (void)hipFree(fac);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  };
  if (hipMalloc((void**)&v, sizeof(double) * N) != hipSuccess ||
      hipMalloc((void**)&rho, sizeof(double) * N) != hipSuccess ||
      hipMalloc((void**)&fac, sizeof(double) * 2 * nb) != hipSuccess) {
    cleanup();
    return c->fail(QBG_E_NOMEM, "hipMalloc failed (tuner scratch)");
  }
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess ||
      hipMemsetAsync(v, 0, sizeof(double) * N, s) != hipSuccess ||
      hipMemsetAsync(rho, 0, sizeof(double) * N, s) != hipSuccess ||
      hipMemsetAsync(fac, 0, sizeof(double) * 2 * nb, s) != hipSuccess ||
      hipMemsetAsync(g->d_xbuf, 0, (size_t)nb * xpair_bytes(g, QBG_PREC_FP64), s) != hipSuccess ||
      hipMemsetAsync(g->d_zbuf, 0, g->ws_bytes - (size_t)g->ws_pairs *
                                                     xpair_bytes(g, QBG_PREC_FP64), s) !=
          hipSuccess) {
    cleanup();
    return c->fail(QBG_E_HIP, "tuner set-up");
  }
  void* xws_saved = g->xws;
  g->xws = g->d_xbuf;
  const F4Ops* fx_saved = g->fx[pi];
  const F4Ops* fy_saved = g->fy[pi];
  const int yp_saved = g->yppb_t[pi];
  float t0 = 0.f;
  // x pass: backward + forward per candidate split of np0
  for (int i = 0; i < nx && !rc && nx > 1; i++) {
    g->fx[pi] = cx[i];
    float ms = 0.f;
    rc = time_ms(s, a, b, [&]() -> int {
      int r = xbwd(g, prec, nb, s);
      return r ? r : xfwd(g, prec, nb, s);
    }, &ms);
    if (i == 0) t0 = ms;
    else if (!rc && ms < kMargin * t0) { best->fx = cx[i]; t0 = ms; }
  }
  // y pass: potential multiply (per pairs-per-block) + density, per split of np1
  float ty = 0.f;
  bool first = true;
  for (int i = 0; i < ny && !rc; i++) {
    g->fy[pi] = cy[i];
    float td = 0.f;
    rc = time_ms(s, a, b, [&]() -> int {
      return ypass(g, prec, Y_DENSITY, nb, nullptr, fac, rho, nullptr, 1.0, s);
    }, &td);
    for (int j = 0; j < nppb && !rc; j++) {
      g->yppb_t[pi] = g->yppb_fixed ? g->yppb : ppbs[j];
      float tv = 0.f;
      rc = time_ms(s, a, b, [&]() -> int {
        return ypass(g, prec, Y_VMUL, nb, v, nullptr, nullptr, nullptr, 1.0, s);
      }, &tv);
      const float t = tv + td;
      if (!rc && (first || t < kMargin * ty)) {
        best->fy = cy[i];
        best->yppb = g->yppb_t[pi];
        ty = t;
        first = false;
      }
    }
  }
  g->fx[pi] = fx_saved;
  g->fy[pi] = fy_saved;
  g->yppb_t[pi] = yp_saved;
  g->xws = xws_saved;
  cleanup();
  return rc;
}

}  // namespace

int grid_tune(qbg_grid* g, int prec) {
  const int pi = prec == QBG_PREC_FP64 ? 0 : 1;
  if (g->tuned[pi] || !g->f4[0] || !g->f4[1]) return QBG_OK;
  if (g->ctx->capturing || !g->d_xbuf || !g->d_zbuf || g->ws_pairs < 1) return QBG_OK;
  int dev = g->ctx->device;
  const Key key{dev, g->np0, g->np1, g->np2, g->nact, g->ncols, pi, g->yppb_fixed ? g->yppb : 0};
  Choice ch{};
  bool have = false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      ch = it->second;
      have = true;
    }
  }
  if (!have) {
    const int rc = measure(g, prec, pi, &ch);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache.emplace(key, ch);
  }
  g->fx[pi] = ch.fx;
  g->fy[pi] = ch.fy;
  g->yppb_t[pi] = ch.yppb;
  g->tuned[pi] = true;
  return QBG_OK;
}

}  // namespace qbg

extern "C" int qbg_grid_tuning(const qbg_grid* g, int prec, int* out) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    if (!g || !out) return QBG_E_INVAL;
    const int pi = prec == QBG_PREC_FP64 ? 0 : 1;
    const qbg::F4Ops* fx = g->fx[pi] ? g->fx[pi] : g->f4[0];
    const qbg::F4Ops* fy = g->fy[pi] ? g->fy[pi] : g->f4[1];
    out[0] = g->tuned[pi] ? 1 : 0;
    out[1] = fx ? fx->n1 : 0;
    out[2] = fx ? fx->n2 : 0;
    out[3] = fy ? fy->n1 : 0;
    out[4] = fy ? fy->n2 : 0;
    out[5] = g->yppb_t[pi] > 0 ? g->yppb_t[pi] : g->yppb;
    return QBG_OK;
  });
}

extern "C" int qbg_grid_plan(qbg_grid* g, int prec, int x_n1, int y_n1, int yppb) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    if (!g || yppb < 0 || x_n1 < 0 || y_n1 < 0) return QBG_E_INVAL;
    if (prec < QBG_PREC_FP64 || prec > QBG_PREC_BF16X2) return QBG_E_INVAL;
    if (!g->f4[0] || !g->f4[1])
      return g->ctx->fail(QBG_E_INVAL, "grid_plan: the grid runs the generic kernels");
    const int pi = prec == QBG_PREC_FP64 ? 0 : 1;
    auto pick = [](int n, int n1, const qbg::F4Ops* dflt) -> const qbg::F4Ops* {
      if (n1 == 0) return dflt;
      const qbg::F4Ops* c[8];
      const int k = qbg::f4_candidates(n, c, 8);
      for (int i = 0; i < k; i++)
        if (c[i]->n1 == n1) return c[i];
      return nullptr;
    };
    const qbg::F4Ops* fx = pick(g->np0, x_n1, g->f4[0]);
    const qbg::F4Ops* fy = pick(g->np1, y_n1, g->f4[1]);
    if (!fx || !fy) return g->ctx->fail(QBG_E_INVAL, "grid_plan: no compiled split with that n1");
    g->fx[pi] = fx;
    g->fy[pi] = fy;
    g->yppb_t[pi] = yppb;
    g->tuned[pi] = true;
    return QBG_OK;
  });
}

extern "C" int qbg_fft_splits(int n, int* n1, int cap) {
  return qbg::guarded(nullptr, [&]() -> int {
    if (!n1 || cap < 0) return QBG_E_INVAL;
    const qbg::F4Ops* c[8];
    const int k = qbg::f4_candidates(n, c, 8);
    for (int i = 0; i < k && i < cap; i++) n1[i] = c[i]->n1;
    return k;
  });
}

// grid.cpp -- context, device memory and the FourierTransform drop-in
// (grid creation from Qbox's maps, batched transforms, density and
// rs_mul_add entry points).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include "internal.h"

using namespace qbg;

namespace qbg {

// radices applied in order 4,...,4,2,3,...,5,7,11
FftPlan make_plan(int n) {
  FftPlan p{};
  p.n = n;
  p.nstage = 0;
  int r = n;
  const int rads[] = {4, 2, 3, 5, 7, 11};
  for (int f : rads)
    while (r % f == 0 && p.nstage < 12) {
      p.radix[p.nstage++] = f;
      r /= f;
    }
  if (r != 1) p.nstage = -1;
  return p;
}

}  // namespace qbg

hipEvent_t qbg_ctx::prof_event() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void qbg_ctx::prof_begin(int kind) {
  if (!prof || capturing) return;
  Rec r{kind, prof_event(), prof_event()};
  (void)hipEventRecord(r.a, stream);
  recs.push_back(r);
}

void qbg_ctx::prof_end() {
  if (!prof || capturing || recs.empty()) return;
  (void)hipEventRecord(recs.back().b, stream);
  if (recs.size() > 4096) prof_collect();
}

void qbg_ctx::prof_collect() {
  if (recs.empty()) return;
  (void)hipEventSynchronize(recs.back().b);
  for (auto& r : recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      prof_ms[r.kind] += ms;
      prof_n[r.kind] += 1;
    }
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  recs.clear();
}

extern "C" {

int qbg_prof_enable(qbg_ctx* c, int on) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    c->prof_collect();
    c->prof = on != 0;
    return QBG_OK;
  });
}

int qbg_prof_reset(qbg_ctx* c) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    c->prof_collect();
    for (int k = 0; k < QBG_K_COUNT; k++) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
    return QBG_OK;
  });
}

int qbg_prof_read(qbg_ctx* c, int kind, double* ms, int* n) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || kind < 0 || kind >= QBG_K_COUNT) return QBG_E_INVAL;
    c->prof_collect();
    if (ms) *ms = c->prof_ms[kind];
    if (n) *n = c->prof_n[kind];
    return QBG_OK;
  });
}

int qbg_init(int device, qbg_ctx** out) {
  return qbg::guarded(nullptr, [&]() -> int {
    if (!out) return QBG_E_INVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return QBG_E_HIP;
    qbg_ctx* c = new (std::nothrow) qbg_ctx;
    if (!c) return QBG_E_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
      delete c;
      return QBG_E_HIP;
    }
    *out = c;
    return QBG_OK;
  });
}

int qbg_finalize(qbg_ctx* c) {
  if (!c) return QBG_E_INVAL;
  if (hipSetDevice(c->device) != hipSuccess) return QBG_E_HIP;
  (void)hipStreamSynchronize(c->stream);
  try {
    c->prof_collect();
  } catch (...) {
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  if (c->join_ev) (void)hipEventDestroy(c->join_ev);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return QBG_OK;
}

const char* qbg_last_error(const qbg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int qbg_sync(qbg_ctx* c) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "sync during graph capture");
    QBG_HIP(c, hipStreamSynchronize(c->stream));
    return QBG_OK;
  });
}

int qbg_stream_join(qbg_ctx* before, qbg_ctx* after) {
  return qbg::guarded(after, [&]() -> int {
    if (!before || !after || before == after) return QBG_E_INVAL;
    if (before->device != after->device)
      return after->fail(QBG_E_INVAL, "stream_join: contexts on different devices");
    QBG_HIP(after, hipSetDevice(after->device));
    if (!before->join_ev)
      QBG_HIP(after, hipEventCreateWithFlags(&before->join_ev, hipEventDisableTiming));
    QBG_HIP(after, hipEventRecord(before->join_ev, before->stream));
    QBG_HIP(after, hipStreamWaitEvent(after->stream, before->join_ev, 0));
    if (before->capturing && !after->capturing) {
      // after's stream is now part of before's capture
      after->capturing = true;
      after->capture_origin = before;
    } else if (after->capturing && before->capture_origin == after) {
      // a branch joined back into the capturing stream
      before->capturing = false;
      before->capture_origin = nullptr;
    }
    return QBG_OK;
  });
}

int qbg_graph_begin(qbg_ctx* c) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "graph capture already open");
    QBG_HIP(c, hipSetDevice(c->device));
    c->prof_collect();
    QBG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    return QBG_OK;
  });
}

int qbg_graph_end(qbg_ctx* c, qbg_graph** out) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !out) return QBG_E_INVAL;
    *out = nullptr;
    if (!c->capturing || c->capture_origin)
      return c->fail(QBG_E_STATE, "no graph capture open on this context");
    c->capturing = false;
    hipGraph_t gr = nullptr;
    QBG_HIP(c, hipStreamEndCapture(c->stream, &gr));
    qbg_graph* g = new (std::nothrow) qbg_graph;
    if (!g) { (void)hipGraphDestroy(gr); return QBG_E_NOMEM; }
    g->ctx = c;
    g->graph = gr;
    if (hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(gr);
      delete g;
      return c->fail(QBG_E_HIP, "hipGraphInstantiate failed");
    }
    *out = g;
    return QBG_OK;
  });
}

int qbg_graph_launch(qbg_graph* g) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    if (!g) return QBG_E_INVAL;
    QBG_HIP(g->ctx, hipGraphLaunch(g->exec, g->ctx->stream));
    return QBG_OK;
  });
}

int qbg_graph_destroy(qbg_graph* g) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    if (!g) return QBG_E_INVAL;
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return QBG_OK;
  });
}

int qbg_malloc(qbg_ctx* c, size_t bytes, void** p) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !p) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "not allowed during graph capture");
    QBG_HIP(c, hipSetDevice(c->device));
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
      *p = nullptr;
      return c->fail(QBG_E_NOMEM, "hipMalloc failed");
    }
    return QBG_OK;
  });
}

int qbg_free(qbg_ctx* c, void* p) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "not allowed during graph capture");
    QBG_HIP(c, hipStreamSynchronize(c->stream));
    QBG_HIP(c, hipFree(p));
    return QBG_OK;
  });
}

int qbg_memcpy_h2d(qbg_ctx* c, void* d, const void* h, size_t n) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "not allowed during graph capture");
    QBG_HIP(c, hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, c->stream));
    QBG_HIP(c, hipStreamSynchronize(c->stream));
    return QBG_OK;
  });
}

int qbg_memcpy_d2h(qbg_ctx* c, void* h, const void* d, size_t n) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "not allowed during graph capture");
    QBG_HIP(c, hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, c->stream));
    QBG_HIP(c, hipStreamSynchronize(c->stream));
    return QBG_OK;
  });
}

int qbg_memset(qbg_ctx* c, void* d, int v, size_t n) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "not allowed during graph capture");
    QBG_HIP(c, hipMemsetAsync(d, v, n, c->stream));
    return QBG_OK;
  });
}

int qbg_timer_start(qbg_ctx* c) {
  return qbg::guarded(c, [&]() -> int {
    if (!c) return QBG_E_INVAL;
    QBG_HIP(c, hipEventRecord(c->ev0, c->stream));
    return QBG_OK;
  });
}

int qbg_timer_stop(qbg_ctx* c, float* ms) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !ms) return QBG_E_INVAL;
    if (c->capturing) return c->fail(QBG_E_STATE, "not allowed during graph capture");
    QBG_HIP(c, hipEventRecord(c->ev1, c->stream));
    QBG_HIP(c, hipEventSynchronize(c->ev1));
    QBG_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
    return QBG_OK;
  });
}

}  // extern "C"

static void grid_free(qbg_grid* g) {
  void* ptrs[] = {g->d_ip, g->d_im, g->d_glist, g->d_tiles, g->d_row_start, g->d_rc_ivec, g->d_rc_x,
                  g->d_act_row, g->d_colmap, g->d_jrow, g->d_tw64[0], g->d_tw64[1], g->d_tw64[2], g->d_tw32[0],
                  g->d_tw32[1], g->d_tw32[2], g->d_zbuf, g->d_xbuf, g->d_fac, g->d_xres};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

template <typename T>
static int upload(qbg_ctx* c, T** d, const std::vector<T>& h) {
  if (hipMalloc((void**)d, sizeof(T) * std::max<size_t>(1, h.size())) != hipSuccess)
    return c->fail(QBG_E_NOMEM, "hipMalloc failed (grid tables)");
  if (!h.empty() &&
      hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    return c->fail(QBG_E_HIP, "hipMemcpy failed (grid tables)");
  return QBG_OK;
}

extern "C" {

int qbg_grid_create(qbg_ctx* c, int np0, int np1, int np2, int ng, int nvec,
                    const int32_t* ip, const int32_t* im, const int32_t* z2v, qbg_grid** out) {
  return qbg::guarded(c, [&]() -> int {
    if (!c || !out || !ip || !im || !z2v) return QBG_E_INVAL;
    *out = nullptr;
    if (np0 <= 0 || np1 <= 0 || np2 <= 0 || ng <= 0 || nvec <= 0 || nvec % 2 == 0)
      return c->fail(QBG_E_INVAL, "grid: bad dimensions (real basis needs odd nvec)");
    FftPlan p0 = make_plan(np0), p1 = make_plan(np1), p2 = make_plan(np2);
    if (p0.nstage < 0 || p1.nstage < 0 || p2.nstage < 0)
      return c->fail(QBG_E_INVAL, "grid: FFT length not a product of 2,3,5,7,11");
    const long nslot = (long)nvec * np2;
    // structure of the real-basis maps (src/BasisMapping.cpp:103-152):
    // rod 0 on column 0, rod r>0 on columns 2r-1 (+G) and 2r (-G), rods in G order
    std::vector<int> gfirst((nvec + 1) / 2, -1);
    int prev_grp = 0;
    for (int ig = 0; ig < ng; ig++) {
      if (ip[ig] < 0 || ip[ig] >= nslot || im[ig] < 0 || im[ig] >= nslot)
        return c->fail(QBG_E_INVAL, "grid: map entry out of range");
      const int cp = ip[ig] / np2, cm = im[ig] / np2;
      int grp;
      if (cp == 0 && cm == 0) grp = 0;
      else if (cp % 2 == 1 && cm == cp + 1) grp = (cp + 1) / 2;
      else return c->fail(QBG_E_INVAL, "grid: maps do not have the real-basis rod structure");
      if (grp < prev_grp) return c->fail(QBG_E_INVAL, "grid: G vectors not grouped by rod");
      if (gfirst[grp] < 0) gfirst[grp] = ig;
      prev_grp = grp;
    }
    if (gfirst[0] != 0) return c->fail(QBG_E_INVAL, "grid: G=0 rod must come first");
    for (size_t r = 0; r < gfirst.size(); r++)
      if (gfirst[r] < 0) return c->fail(QBG_E_INVAL, "grid: empty rod");
    // slots must be distinct except +G/-G of G=0 (second write wins)
    {
      std::vector<unsigned char> used(nslot, 0);
      for (int ig = 0; ig < ng; ig++) {
        if (used[ip[ig]]) return c->fail(QBG_E_INVAL, "grid: duplicate map slot");
        used[ip[ig]] = 1;
        if (im[ig] != ip[ig]) {
          if (used[im[ig]]) return c->fail(QBG_E_INVAL, "grid: duplicate map slot");
          used[im[ig]] = 1;
        }
      }
    }
    qbg_grid* g = new (std::nothrow) qbg_grid;
    if (!g) return QBG_E_NOMEM;
    g->ctx = c;
    g->np0 = np0; g->np1 = np1; g->np2 = np2; g->ng = ng; g->nvec = nvec;
    g->plan0 = p0; g->plan1 = p1; g->plan2 = p2;
    // LDS: two buffers of (lines x n) complex doubles within 64 KiB
    auto lines = [](int n) { return std::max(1, std::min(8, 65536 / (2 * 16 * n))); };
    g->zcols = lines(np2);
    // four-step kernels where compiled for the axis length (fft4_*.hip);
    // their z tile: zcols*N + N complex doubles of LDS, zcols*max(n1,n2) threads
    // (QBG_GENERIC_FFT=1 selects the generic Stockham kernels everywhere)
    const char* gen = std::getenv("QBG_GENERIC_FFT");
    const bool use_f4 = !(gen && gen[0] == '1');
    g->f4[0] = use_f4 ? qbg::f4_ops(np0) : nullptr;
    g->f4[1] = use_f4 ? qbg::f4_ops(np1) : nullptr;
    g->f4[2] = use_f4 ? qbg::f4_ops(np2) : nullptr;
    // the four-step kernels share blocked workspace layouts: all axes or none
    if (!g->f4[0] || !g->f4[1] || !g->f4[2]) g->f4[0] = g->f4[1] = g->f4[2] = nullptr;
    if (g->f4[2]) {
      const int m = std::max(g->f4[2]->n1, g->f4[2]->n2);
      g->zcols = std::min(256 / m, 4096 / np2 - 1);
    }
    // launch-shape override for tuning sweeps: QBG_YPPB (y-pass pairs per
    // block of the v multiply)
    if (const char* e = std::getenv("QBG_YPPB")) {
      g->yppb = std::max(1, std::atoi(e));
      g->yppb_fixed = true;
    }
    g->xrows = lines(np0);
    g->ytile = lines(np1);
    if (np1 * g->ytile > 256 * 16) g->ytile = std::max(1, 256 * 16 / np1);
    if (g->zcols < 3) { delete g; return c->fail(QBG_E_INVAL, "grid: np2 too large"); }
    // active rows: y in [0, ntrans0) and [np1 - ntrans0, np1)
    int nt0 = 0;
    for (int iv = 0; iv < nvec; iv++) {
      const int xy = z2v[iv];
      if (xy < 0 || xy >= np0 * np1) { delete g; return c->fail(QBG_E_INVAL, "grid: zvec_to_val out of range"); }
      const int y = xy / np0;
      const int k = y == 0 ? 0 : std::min(y, np1 - y);
      nt0 = std::max(nt0, k + 1);
    }
    if (2 * nt0 > np1) { delete g; return c->fail(QBG_E_INVAL, "grid: basis does not fit in grid (y)"); }
    g->ntrans0 = nt0;
    g->nact = 2 * nt0;
    std::vector<int32_t> act_row(g->nact), jrow(np1, -1);
    for (int r = 0; r < g->nact; r++) {
      act_row[r] = r < nt0 ? r : np1 - 2 * nt0 + r;
      jrow[act_row[r]] = r;
    }
    // Column numbering of the pass workspace.  Qbox numbers columns by rod
    // (rods sorted by size); here the rods are ranked by the (row, x) of
    // their +G column, rod of rank i owns column i (+G) and column nrods + i
    // (-G; unused for rod (0,0)).  The columns of one grid row then form two
    // runs in x order (ascending +G, descending -G), so the x passes read and
    // write coalesced; a z tile is a run of ranks [t0, t1) and owns columns
    // [t0, t1) and [nrods + t0, nrods + t1), its G vectors listed in glist.
    const int nrods = (nvec + 1) / 2;
    g->ncols = 2 * nrods;
    std::vector<int> rank(nrods), byrank(nrods);
    {
      auto pcol = [&](int r) { return r == 0 ? 0 : 2 * r - 1; };
      for (int r = 0; r < nrods; r++) byrank[r] = r;
      std::stable_sort(byrank.begin(), byrank.end(), [&](int a, int b) {
        const int ya = jrow[z2v[pcol(a)] / np0], yb = jrow[z2v[pcol(b)] / np0];
        if (ya != yb) return ya < yb;
        return z2v[pcol(a)] % np0 < z2v[pcol(b)] % np0;
      });
      for (int i = 0; i < nrods; i++) rank[byrank[i]] = i;
    }
    auto qcol = [&](int iv) {   // workspace column of Qbox column iv
      if (iv == 0) return rank[0];
      const int r = (iv + 1) / 2;
      return iv % 2 == 1 ? rank[r] : nrods + rank[r];
    };
    auto gend = [&](int r) { return r + 1 < nrods ? gfirst[r + 1] : ng; };
    std::vector<qbg_tile> tiles;
    std::vector<int32_t> glist;
    glist.reserve(ng);
    std::vector<int32_t> vip(ng), vim(ng);
    {
      const int rmax = std::max(1, g->zcols / 2);
      for (int t0 = 0; t0 < nrods; t0 += rmax) {
        const int t1 = std::min(nrods, t0 + rmax), nr = t1 - t0;
        qbg_tile t;
        t.col0 = t0;
        t.ncol = 2 * nr;
        t.np = nr;
        t.colm = nrods + t0;
        t.g0 = (int)glist.size();
        for (int i = t0; i < t1; i++) {
          const int r = byrank[i];
          for (int ig = gfirst[r]; ig < gend(r); ig++) {
            glist.push_back(ig);
            // tile-local slots offset by col0: local column (i - t0) for +G,
            // nr + (i - t0) for -G (rod (0,0): both in its +G column)
            vip[ig] = i * np2 + ip[ig] % np2;
            vim[ig] = (r == 0 ? i : nr + i) * np2 + im[ig] % np2;
          }
        }
        t.g1 = (int)glist.size();
        tiles.push_back(t);
      }
    }
    g->ntiles = (int)tiles.size();
    std::vector<int32_t> row_start(g->nact + 1, 0), rc_ivec(nvec), rc_x(nvec);
    {
      std::vector<int> rowof(nvec);
      for (int iv = 0; iv < nvec; iv++) {
        const int y = z2v[iv] / np0;
        rowof[iv] = jrow[y];
        row_start[rowof[iv] + 1]++;
      }
      for (int r = 0; r < g->nact; r++) row_start[r + 1] += row_start[r];
      std::vector<int> order(nvec);
      for (int iv = 0; iv < nvec; iv++) order[iv] = iv;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (rowof[a] != rowof[b]) return rowof[a] < rowof[b];
        return z2v[a] % np0 < z2v[b] % np0;
      });
      for (int e = 0; e < nvec; e++) {
        rc_ivec[e] = qcol(order[e]);
        rc_x[e] = z2v[order[e]] % np0;
      }
      // columns map to distinct grid points
      for (int e = 1; e < nvec; e++)
        if (rowof[order[e]] == rowof[order[e - 1]] && rc_x[e] == rc_x[e - 1]) {
          delete g;
          return c->fail(QBG_E_INVAL, "grid: two columns on one grid point");
        }
    }
    // dense (row, x) -> workspace column map for the four-step x passes
    std::vector<int32_t> colmap((size_t)g->nact * np0, -1);
    for (int iv = 0; iv < nvec; iv++) {
      const int x = z2v[iv] % np0, y = z2v[iv] / np0;
      colmap[(size_t)jrow[y] * np0 + x] = qcol(iv);
    }
    QBG_HIP(c, hipSetDevice(c->device));
    int rc = QBG_OK;
    if ((rc = upload(c, &g->d_ip, vip)) || (rc = upload(c, &g->d_im, vim)) ||
        (rc = upload(c, &g->d_glist, glist)) ||
        (rc = upload(c, &g->d_tiles, tiles)) || (rc = upload(c, &g->d_row_start, row_start)) ||
        (rc = upload(c, &g->d_rc_ivec, rc_ivec)) || (rc = upload(c, &g->d_rc_x, rc_x)) ||
        (rc = upload(c, &g->d_act_row, act_row)) || (rc = upload(c, &g->d_colmap, colmap)) ||
        (rc = upload(c, &g->d_jrow, jrow))) {
      grid_free(g);
      delete g;
      return rc;
    }
    const int ns[3] = {np0, np1, np2};
    for (int a = 0; a < 3; a++) {
      const int n = ns[a];
      std::vector<double> t64(2 * n);
      std::vector<float> t32(2 * n);
      const long double pi = 3.14159265358979323846264338327950288L;
      for (int k = 0; k < n; k++) {
        const long double ang = -2.0L * pi * k / n;
        t64[2 * k] = (double)cosl(ang);
        t64[2 * k + 1] = (double)sinl(ang);
        t32[2 * k] = (float)cosl(ang);
        t32[2 * k + 1] = (float)sinl(ang);
      }
      if ((rc = upload(c, &g->d_tw64[a], t64)) || (rc = upload(c, &g->d_tw32[a], t32))) {
        grid_free(g);
        delete g;
        return rc;
      }
    }
    *out = g;
    return QBG_OK;
  });
}

int qbg_grid_info(const qbg_grid* g, int* out) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    if (!g || !out) return QBG_E_INVAL;
    out[0] = g->np0; out[1] = g->np1; out[2] = g->np2;
    out[3] = g->ng; out[4] = g->nvec; out[5] = g->ntrans0;
    out[6] = g->f4[0] ? 1 : 0;
    for (int a = 0; a < 3; a++) {
      out[7 + 2 * a] = g->f4[a] ? g->f4[a]->n1 : 0;
      out[8 + 2 * a] = g->f4[a] ? g->f4[a]->n2 : 0;
    }
    return QBG_OK;
  });
}

int qbg_grid_destroy(qbg_grid* g) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    if (!g) return QBG_E_INVAL;
    QBG_HIP(g->ctx, hipSetDevice(g->ctx->device));
    (void)hipStreamSynchronize(g->ctx->stream);
    grid_free(g);
    delete g;
    return QBG_OK;
  });
}

}  // extern "C"

// workspace for up to npairs pairs (sized for FP64; FP32 uses half).  The
// pass launchers refuse to run a batch larger than g->ws_pairs.
int qbg::grid_ensure_ws(qbg_grid* g, int npairs) {
  if (npairs <= g->ws_pairs) return QBG_OK;
  if (g->ctx->capturing)
    return g->ctx->fail(QBG_E_STATE, "FFT workspace must exist before graph capture "
                                     "(run the call once uncaptured)");
  // sized for FP64 and for the blocked layouts of the four-step kernels
  // (planes and x padded to multiples of 16)
  const size_t np2p = (size_t)(g->np2 + 15) / 16 * 16, np0p = (size_t)(g->np0 + 15) / 16 * 16;
  const size_t zb = (size_t)npairs * g->ncols * np2p * 16;
  const size_t xb = (size_t)npairs * g->nact * np0p * g->np2 * 16;
  qbg_ctx* c = g->ctx;
  (void)hipStreamSynchronize(c->stream);
  if (g->d_zbuf) (void)hipFree(g->d_zbuf);
  if (g->d_xbuf) (void)hipFree(g->d_xbuf);
  g->d_zbuf = g->d_xbuf = nullptr;
  g->ws_bytes = 0;
  g->ws_pairs = 0;
  g->xws = nullptr;
  if (hipMalloc(&g->d_zbuf, zb) != hipSuccess || hipMalloc(&g->d_xbuf, xb) != hipSuccess)
    return c->fail(QBG_E_NOMEM, "hipMalloc failed (FFT workspace)");
  g->ws_bytes = zb + xb;
  g->ws_pairs = npairs;
  g->xws = g->d_xbuf;
  return QBG_OK;
}

// one pair's x workspace: nact rows x np0 (padded to the x blocks) x np2
// planes; covers the four-step layout and the generic one (np0 unpadded)
size_t qbg::xpair_bytes(const qbg_grid* g, int prec) {
  const size_t np0p = (size_t)(g->np0 + 15) / 16 * 16;
  return (size_t)g->nact * np0p * g->np2 * (prec == QBG_PREC_FP64 ? 16 : 8);
}

bool qbg::grid_ensure_xres(qbg_grid* g, int npairs) {
  // sized for FP64 whatever the precision asked for now, so that a later
  // FP64 evaluation does not reallocate under graphs captured earlier
  const size_t need = (size_t)npairs * xpair_bytes(g, QBG_PREC_FP64);
  if (need <= g->xres_bytes) return true;
  if (g->ctx->capturing) return false;
  (void)hipStreamSynchronize(g->ctx->stream);
  if (g->d_xres) (void)hipFree(g->d_xres);
  g->d_xres = nullptr;
  g->xres_bytes = 0;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || need > free_b / 2) return false;
  if (hipMalloc(&g->d_xres, need) != hipSuccess) {
    (void)hipGetLastError();
    g->d_xres = nullptr;
    return false;
  }
  g->xres_bytes = need;
  return true;
}
static int ensure_ws(qbg_grid* g, int npairs) { return qbg::grid_ensure_ws(g, npairs); }

static int ensure_fac(qbg_grid* g, int n) {
  if (n <= g->fac_cap) return QBG_OK;
  if (g->ctx->capturing) return g->ctx->fail(QBG_E_STATE, "allocation during graph capture");
  (void)hipStreamSynchronize(g->ctx->stream);
  if (g->d_fac) (void)hipFree(g->d_fac);
  g->d_fac = nullptr;
  g->fac_cap = 0;
  if (hipMalloc((void**)&g->d_fac, sizeof(double) * n) != hipSuccess)
    return g->ctx->fail(QBG_E_NOMEM, "hipMalloc failed (factors)");
  g->fac_cap = n;
  return QBG_OK;
}

#define TRY(x)                  \
  do {                          \
    int r_ = (x);               \
    if (r_ != QBG_OK) return r_; \
  } while (0)

#define PROF(kind, call)                                                   \
  do {                                                                     \
    g->ctx->prof_begin(kind);                                              \
    const int r_ = (call);                                                 \
    g->ctx->prof_end();                                                    \
    if (r_ != QBG_OK) return g->ctx->fail(r_, "FFT pass launch refused"); \
  } while (0)

static int checkprec(qbg_grid* g, int prec) {
  if (!g) return QBG_E_INVAL;
  if (prec != QBG_PREC_FP64 && prec != QBG_PREC_FP32)
    return g->ctx->fail(QBG_E_INVAL, "unknown precision");
  QBG_HIP(g->ctx, hipSetDevice(g->ctx->device));
  return QBG_OK;
}

static int last_launch(qbg_grid* g) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return g->ctx->fail(QBG_E_HIP, hipGetErrorString(e));
  return QBG_OK;
}

// rho += sum over band pairs of fac[n] |psi_n(r)|^2, the pairs in batches of
// up to g->batch (src/SlaterDet.cpp:232-283); fac is a device array of
// 2 * ceil(nst / 2) factors, so the loop copies nothing from the host
namespace {
// points the pass launchers at a slice of the resident x array for one
// batch and back at the batch workspace when the scope ends (also on error)
struct XwsScope {
  qbg_grid* g;
  XwsScope(qbg_grid* g_, void* p) : g(g_) { if (p) g->xws = p; }
  ~XwsScope() { g->xws = g->d_xbuf; }
};
}  // namespace

int qbg::density_batches(qbg_grid* g, const double* c, int ldc, int nst, const double* fac,
                         double* rho, int prec, bool resident, int pair_base) {
  const int npairs = (nst + 1) / 2;
  const int B = std::min(npairs, g->batch);
  TRY(ensure_ws(g, B));
  TRY(qbg::grid_tune(g, prec));
  const size_t pb = xpair_bytes(g, prec);
  if (resident && g->xres_bytes < (size_t)(pair_base + npairs) * pb)
    return g->ctx->fail(QBG_E_STATE, "resident x array too small");
  hipStream_t s = g->ctx->stream;
  for (int p0 = 0; p0 < npairs; p0 += B) {
    const int nb = std::min(B, npairs - p0);
    const int band0 = 2 * p0;
    const int last_has2 = (band0 + 2 * nb) <= nst ? 1 : 0;
    XwsScope xs(g, resident ? (char*)g->d_xres + (size_t)(pair_base + p0) * pb : nullptr);
    PROF(QBG_K_ZBWD, zbwd(g, prec, Z_PAIR, c + 2L * ldc * band0, 2L * ldc, ldc, last_has2, nb, s));
    PROF(QBG_K_XBWD, xbwd(g, prec, nb, s));
    PROF(QBG_K_YDENS, ypass(g, prec, Y_DENSITY, nb, nullptr, fac + 2 * p0, rho, nullptr, 1.0, s));
  }
  return last_launch(g);
}

// the completion arguments of the batch starting at band0
static qbg::ZFin fin_batch(const qbg::ZFin* fin, int band0) {
  qbg::ZFin f = *fin;
  f.c = fin->c + fin->cpst * band0;
  f.y = fin->y + fin->ldy * band0;
  return f;
}

int qbg::hpsi_resident(qbg_grid* g, const double* v, int nst, double* hpsi, int ldh, int prec,
                       int pair_base, const ZFin* fin) {
  const int npairs = (nst + 1) / 2;
  const int B = std::min(npairs, g->batch);
  TRY(ensure_ws(g, B));
  TRY(qbg::grid_tune(g, prec));
  const size_t pb = xpair_bytes(g, prec);
  if (g->xres_bytes < (size_t)(pair_base + npairs) * pb)
    return g->ctx->fail(QBG_E_STATE, "resident x array too small");
  hipStream_t s = g->ctx->stream;
  for (int p0 = 0; p0 < npairs; p0 += B) {
    const int nb = std::min(B, npairs - p0);
    const int band0 = 2 * p0;
    const int last_has2 = (band0 + 2 * nb) <= nst ? 1 : 0;
    XwsScope xs(g, (char*)g->d_xres + (size_t)(pair_base + p0) * pb);
    PROF(QBG_K_YVMUL, ypass(g, prec, Y_VMUL, nb, v, nullptr, nullptr, nullptr, 1.0, s));
    PROF(QBG_K_XFWD, xfwd(g, prec, nb, s));
    const ZFin fb = fin ? fin_batch(fin, band0) : ZFin{};
    PROF(QBG_K_ZFWD, zfwd(g, prec, Z_PAIR, OUT_ACCUM, hpsi + 2L * ldh * band0, 2L * ldh, ldh,
                          last_has2, nb, s, fin ? &fb : nullptr));
  }
  return last_launch(g);
}

int qbg::rs_mul_add_bands(qbg_grid* g, const double* v, const double* c, int ldc, int nst,
                          double* hpsi, int ldh, int prec, const ZFin* fin) {
  const int npairs = (nst + 1) / 2;
  const int B = std::min(npairs, g->batch);
  TRY(ensure_ws(g, B));
  TRY(qbg::grid_tune(g, prec));
  hipStream_t s = g->ctx->stream;
  for (int p0 = 0; p0 < npairs; p0 += B) {
    const int nb = std::min(B, npairs - p0);
    const int band0 = 2 * p0;
    const int last_has2 = (band0 + 2 * nb) <= nst ? 1 : 0;
    PROF(QBG_K_ZBWD, zbwd(g, prec, Z_PAIR, c + 2L * ldc * band0, 2L * ldc, ldc, last_has2, nb, s));
    PROF(QBG_K_XBWD, xbwd(g, prec, nb, s));
    PROF(QBG_K_YVMUL, ypass(g, prec, Y_VMUL, nb, v, nullptr, nullptr, nullptr, 1.0, s));
    PROF(QBG_K_XFWD, xfwd(g, prec, nb, s));
    const ZFin fb = fin ? fin_batch(fin, band0) : ZFin{};
    PROF(QBG_K_ZFWD, zfwd(g, prec, Z_PAIR, OUT_ACCUM, hpsi + 2L * ldh * band0, 2L * ldh, ldh,
                          last_has2, nb, s, fin ? &fb : nullptr));
  }
  return last_launch(g);
}

extern "C" {

int qbg_fft_backward(qbg_grid* g, const double* c, double* f, int prec) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    TRY(checkprec(g, prec));
    TRY(ensure_ws(g, 1));
    hipStream_t s = g->ctx->stream;
    PROF(QBG_K_ZBWD, zbwd(g, prec, Z_SINGLE, c, 0, 0, 0, 1, s));
    PROF(QBG_K_XBWD, xbwd(g, prec, 1, s));
    PROF(QBG_K_YOTHER, ypass(g, prec, Y_BWD_FULL, 1, nullptr, nullptr, nullptr, f, 1.0, s));
    return last_launch(g);
  });
}

int qbg_fft_backward_pair(qbg_grid* g, const double* c1, const double* c2, double* f, int prec) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    TRY(checkprec(g, prec));
    TRY(ensure_ws(g, 1));
    hipStream_t s = g->ctx->stream;
    const long off2 = (long)(c2 - c1) / 2;
    if ((c2 - c1) % 2) return g->ctx->fail(QBG_E_INVAL, "misaligned band pointers");
    PROF(QBG_K_ZBWD, zbwd(g, prec, Z_PAIR, c1, 0, off2, 1, 1, s));
    PROF(QBG_K_XBWD, xbwd(g, prec, 1, s));
    PROF(QBG_K_YOTHER, ypass(g, prec, Y_BWD_FULL, 1, nullptr, nullptr, nullptr, f, 1.0, s));
    return last_launch(g);
  });
}

int qbg_fft_forward(qbg_grid* g, double* f, double* c, int prec) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    TRY(checkprec(g, prec));
    TRY(ensure_ws(g, 1));
    hipStream_t s = g->ctx->stream;
    PROF(QBG_K_YOTHER, ypass(g, prec, Y_FWD_FULL, 1, nullptr, nullptr, nullptr, f, 1.0, s));
    PROF(QBG_K_XFWD, xfwd(g, prec, 1, s));
    PROF(QBG_K_ZFWD, zfwd(g, prec, Z_SINGLE, OUT_STORE, c, 0, 0, 0, 1, s));
    return last_launch(g);
  });
}

int qbg_fft_forward_pair(qbg_grid* g, double* f, double* c1, double* c2, int prec) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    TRY(checkprec(g, prec));
    TRY(ensure_ws(g, 1));
    hipStream_t s = g->ctx->stream;
    if ((c2 - c1) % 2) return g->ctx->fail(QBG_E_INVAL, "misaligned band pointers");
    const long off2 = (long)(c2 - c1) / 2;
    PROF(QBG_K_YOTHER, ypass(g, prec, Y_FWD_FULL, 1, nullptr, nullptr, nullptr, f, 1.0, s));
    PROF(QBG_K_XFWD, xfwd(g, prec, 1, s));
    PROF(QBG_K_ZFWD, zfwd(g, prec, Z_PAIR, OUT_STORE, c1, 0, off2, 1, 1, s));
    return last_launch(g);
  });
}

int qbg_compute_density(qbg_grid* g, const double* c, int ldc, int nst, const double* occ,
                        double weight, double omega, double* rho, int prec) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    TRY(checkprec(g, prec));
    if (nst <= 0 || ldc < g->ng || !occ || !(omega > 0.0))
      return g->ctx->fail(QBG_E_INVAL, "compute_density: bad arguments");
    if (g->ctx->capturing)
      return g->ctx->fail(QBG_E_STATE, "compute_density copies host factors: not capturable");
    const int npairs = (nst + 1) / 2;
    TRY(ensure_fac(g, 2 * npairs));
    // fac = weight * occ / omega (src/SlaterDet.cpp:232-244); a missing second
    // band of the last pair contributes 0
    std::vector<double> fac(2 * npairs, 0.0);
    for (int n = 0; n < nst; n++) fac[n] = weight * (1.0 / omega) * occ[n];
    hipStream_t s = g->ctx->stream;
    QBG_HIP(g->ctx, hipMemcpyAsync(g->d_fac, fac.data(), sizeof(double) * fac.size(),
                                   hipMemcpyHostToDevice, s));
    TRY(qbg::density_batches(g, c, ldc, nst, g->d_fac, rho, prec));
    // the host factor array must outlive the async copy
    QBG_HIP(g->ctx, hipStreamSynchronize(s));
    return QBG_OK;
  });
}
int qbg_rs_mul_add(qbg_grid* g, const double* v, const double* c, int ldc, int nst,
                   double* hpsi, int ldh, int prec) {
  return qbg::guarded((g ? g->ctx : nullptr), [&]() -> int {
    TRY(checkprec(g, prec));
    if (nst <= 0 || ldc < g->ng || ldh < g->ng)
      return g->ctx->fail(QBG_E_INVAL, "rs_mul_add: bad arguments");
    return qbg::rs_mul_add_bands(g, v, c, ldc, nst, hpsi, ldh, prec);
  });
}

}  // extern "C"

// internal.h -- shared definitions of libqbg (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>
#include "qbg.h"

namespace qbg {

struct Error {
  int code;
  std::string msg;
};

// FFT factorization of one line length: radices applied in order
struct FftPlan {
  int n;
  int nstage;
  int radix[12];
};

FftPlan make_plan(int n);

template <typename T>
struct cpx {
  T x, y;
};

}  // namespace qbg

struct qbg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool capturing = false;           // between qbg_graph_begin and qbg_graph_end
  // the context whose capture this one's stream joined (qbg_stream_join)
  qbg_ctx* capture_origin = nullptr;
  hipEvent_t join_ev = nullptr;     // recorded by qbg_stream_join(this, ...)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // per-kernel event profiler
  bool prof = false;
  struct Rec { int kind; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double prof_ms[QBG_K_COUNT] = {0};
  int prof_n[QBG_K_COUNT] = {0};
  hipEvent_t prof_event();
  void prof_begin(int kind);
  void prof_end();
  void prof_collect();
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

// column tile of the z pass: ncol workspace columns, the first np of them
// at [col0, col0+np) (+G columns), the rest at [colm, colm+ncol-np) (-G
// columns); its G vectors are glist[g0, g1)
struct qbg_tile {
  int col0, ncol, g0, g1, np, colm;
};

namespace qbg {
// compiled four-step kernels of one line length n = n1*n2 (fft4_*.hip);
// L = lines per block of the x and y passes
// Completion of H*psi in the forward z pass (four-step kernels only): with
// c set, the pass stores out = 0.5 g2 c + y + (the transform) instead of
// adding the transform to out, so the kinetic and nonlocal parts need no
// earlier pass over H*psi.  c and y follow out's pair layout: the pair's
// first band at c + 2 pair cpst (its second at + 2 coff2) and at
// y + 2 pair ldy (second at + ldy), y the FP32 nonlocal product (real view,
// re/im interleaved), g2 the wavefunction basis |G|^2.
struct ZFin {
  const double* c = nullptr;
  long cpst = 0, coff2 = 0;
  const double* g2 = nullptr;
  const float* y = nullptr;
  long ldy = 0;
};

struct F4Ops {
  int n, n1, n2, L;
  void (*zbwd)(qbg_grid* g, int prec, int single, const double* c, long pstride, long off2,
               int last_has2, int nbatch, hipStream_t s);
  void (*zfwd)(qbg_grid* g, int prec, int single, int accum, double* out, long pstride,
               long off2, int last_has2, int nbatch, hipStream_t s, ZFin fin);
  void (*xpass)(qbg_grid* g, int prec, bool fwd, int nbatch, hipStream_t s);
  void (*ypass)(qbg_grid* g, int prec, int mode, int nbatch, int ppb, const double* v,
                const double* fac, double* rho, double* fgrid, double scale, hipStream_t s);
};
const F4Ops* f4_ops(int n);   // default split of n, nullptr if n has no compiled kernels
// every compiled split of n (default first), at most cap; returns the count
int f4_candidates(int n, const F4Ops** out, int cap);
}  // namespace qbg

struct qbg_grid {
  qbg_ctx* ctx = nullptr;
  int np0 = 0, np1 = 0, np2 = 0, ng = 0, nvec = 0;
  int ntrans0 = 0;     // active y rows: [0, ntrans0) and [np1-ntrans0, np1)
  int nact = 0;        // number of active rows (2*ntrans0, capped at np1)
  int ntiles = 0;
  int zcols = 8;       // columns per z tile
  int xrows = 8;       // rows per x-pass block
  int ytile = 8;       // x values per y-pass block
  qbg::FftPlan plan0, plan1, plan2;
  // device tables
  int ncols = 0;       // workspace columns per plane (2 * rods)
  int32_t* d_ip = nullptr;          // tile-local slots of +G (offset by col0*np2)
  int32_t* d_im = nullptr;          // tile-local slots of -G
  int32_t* d_glist = nullptr;       // G indices of the tiles, tile by tile
  qbg_tile* d_tiles = nullptr;
  int32_t* d_row_start = nullptr;   // nact+1
  int32_t* d_rc_ivec = nullptr;     // nvec, workspace columns sorted by (row, x)
  int32_t* d_rc_x = nullptr;        // nvec
  int32_t* d_act_row = nullptr;     // nact: y index of active row r
  int32_t* d_colmap = nullptr;      // nact*np0: workspace column at (row, x) or -1
  int32_t* d_jrow = nullptr;        // np1: active row of y or -1
  const qbg::F4Ops* f4[3] = {nullptr, nullptr, nullptr};   // per axis, or generic
  double* d_tw64[3] = {nullptr, nullptr, nullptr};   // exp(-2 pi i k / n)
  float* d_tw32[3] = {nullptr, nullptr, nullptr};
  // workspace for batched pairs (grown on demand)
  void* d_zbuf = nullptr;
  void* d_xbuf = nullptr;
  size_t ws_bytes = 0;
  int ws_pairs = 0;                 // pairs the workspace holds
  // the x workspace the pass launchers use: d_xbuf, or a batch's slice of
  // the resident array while a resident density or H*psi loop runs
  void* xws = nullptr;
  // resident x-pass output of every pair of a band set (grown on demand):
  // the backward z and x transforms of update_density, kept for the H*psi
  // of the same coefficients
  void* d_xres = nullptr;
  size_t xres_bytes = 0;
  double* d_fac = nullptr;          // per-band density factors
  int fac_cap = 0;
  int batch = 64;                   // max pairs per batch
  int yppb = 16;                     // y-pass pairs per block for the v multiply
  // the tuner's choices per precision (index 0 FP64, 1 FP32): the split of
  // the x-pass (axis 0) and y-pass (axis 1) kernels among the compiled ones
  // for the axis length, and the y-pass pairs per block of the v multiply;
  // until tuned, the defaults (f4[0], f4[1], yppb)
  const qbg::F4Ops* fx[2] = {nullptr, nullptr};
  const qbg::F4Ops* fy[2] = {nullptr, nullptr};
  int yppb_t[2] = {0, 0};
  bool tuned[2] = {false, false};
  bool yppb_fixed = false;           // set by QBG_YPPB: no tuning of yppb
};

// launchers (fft.hip)
namespace qbg {
enum ZMode { Z_PAIR = 0, Z_SINGLE = 1 };
enum YMode { Y_VMUL = 0, Y_DENSITY = 1, Y_BWD_FULL = 2, Y_FWD_FULL = 3,
             Y_BWD_REAL_ADD = 4, Y_FWD_REAL = 5 };
enum OutMode { OUT_STORE = 0, OUT_ACCUM = 1 };

// batched pass launchers (fft.hip).  Pair b of a batch reads/writes the
// bands at c + 2*b*pstride (first) and c + 2*(b*pstride + off2) (second,
// absent for the last pair when last_has2 == 0).
// allocate the pass workspace of g for npairs band pairs (no-op if it holds
// that many already); every launcher below returns QBG_E_STATE without
// launching when the workspace is missing or smaller than its batch.
int grid_ensure_ws(qbg_grid* g, int npairs);
// time the compiled x- and y-pass variants on the grid's workspace and keep
// the fastest for prec (once per grid and precision; not while capturing;
// the workspace must exist and holds no data across this call)
int grid_tune(qbg_grid* g, int prec);
// bytes of one pair's x workspace at precision prec (both pass families)
size_t xpair_bytes(const qbg_grid* g, int prec);
// allocate the resident x array for npairs pairs (sized for FP64); returns
// false (allocating nothing) when it would take more than half of the free
// device memory or during graph capture, and the caller keeps the
// recomputing path
bool grid_ensure_xres(qbg_grid* g, int npairs);
int zbwd(qbg_grid* g, int prec, int mode, const double* c, long pstride, long off2,
         int last_has2, int nbatch, hipStream_t s);
int xbwd(qbg_grid* g, int prec, int nbatch, hipStream_t s);
int xfwd(qbg_grid* g, int prec, int nbatch, hipStream_t s);
int ypass(qbg_grid* g, int prec, int mode, int nbatch, const double* v, const double* fac,
          double* rho, double* fgrid, double scale, hipStream_t s);
int zfwd(qbg_grid* g, int prec, int mode, int outmode, double* out, long pstride, long off2,
         int last_has2, int nbatch, hipStream_t s, const ZFin* fin = nullptr);
// whether zfwd can complete H*psi (ZFin) on this grid: four-step z kernels
bool zfwd_fin_ok(const qbg_grid* g);
// band-pair density passes with per-band factors already on the device
// (fac[2*npairs], zero for a missing last band); no host synchronization.
// resident: the x-pass output of every pair stays in g->d_xres, from pair
// slot pair_base on (the array must hold pair_base + (nst+1)/2 pairs at
// prec), for hpsi_resident
int density_batches(qbg_grid* g, const double* c, int ldc, int nst, const double* fac,
                    double* rho, int prec, bool resident = false, int pair_base = 0);
// hpsi += the local-potential part (rs_mul_add) of the bands whose x-pass
// output density_batches(..., resident = true, pair_base) left in g->d_xres;
// consumes it (the y pass transforms it in place)
int hpsi_resident(qbg_grid* g, const double* v, int nst, double* hpsi, int ldh, int prec,
                  int pair_base = 0, const ZFin* fin = nullptr);
// qbg_rs_mul_add, optionally completing H*psi (fin: c of the same bands,
// ldc, y, ldy, g2; then hpsi is stored, not added to)
int rs_mul_add_bands(qbg_grid* g, const double* v, const double* c, int ldc, int nst,
                     double* hpsi, int ldh, int prec, const ZFin* fin = nullptr);
// util.hip
void fill_zero(double* p, size_t n, hipStream_t s);
void fill_value(double* p, size_t n, double v, hipStream_t s);
void add2(const double* a, const double* b, double* out, size_t n, hipStream_t s);
void gather_segments(const double* const* seg_ptr, const int* seg_n, int nseg, double* dst,
                     hipStream_t s);
}  // namespace qbg

struct qbg_graph {
  qbg_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace qbg {
// Body of an exported entry point: no C++ exception crosses the C ABI.  A
// failed host allocation returns QBG_E_NOMEM, any other exception
// QBG_E_INVAL, each with its text in the context's last error (when there
// is a context to hold it).
template <class F>
int guarded(qbg_ctx* c, F&& body) noexcept {
  const char* what = nullptr;
  int code = QBG_E_INVAL;
  try {
    return body();
  } catch (const std::bad_alloc&) {
    code = QBG_E_NOMEM;
    what = "host allocation failed";
  } catch (const std::exception& e) {
    what = e.what();
  } catch (...) {
    what = "unknown exception";
  }
  if (c) {
    try {
      c->err = what;
    } catch (...) {
    }
  }
  return code;
}
}  // namespace qbg

#define QBG_HIP(ctx, call)                                                   \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return (ctx)->fail(QBG_E_HIP, std::string(#call) + ": " +             \
                                        hipGetErrorString(e_));              \
  } while (0)

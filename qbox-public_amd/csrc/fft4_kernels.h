// fft4_kernels.h -- the five pass kernels with four-step line transforms of
// compile-time length (N1 x N2), and their launchers.  Instantiated per line
// length in fft4_sizes.hip; dispatched from fft.hip through f4_ops().
//
// Workspace layouts (blocked so that every block moves long contiguous
// runs): zbuf[pair][kb][col][kk] holds plane k = KB*kb + kk of workspace
// column col; xbuf[pair][k][xb][r][xi] holds x = XB*xb + xi of active row r
// of plane k.  Real-space grids are i + np0*(j + np1*k).  Within a line of length N,
// phase A holds x[N2*n1 + n2] (thread n2) and phase B holds X[k1 + N1*k2]
// (thread k1); a forward transform starting from the phase-B layout
// (thread k1 holding y = k1 + N1*k2) ends in the phase-A layout.
#pragma once
#include "fft4.h"

namespace qbg {
namespace f4 {

constexpr int KB = 16;   // planes per zbuf block
constexpr int XB = 16;   // x per xbuf block
__host__ __device__ constexpr int nblk(int n, int b) { return (n + b - 1) / b; }
// threads of an x or y block: 16 lines of max(N1, N2) threads, whole waves
__host__ __device__ constexpr int xthreads(int n1, int n2) {
  return ((16 * (n1 > n2 ? n1 : n2) + 63) / 64) * 64;
}

// ------------------------------------------------------------------ z-bwd
template <typename T, int N1, int N2>
__global__ __launch_bounds__(256) void kz_bwd(const double* __restrict__ c, long pstride,
                                              long off2, int last_has2, int nbatch, int single,
                                              const int32_t* __restrict__ ip,
                                              const int32_t* __restrict__ im,
                                              const int32_t* __restrict__ glist,
                                              const qbg_tile* __restrict__ tiles, int ncols, int TC,
                                              const cpx<T>* __restrict__ tw_g,
                                              cpx<T>* __restrict__ zbuf) {
  constexpr int N = N1 * N2;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* tw = A + TC * N;
  const int tid = threadIdx.x;
  const qbg_tile t = tiles[blockIdx.x];
  const int pair = blockIdx.y;
  for (int i = tid; i < TC * N; i += 256) A[i] = {T(0), T(0)};
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  __syncthreads();
  // pack two real-basis bands (or one) onto the columns
  // (src/BasisMapping.cpp:628-700); -G written last so it wins at G=0
  const int slot0 = t.col0 * N;
  const double* c1 = c + 2 * (long)pair * pstride;
  const bool has2 = !single && (pair < nbatch - 1 || last_has2);
  const double* c2 = c1 + 2 * off2;
  for (int j = t.g0 + tid; j < t.g1; j += 256) {
    const int ig = glist[j];
    const double a = c1[2 * ig], b = c1[2 * ig + 1];
    double cc = 0.0, d = 0.0;
    if (has2) { cc = c2[2 * ig]; d = c2[2 * ig + 1]; }
    const int sp = ip[ig] - slot0, sm = im[ig] - slot0;
    A[(sp % N) * TC + sp / N] = {T(a - d), T(b + cc)};
    A[(sm % N) * TC + sm / N] = {T(a + d), T(cc - b)};
  }
  __syncthreads();
  const int col = tid % TC, q = tid / TC;
  using Mp = Map<N1, N2>;
  cpx<T> a[N1];
  const bool va = col < t.ncol && q < N2;
  if (va) {                               // phase A, q = n2
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = A[Mp::in_a(n1, q) * TC + col];
    rdft<T, N1, +1>(a);
    if constexpr (!Mp::pfa) {
#pragma unroll
      for (int k1 = 1; k1 < N1; k1++)
        if (q) a[k1] = mul(a[k1], twl<T, +1>(tw, q * k1));
    }
  }
  __syncthreads();                        // the prime-factor reads cross threads
  if (va) {
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) A[(k1 * N2 + q) * TC + col] = a[k1];
  }
  __syncthreads();
  cpx<T> b[N2];
  const bool vb = col < t.ncol && q < N1;
  if (vb) {                               // phase B, q = k1
#pragma unroll
    for (int n2 = 0; n2 < N2; n2++) b[n2] = A[(q * N2 + n2) * TC + col];
    rdft<T, N2, +1>(b);
  }
  __syncthreads();
  if (vb) {
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) A[Mp::out_b(q, k2) * TC + col] = b[k2];
  }
  __syncthreads();
  // store KB-plane chunks: the tile's +G and -G columns are two runs
  // lane = (column slot, plane kk): 16 slots x KB planes per pass
  // (all NKB chunks of a column unrolled: the LDS reads and the stores
  // issue back to back; the chunk past the line end is dropped)
  constexpr int NKB = nblk(N, KB);
  constexpr uint32_t SZ = sizeof(cpx<T>);
  const auto zr = buf_res(zbuf + (long)pair * NKB * ncols * KB, (uint32_t)(NKB * KB) * ncols * SZ);
  const int skk = tid % KB, scl = tid / KB;
  for (int cl = scl; cl < t.ncol; cl += 256 / KB) {
    const int gcol = cl < t.np ? t.col0 + cl : t.colm + (cl - t.np);
    const uint32_t cb = (uint32_t)(gcol * KB + skk) * SZ;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
      const int z = kb * KB + skk;
      const cpx<T> v = A[(z < N ? z : 0) * TC + cl];
      buf_store<T>(zr, z < N ? cb + (uint32_t)(kb * ncols * KB) * SZ : OOB, v);
    }
  }
}

// ------------------------------------------------------------------ z-fwd
template <typename T, int N1, int N2>
__global__ __launch_bounds__(256) void kz_fwd(const cpx<T>* __restrict__ zbuf,
                                              double* __restrict__ out, long pstride, long off2,
                                              int last_has2, int nbatch, int single, int accum,
                                              const int32_t* __restrict__ ip,
                                              const int32_t* __restrict__ im,
                                              const int32_t* __restrict__ glist,
                                              const qbg_tile* __restrict__ tiles, int ncols, int TC,
                                              const cpx<T>* __restrict__ tw_g, double scale,
                                              ZFin fin) {
  constexpr int N = N1 * N2;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* tw = A + TC * N;
  const int tid = threadIdx.x;
  const qbg_tile t = tiles[blockIdx.x];
  const int pair = blockIdx.y;
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  __syncthreads();
  // load KB-plane chunks of the tile's columns
  // (all NKB chunks of a column in flight before the LDS writes)
  constexpr int NKB = nblk(N, KB);
  constexpr uint32_t SZ = sizeof(cpx<T>);
  {
    const auto zr = buf_res(zbuf + (long)pair * NKB * ncols * KB,
                            (uint32_t)(NKB * KB) * ncols * SZ);
    const int skk = tid % KB, scl = tid / KB;
    for (int cl = scl; cl < t.ncol; cl += 256 / KB) {
      const int gcol = cl < t.np ? t.col0 + cl : t.colm + (cl - t.np);
      const uint32_t cb = (uint32_t)(gcol * KB + skk) * SZ;
      cpx<T> st[NKB];
#pragma unroll
      for (int kb = 0; kb < NKB; kb++) {
        const int z = kb * KB + skk;
        st[kb] = buf_load<T>(zr, z < N ? cb + (uint32_t)(kb * ncols * KB) * SZ : OOB);
      }
#pragma unroll
      for (int kb = 0; kb < NKB; kb++) {
        const int z = kb * KB + skk;
        if (z < N) A[z * TC + cl] = st[kb];
      }
    }
  }
  __syncthreads();
  const int col = tid % TC, q = tid / TC;
  using Mp = Map<N1, N2>;
  cpx<T> b[N2];
  const bool va = col < t.ncol && q < N1;
  if (va) {                               // forward phase A', q = k1
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) b[k2] = A[Mp::out_b(q, k2) * TC + col];
    rdft<T, N2, -1>(b);
    if constexpr (!Mp::pfa) {
#pragma unroll
      for (int m1 = 1; m1 < N2; m1++)
        if (q) b[m1] = mul(b[m1], twl<T, -1>(tw, q * m1));
    }
  }
  __syncthreads();
  if (va) {
#pragma unroll
    for (int m1 = 0; m1 < N2; m1++) A[(m1 * N1 + q) * TC + col] = b[m1];
  }
  __syncthreads();
  cpx<T> cv[N1];
  const bool vb = col < t.ncol && q < N2;
  if (vb) {                               // phase B', q = m1
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) cv[k1] = A[(q * N1 + k1) * TC + col];
    rdft<T, N1, -1>(cv);
  }
  __syncthreads();
  if (vb) {
#pragma unroll
    for (int m2 = 0; m2 < N1; m2++) A[Mp::in_a(m2, q) * TC + col] = cv[m2];
  }
  __syncthreads();
  // scale by 1/np012 (fz, src/FourierTransform.cpp:739-743) and unpack
  // (zvec_to_doublevector / zvec_to_vector, src/BasisMapping.cpp:703-750)
  const int slot0 = t.col0 * N;
  double* o1 = out + 2 * (long)pair * pstride;
  double* o2 = o1 + 2 * off2;
  const bool has2 = !single && (pair < nbatch - 1 || last_has2);
  // completion of H*psi: 0.5 g2 c + y of the pair's bands
  const double* c1 = fin.c ? fin.c + 2 * (long)pair * fin.cpst : nullptr;
  const double* c2 = fin.c ? c1 + 2 * fin.coff2 : nullptr;
  const float* y1 = fin.c ? fin.y + 2 * (long)pair * fin.ldy : nullptr;
  const float* y2 = fin.c ? y1 + fin.ldy : nullptr;
  for (int j = t.g0 + tid; j < t.g1; j += 256) {
    const int ig = glist[j];
    const int sp = ip[ig] - slot0;
    const cpx<T> zpv = A[(sp % N) * TC + sp / N];
    const double a = (double)zpv.x * scale, b = (double)zpv.y * scale;
    if (fin.c) {
      const double kg = 0.5 * fin.g2[ig];
      const double r1 = kg * c1[2 * ig] + (double)y1[2 * ig];
      const double i1 = kg * c1[2 * ig + 1] + (double)y1[2 * ig + 1];
      if (has2) {
        const int sm = im[ig] - slot0;
        const cpx<T> zmv = A[(sm % N) * TC + sm / N];
        const double cc = (double)zmv.x * scale, d = (double)zmv.y * scale;
        o1[2 * ig] = r1 + 0.5 * (a + cc);
        o1[2 * ig + 1] = i1 + 0.5 * (b - d);
        o2[2 * ig] = (kg * c2[2 * ig] + (double)y2[2 * ig]) + 0.5 * (b + d);
        o2[2 * ig + 1] = (kg * c2[2 * ig + 1] + (double)y2[2 * ig + 1]) + 0.5 * (cc - a);
      } else {
        o1[2 * ig] = r1 + a;
        o1[2 * ig + 1] = i1 + b;
      }
      continue;
    }
    if (has2) {
      const int sm = im[ig] - slot0;
      const cpx<T> zmv = A[(sm % N) * TC + sm / N];
      const double cc = (double)zmv.x * scale, d = (double)zmv.y * scale;
      const double r1 = 0.5 * (a + cc), i1 = 0.5 * (b - d);
      const double r2 = 0.5 * (b + d), i2 = 0.5 * (cc - a);
      if (accum) {
        o1[2 * ig] += r1; o1[2 * ig + 1] += i1;
        o2[2 * ig] += r2; o2[2 * ig + 1] += i2;
      } else {
        o1[2 * ig] = r1; o1[2 * ig + 1] = i1;
        o2[2 * ig] = r2; o2[2 * ig + 1] = i2;
      }
    } else {
      if (accum) { o1[2 * ig] += a; o1[2 * ig + 1] += b; }
      else { o1[2 * ig] = a; o1[2 * ig + 1] = b; }
    }
  }
}

// ------------------------------------------------------------------ x-bwd
// Block (plane block kb, active row r, pair): the 16 x-lines of row r in the
// planes of kb, line kk = plane KB*kb + kk.  Lane kk is fastest so that a
// wave reads whole KB-plane chunks of adjacent columns.  colmap[r*N + x] =
// workspace column at (row r, x) or -1.  The result is staged in LDS and
// stored as XB-wide runs of xbuf.
template <typename T, int N1, int N2>
__global__ __launch_bounds__(xthreads(N1, N2)) void kx_bwd(
    const cpx<T>* __restrict__ zbuf, cpx<T>* __restrict__ xbuf, int np2, int ncols, int nact,
    const int32_t* __restrict__ colmap, const cpx<T>* __restrict__ tw_g) {
  constexpr int N = N1 * N2, NP = N + 1, NT = xthreads(N1, N2), NXB = nblk(N, XB);
  using Mp = Map<N1, N2>;
  __shared__ cpx<T> ex[KB * NP];
  __shared__ cpx<T> tw[Mp::pfa ? 1 : N];
  const int tid = threadIdx.x;
  const int nkb = nblk(np2, KB);
  const int kb = blockIdx.x % nkb, r = blockIdx.x / nkb;
  const int pair = blockIdx.y;
  const int nk = min(KB, np2 - kb * KB);           // valid planes of this block
  if constexpr (!Mp::pfa) {
    for (int i = tid; i < N; i += NT) tw[i] = tw_g[i];
    __syncthreads();
  }
  const int kk = tid % KB, p = tid / KB;
  // the row's column map as byte offsets into the block's zbuf run, stored
  // in the order phase A reads them (thread p, element n1): one ds_read with
  // an immediate offset per element; OOB where the row has no column at x
  __shared__ uint32_t cmo[N];
  for (int i = tid; i < N; i += NT) {
    const int pp = i / N1, n1 = i - pp * N1;
    const int c = colmap[(long)r * N + Mp::in_a(n1, pp)];
    cmo[i] = c >= 0 ? (uint32_t)(c * KB) * (uint32_t)sizeof(cpx<T>) : OOB;
  }
  __syncthreads();
  if (p < N2 && kk < nk) {                // phase A, p = n2
    const auto zr = buf_res(zbuf + ((long)pair * nkb + kb) * ncols * KB,
                            (uint32_t)ncols * KB * sizeof(cpx<T>));
    const uint32_t lo = kk * (uint32_t)sizeof(cpx<T>);
    cpx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = buf_load<T>(zr, cmo[p * N1 + n1] + lo);
    rdft<T, N1, +1>(a);
    if constexpr (!Mp::pfa) {
#pragma unroll
      for (int k1 = 1; k1 < N1; k1++)
        if (p) a[k1] = mul(a[k1], twl<T, +1>(tw, p * k1));
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) ex[kk * NP + k1 * N2 + p] = a[k1];
  }
  __syncthreads();
  cpx<T> b[N2];
  const bool vb = p < N1 && kk < nk;
  if (vb) {                               // phase B, p = k1
#pragma unroll
    for (int n2 = 0; n2 < N2; n2++) b[n2] = ex[kk * NP + p * N2 + n2];
    rdft<T, N2, +1>(b);
  }
  __syncthreads();
  if (vb) {
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) ex[kk * NP + Mp::out_b(p, k2)] = b[k2];
  }
  __syncthreads();
  // xbuf[pair][k][xb][r][xi]: runs of XB elements (run j = l*NXB + xb of
  // plane l at j*nact*XB from the block's first); unrolled so that the LDS
  // reads and stores of all runs issue back to back
  constexpr int SQ = NT / XB, NIT = nblk(KB * NXB, SQ);
  cpx<T>* xo = xbuf + (((long)pair * np2 + kb * KB) * NXB * nact + r) * XB;
  const auto xr = buf_res(xo, (uint32_t)(KB * NXB * XB) * nact * (uint32_t)sizeof(cpx<T>));
  const int sxi = tid % XB, sq = tid / XB;
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int j = sq + it * SQ;
    const int l = j / NXB, xb = j - l * NXB;
    const int x = xb * XB + sxi;
    const bool ok = j < nk * NXB && x < N;
    const cpx<T> v = ex[ok ? l * NP + x : 0];
    buf_store<T>(xr, ok ? (uint32_t)(j * nact * XB + sxi) * (uint32_t)sizeof(cpx<T>) : OOB, v);
  }
}

// ------------------------------------------------------------------ x-fwd
template <typename T, int N1, int N2>
__global__ __launch_bounds__(xthreads(N1, N2)) void kx_fwd(
    const cpx<T>* __restrict__ xbuf, cpx<T>* __restrict__ zbuf, int np2, int ncols, int nact,
    const int32_t* __restrict__ colmap, const cpx<T>* __restrict__ tw_g) {
  constexpr int N = N1 * N2, NP = N + 1, NT = xthreads(N1, N2), NXB = nblk(N, XB);
  using Mp = Map<N1, N2>;
  __shared__ cpx<T> ex[KB * NP];
  __shared__ cpx<T> tw[Mp::pfa ? 1 : N];
  const int tid = threadIdx.x;
  const int nkb = nblk(np2, KB);
  const int kb = blockIdx.x % nkb, r = blockIdx.x / nkb;
  const int pair = blockIdx.y;
  const int nk = min(KB, np2 - kb * KB);
  if constexpr (!Mp::pfa)
    for (int i = tid; i < N; i += NT) tw[i] = tw_g[i];
  // byte offsets of the row's columns in phase B' order (as in kx_bwd)
  __shared__ uint32_t cmo[N];
  for (int i = tid; i < N; i += NT) {
    const int pp = i / N1, n1 = i - pp * N1;
    const int c = colmap[(long)r * N + Mp::in_a(n1, pp)];
    cmo[i] = c >= 0 ? (uint32_t)(c * KB) * (uint32_t)sizeof(cpx<T>) : OOB;
  }
  {
    // all runs' loads in flight before the LDS writes (layout as in kx_bwd)
    constexpr int SQ = NT / XB, NIT = nblk(KB * NXB, SQ);
    const cpx<T>* xo = xbuf + (((long)pair * np2 + kb * KB) * NXB * nact + r) * XB;
    const auto xr = buf_res(xo, (uint32_t)(KB * NXB * XB) * nact * (uint32_t)sizeof(cpx<T>));
    const int sxi = tid % XB, sq = tid / XB;
    cpx<T> st[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int j = sq + it * SQ;
      const int xb = j % NXB;
      const bool ok = j < nk * NXB && xb * XB + sxi < N;
      st[it] = buf_load<T>(xr, ok ? (uint32_t)(j * nact * XB + sxi) * (uint32_t)sizeof(cpx<T>)
                                  : OOB);
    }
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int j = sq + it * SQ;
      const int l = j / NXB, xb = j - l * NXB;
      const int x = xb * XB + sxi;
      if (j < nk * NXB && x < N) ex[l * NP + x] = st[it];
    }
  }
  __syncthreads();
  const int kk = tid % KB, p = tid / KB;
  cpx<T> b[N2];
  const bool va = p < N1 && kk < nk;
  if (va) {                               // forward phase A', p = k1
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) b[k2] = ex[kk * NP + Mp::out_b(p, k2)];
    rdft<T, N2, -1>(b);
    if constexpr (!Mp::pfa) {
#pragma unroll
      for (int m1 = 1; m1 < N2; m1++)
        if (p) b[m1] = mul(b[m1], twl<T, -1>(tw, p * m1));
    }
  }
  __syncthreads();
  if (va) {
#pragma unroll
    for (int m1 = 0; m1 < N2; m1++) ex[kk * NP + m1 * N1 + p] = b[m1];
  }
  __syncthreads();
  if (p < N2 && kk < nk) {                // phase B', p = m1
    cpx<T> cv[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) cv[k1] = ex[kk * NP + p * N1 + k1];
    rdft<T, N1, -1>(cv);
    const auto zr = buf_res(zbuf + ((long)pair * nkb + kb) * ncols * KB,
                            (uint32_t)ncols * KB * sizeof(cpx<T>));
    const uint32_t lo = kk * (uint32_t)sizeof(cpx<T>);
#pragma unroll
    for (int m2 = 0; m2 < N1; m2++) buf_store<T>(zr, cmo[p * N1 + m2] + lo, cv[m2]);
  }
}

// ------------------------------------------------------------------ y pass
// Block (x block xb, plane k, chunk of PPB pairs): the XB y-lines at
// x = XB*xb + xi, one per lane group; lane xi is fastest, so a wave reads and
// writes whole XB-wide runs of xbuf (adjacent rows are adjacent runs) and of
// the real-space grids.  jrow[y] = active row of y or -1.
// MC: mode class, 0 = Y_VMUL, 1 = Y_DENSITY, 2 = the single full-grid
// transforms (mode selects which); one instantiation per class keeps each
// kernel's register allocation to what its mode needs.
template <typename T, int N1, int N2, int MC>
__global__ __launch_bounds__(xthreads(N1, N2)) void ky(
    cpx<T>* __restrict__ xbuf, int np0, int np2, int nact, const int32_t* __restrict__ jrow,
    int nbatch, int PPB, const cpx<T>* __restrict__ tw_g, const double* __restrict__ vin,
    const double* __restrict__ fac, double* __restrict__ rho, double* __restrict__ fgrid,
    double scale, int mode) {
  constexpr int N = N1 * N2, NP = N + 1, NT = xthreads(N1, N2);
  using Mp = Map<N1, N2>;
  __shared__ cpx<T> ex[XB * NP];
  __shared__ cpx<T> tw[Mp::pfa ? 1 : N];
  // byte offsets of the active rows within a pair's run of this block, in
  // the order phase A reads them (thread q, element n1); OOB for the rows
  // outside the active band: one ds_read per element, no branch
  __shared__ uint32_t jo[N];
  const int tid = threadIdx.x;
  const int xb = blockIdx.x, k = blockIdx.y;
  const int nxb = nblk(np0, XB);
  if constexpr (!Mp::pfa)
    for (int i = tid; i < N; i += NT) tw[i] = tw_g[i];
  for (int i = tid; i < N; i += NT) {
    const int qq = i / N1, n1 = i - qq * N1;
    const int r = jrow[Mp::in_a(n1, qq)];
    jo[i] = r >= 0 ? (uint32_t)(r * XB) * (uint32_t)sizeof(cpx<T>) : OOB;
  }
  __syncthreads();
  const int xi = tid % XB, q = tid / XB;
  const int x = xb * XB + xi;
  const bool vx = x < np0;
  const bool vA = vx && q < N2;           // phase A / B' role (N2 per line)
  const bool vB = vx && q < N1;           // phase B / A' role (N1 per line)
  const long plane = (long)k * np0 * N;
  // pair p's run of this block: nact rows of XB elements
  const long pstride = (long)np2 * nxb * nact * XB;
  cpx<T>* const run0 = xbuf + (((long)k * nxb + xb) * nact) * XB;
  const uint32_t runbytes = (uint32_t)nact * XB * sizeof(cpx<T>);
  const uint32_t lo = xi * (uint32_t)sizeof(cpx<T>);
  const uint32_t* const jq = jo + q * N1;

  if (MC == 0) mode = Y_VMUL;
  if (MC == 1) mode = Y_DENSITY;
  if (MC == 2 && (mode == Y_FWD_FULL || mode == Y_FWD_REAL)) {
    cpx<T> b[N2];
    if (vB) {
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) {
        const long gi = plane + (long)Mp::out_b(q, k2) * np0 + x;
        b[k2] = mode == Y_FWD_FULL ? cpx<T>{T(fgrid[2 * gi]), T(fgrid[2 * gi + 1])}
                                   : cpx<T>{T(scale * vin[gi]), T(0)};
      }
      rdft<T, N2, -1>(b);
      if constexpr (!Mp::pfa) {
        if constexpr (!Mp::pfa) {
#pragma unroll
          for (int m1 = 1; m1 < N2; m1++)
            if (q) b[m1] = mul(b[m1], twl<T, -1>(tw, q * m1));
        }
      }
#pragma unroll
      for (int m1 = 0; m1 < N2; m1++) ex[xi * NP + m1 * N1 + q] = b[m1];
    }
    __syncthreads();
    if (vA) {
      cpx<T> cv[N1];
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) cv[k1] = ex[xi * NP + q * N1 + k1];
      rdft<T, N1, -1>(cv);
      const auto xr = buf_res(run0, runbytes);
#pragma unroll
      for (int m2 = 0; m2 < N1; m2++) buf_store<T>(xr, jq[m2] + lo, cv[m2]);
    }
    return;
  }

  constexpr int NACC = MC == 1 ? N2 : 1;
  double acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; i++) acc[i] = 0.0;
  // the potential of this thread's points, read once for all pairs of the block
  constexpr int NV = MC == 0 ? N2 : 1;
  T vreg[NV];
  if (MC == 0 && vB) {
#pragma unroll
    for (int k2 = 0; k2 < NV; k2++) vreg[k2] = (T)vin[plane + (long)Mp::out_b(q, k2) * np0 + x];
  }
  const int p0 = blockIdx.z * PPB;
  const int p1 = min(nbatch, p0 + PPB);
  for (int pair = p0; pair < p1; pair++) {
    __syncthreads();                      // previous pair's readers of ex are done
    const auto xr = buf_res(run0 + pair * pstride, runbytes);
    if (vA) {                             // phase A, q = n2 (bwd)
      cpx<T> a[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; n1++) a[n1] = buf_load<T>(xr, jq[n1] + lo);
      rdft<T, N1, +1>(a);
      if constexpr (!Mp::pfa) {
#pragma unroll
        for (int k1 = 1; k1 < N1; k1++)
          if (q) a[k1] = mul(a[k1], twl<T, +1>(tw, q * k1));
      }
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) ex[xi * NP + k1 * N2 + q] = a[k1];
    }
    __syncthreads();
    cpx<T> b[N2];
    if (vB) {                             // phase B, q = k1: psi at y = q + N1*k2
#pragma unroll
      for (int n2 = 0; n2 < N2; n2++) b[n2] = ex[xi * NP + q * N2 + n2];
      rdft<T, N2, +1>(b);
    }
    if (MC == 0) {
      // tmp *= v (src/SlaterDet.cpp:432-434), then forward along y
      if (vB) {
#pragma unroll
        for (int k2 = 0; k2 < NV; k2++) b[k2] = {b[k2].x * vreg[k2], b[k2].y * vreg[k2]};
        rdft<T, N2, -1>(b);
        if constexpr (!Mp::pfa) {
#pragma unroll
          for (int m1 = 1; m1 < N2; m1++)
            if (q) b[m1] = mul(b[m1], twl<T, -1>(tw, q * m1));
        }
      }
      __syncthreads();
      if (vB) {
#pragma unroll
        for (int m1 = 0; m1 < N2; m1++) ex[xi * NP + m1 * N1 + q] = b[m1];
      }
      __syncthreads();
      if (vA) {
        cpx<T> cv[N1];
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) cv[k1] = ex[xi * NP + q * N1 + k1];
        rdft<T, N1, -1>(cv);
#pragma unroll
        for (int m2 = 0; m2 < N1; m2++) buf_store<T>(xr, jq[m2] + lo, cv[m2]);
      }
    } else if (MC == 1) {
      if (vB) {
        const double f1 = fac[2 * pair], f2 = fac[2 * pair + 1];
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) {
          const double p1v = (double)b[k2].x, p2v = (double)b[k2].y;
          acc[k2 % NACC] += f1 * p1v * p1v + f2 * p2v * p2v;
        }
      }
    } else if (mode == Y_BWD_FULL) {
      if (vB) {
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) {
          const long gi = plane + (long)Mp::out_b(q, k2) * np0 + x;
          fgrid[2 * gi] = (double)b[k2].x;
          fgrid[2 * gi + 1] = (double)b[k2].y;
        }
      }
    } else if (mode == Y_BWD_REAL_ADD) {
      if (vB) {
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) {
          const long gi = plane + (long)Mp::out_b(q, k2) * np0 + x;
          fgrid[gi] = (vin ? vin[gi] : 0.0) + (double)b[k2].x * scale;
        }
      }
    }
  }
  if (MC == 1 && vB) {
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) rho[plane + (long)Mp::out_b(q, k2) * np0 + x] += acc[k2 % NACC];
  }
}

// ---------------------------------------------------------------- launchers
template <typename T>
inline const cpx<T>* twp(const qbg_grid* g, int axis) {
  if constexpr (sizeof(T) == 8) return reinterpret_cast<const cpx<T>*>(g->d_tw64[axis]);
  else return reinterpret_cast<const cpx<T>*>(g->d_tw32[axis]);
}

template <int N1, int N2>
struct Launch {
  static constexpr int N = N1 * N2;
  static constexpr int L = 256 / cmax(N1, N2);

  template <typename T>
  static void zbwd_t(qbg_grid* g, int single, const double* c, long pstride, long off2,
                     int last_has2, int nbatch, hipStream_t s) {
    const size_t lds = sizeof(cpx<T>) * ((size_t)g->zcols * N + N);
    hipLaunchKernelGGL((kz_bwd<T, N1, N2>), dim3(g->ntiles, nbatch), dim3(256), lds, s, c,
                       pstride, off2, last_has2, nbatch, single, g->d_ip, g->d_im, g->d_glist,
                       g->d_tiles, g->ncols, g->zcols, twp<T>(g, 2), (cpx<T>*)g->d_zbuf);
  }
  static void zbwd(qbg_grid* g, int prec, int single, const double* c, long pstride, long off2,
                   int last_has2, int nbatch, hipStream_t s) {
    if (prec == QBG_PREC_FP64) zbwd_t<double>(g, single, c, pstride, off2, last_has2, nbatch, s);
    else zbwd_t<float>(g, single, c, pstride, off2, last_has2, nbatch, s);
  }

  template <typename T>
  static void zfwd_t(qbg_grid* g, int single, int accum, double* out, long pstride, long off2,
                     int last_has2, int nbatch, hipStream_t s, ZFin fin) {
    const size_t lds = sizeof(cpx<T>) * ((size_t)g->zcols * N + N);
    const double scale = 1.0 / ((double)g->np0 * g->np1 * g->np2);
    hipLaunchKernelGGL((kz_fwd<T, N1, N2>), dim3(g->ntiles, nbatch), dim3(256), lds, s,
                       (const cpx<T>*)g->d_zbuf, out, pstride, off2, last_has2, nbatch, single,
                       accum, g->d_ip, g->d_im, g->d_glist, g->d_tiles, g->ncols, g->zcols, twp<T>(g, 2),
                       scale, fin);
  }
  static void zfwd(qbg_grid* g, int prec, int single, int accum, double* out, long pstride,
                   long off2, int last_has2, int nbatch, hipStream_t s, ZFin fin) {
    if (prec == QBG_PREC_FP64)
      zfwd_t<double>(g, single, accum, out, pstride, off2, last_has2, nbatch, s, fin);
    else
      zfwd_t<float>(g, single, accum, out, pstride, off2, last_has2, nbatch, s, fin);
  }

  static void xpass(qbg_grid* g, int prec, bool fwd, int nbatch, hipStream_t s) {
    constexpr int NT = xthreads(N1, N2);
    dim3 grid(nblk(g->np2, KB) * g->nact, nbatch);
    if (prec == QBG_PREC_FP64) {
      if (!fwd)
        hipLaunchKernelGGL((kx_bwd<double, N1, N2>), grid, dim3(NT), 0, s,
                           (const cpx<double>*)g->d_zbuf, (cpx<double>*)g->xws, g->np2,
                           g->ncols, g->nact, g->d_colmap, twp<double>(g, 0));
      else
        hipLaunchKernelGGL((kx_fwd<double, N1, N2>), grid, dim3(NT), 0, s,
                           (const cpx<double>*)g->xws, (cpx<double>*)g->d_zbuf, g->np2,
                           g->ncols, g->nact, g->d_colmap, twp<double>(g, 0));
    } else {
      if (!fwd)
        hipLaunchKernelGGL((kx_bwd<float, N1, N2>), grid, dim3(NT), 0, s,
                           (const cpx<float>*)g->d_zbuf, (cpx<float>*)g->xws, g->np2,
                           g->ncols, g->nact, g->d_colmap, twp<float>(g, 0));
      else
        hipLaunchKernelGGL((kx_fwd<float, N1, N2>), grid, dim3(NT), 0, s,
                           (const cpx<float>*)g->xws, (cpx<float>*)g->d_zbuf, g->np2,
                           g->ncols, g->nact, g->d_colmap, twp<float>(g, 0));
    }
  }

  static void ypass(qbg_grid* g, int prec, int mode, int nbatch, int ppb, const double* v,
                    const double* fac, double* rho, double* fgrid, double scale, hipStream_t s) {
    const bool fwd = mode == Y_FWD_FULL || mode == Y_FWD_REAL;
    if (fwd) { nbatch = 1; ppb = 1; }
    dim3 grid(nblk(g->np0, XB), g->np2, (nbatch + ppb - 1) / ppb);
    const int mc = mode == Y_VMUL ? 0 : (mode == Y_DENSITY ? 1 : 2);
    if (prec == QBG_PREC_FP64) ylaunch<double>(mc, grid, g, nbatch, ppb, v, fac, rho, fgrid, scale, mode, s);
    else ylaunch<float>(mc, grid, g, nbatch, ppb, v, fac, rho, fgrid, scale, mode, s);
  }
  template <typename T>
  static void ylaunch(int mc, dim3 grid, qbg_grid* g, int nbatch, int ppb, const double* v,
                      const double* fac, double* rho, double* fgrid, double scale, int mode,
                      hipStream_t s) {
    constexpr int NT = xthreads(N1, N2);
#define QBG_KY(MC)                                                                            \
  hipLaunchKernelGGL((ky<T, N1, N2, MC>), grid, dim3(NT), 0, s, (cpx<T>*)g->xws, g->np0,     \
                     g->np2, g->nact, g->d_jrow, nbatch, ppb, twp<T>(g, 1), v, fac, rho, fgrid, \
                     scale, mode)
    if (mc == 0) QBG_KY(0);
    else if (mc == 1) QBG_KY(1);
    else QBG_KY(2);
#undef QBG_KY
  }
};

}  // namespace f4
}  // namespace qbg

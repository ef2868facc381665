// fft.hip -- batched sparse 3D FFT passes for gfx950.
//
// One band-pair transform (FourierTransform::backward(c1,c2,f) followed by
// ...forward(f,c1,c2), src/FourierTransform.cpp:147-239) is split into five
// HBM passes over a batch of pairs:
//   z-bwd : pack two real-basis bands onto the G-sphere columns
//           (BasisMapping::doublevector_to_zvec, src/BasisMapping.cpp:667-700)
//           inside LDS and transform the columns along z;
//   x-bwd : scatter the columns onto the active y rows of each z plane
//           (transpose_bwd, :445-467) and transform along x;
//   y     : transform along y for every x; fused with the potential
//           multiply (SlaterDet::rs_mul_add, src/SlaterDet.cpp:432-434) and
//           the forward y transform, or with the density accumulation
//           (src/SlaterDet.cpp:239-257);
//   x-fwd : transform the active rows along x and gather the columns;
//   z-fwd : transform the columns, scale by 1/np012 (fz, :679-761) and
//           unpack into (or accumulate onto) the two bands
//           (zvec_to_doublevector, :721-750).
// Only the active rows (2*ntrans0 per plane) and the columns ever reach HBM;
// the columns are numbered row by row (grid.cpp).  Each 1D transform is a
// Stockham mixed-radix FFT in LDS.
#include <hip/hip_runtime.h>
#include "internal.h"

namespace qbg {

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }

// small DFT of size R: out[k] = sum_j v[j] exp(SIGN 2 pi i j k / R);
// tw[m] = exp(-2 pi i m / n) of the line length n, R | n.
template <typename T, int R, int SIGN>
__device__ __forceinline__ void small_dft(cpx<T>* v, const cpx<T>* __restrict__ tw, int n) {
  if constexpr (R == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cpx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cpx<T> b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
    // SIGN*i*b1
    cpx<T> jb1 = SIGN > 0 ? cpx<T>{-b1.y, b1.x} : cpx<T>{b1.y, -b1.x};
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, jb1);
    v[3] = csub(a1, jb1);
  } else {
    cpx<T> out[R];
    const int step = n / R;
#pragma unroll
    for (int k = 0; k < R; k++) {
      cpx<T> s = v[0];
#pragma unroll
      for (int j = 1; j < R; j++) {
        cpx<T> w = tw[((j * k) % R) * step];
        if (SIGN > 0) w.y = -w.y;
        s = cadd(s, cmul(v[j], w));
      }
      out[k] = s;
    }
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = out[k];
  }
}

// One Stockham stage (decimation in time) for butterfly j of one line:
// reads x[j + r*n/R], twiddles by w^(r*(j%Ns)*n/(Ns*R)), writes
// y[(j/Ns)*Ns*R + j%Ns + r*Ns].
template <typename T, int R, int SIGN>
__device__ __forceinline__ void butterfly(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ dst,
                                          int j, int n, int Ns, int es,
                                          const cpx<T>* __restrict__ tw) {
  const int nR = n / R;
  const int jm = j % Ns;
  const int tstep = jm * (nR / Ns);
  cpx<T> v[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    cpx<T> t = src[(j + r * nR) * es];
    if (r > 0 && jm > 0) {
      cpx<T> w = tw[r * tstep];
      if (SIGN > 0) w.y = -w.y;
      t = cmul(t, w);
    }
    v[r] = t;
  }
  small_dft<T, R, SIGN>(v, tw, n);
  const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
  for (int r = 0; r < R; r++) dst[(base + r * Ns) * es] = v[r];
}

// Transform nlines lines of length plan.n held in LDS buffer a (element e of
// line l at a[l*ls + e*es]); b is scratch of the same size.  Returns the
// buffer holding the result.  All threads of the block must call it.
template <typename T, int SIGN>
__device__ cpx<T>* lds_fft(cpx<T>* a, cpx<T>* b, const FftPlan& pl, int nlines, int ls,
                           int es, bool lines_fast, const cpx<T>* __restrict__ tw) {
  const int n = pl.n;
  cpx<T>* src = a;
  cpx<T>* dst = b;
  int Ns = 1;
  for (int s = 0; s < pl.nstage; s++) {
    const int R = pl.radix[s];
    const int nR = n / R;
    const int total = nlines * nR;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      int line, j;
      if (lines_fast) { line = t % nlines; j = t / nlines; }
      else { j = t % nR; line = t / nR; }
      const cpx<T>* sl = src + line * ls;
      cpx<T>* dl = dst + line * ls;
      switch (R) {
        case 2: butterfly<T, 2, SIGN>(sl, dl, j, n, Ns, es, tw); break;
        case 3: butterfly<T, 3, SIGN>(sl, dl, j, n, Ns, es, tw); break;
        case 4: butterfly<T, 4, SIGN>(sl, dl, j, n, Ns, es, tw); break;
        case 5: butterfly<T, 5, SIGN>(sl, dl, j, n, Ns, es, tw); break;
        case 7: butterfly<T, 7, SIGN>(sl, dl, j, n, Ns, es, tw); break;
        case 11: butterfly<T, 11, SIGN>(sl, dl, j, n, Ns, es, tw); break;
        default: break;
      }
    }
    __syncthreads();
    cpx<T>* tmp = src; src = dst; dst = tmp;
    Ns *= R;
  }
  return src;
}

template <typename T>
__device__ __forceinline__ void zero_lds(cpx<T>* a, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = {T(0), T(0)};
}

// ----------------------------------------------------------------- z-bwd
// band pointers of pair b: c1 = c + 2*(b*pstride), c2 = c1 + 2*off2
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_zbwd(
    const double* __restrict__ c, long pstride, long off2, int last_has2, int nbatch,
    const int32_t* __restrict__ ip, const int32_t* __restrict__ im,
    const int32_t* __restrict__ glist, const qbg_tile* __restrict__ tiles, int np2, int ncols,
    int TC, FftPlan plan, const cpx<T>* __restrict__ tw, cpx<T>* __restrict__ zbuf) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + TC * np2;
  const qbg_tile t = tiles[blockIdx.x];
  const int pair = blockIdx.y;
  zero_lds(A, TC * np2);
  __syncthreads();
  const int slot0 = t.col0 * np2;
  const double* c1 = c + 2 * (long)pair * pstride;
  const bool has2 = (MODE == Z_PAIR) && (pair < nbatch - 1 || last_has2);
  const double* c2 = c1 + 2 * off2;
  for (int j = t.g0 + threadIdx.x; j < t.g1; j += blockDim.x) {
    const int ig = glist[j];
    const double a = c1[2 * ig], b = c1[2 * ig + 1];
    double cc = 0.0, d = 0.0;
    if (has2) { cc = c2[2 * ig]; d = c2[2 * ig + 1]; }
    const int sp = ip[ig] - slot0, sm = im[ig] - slot0;
    const int lp = (sp % np2) * TC + sp / np2;
    const int lm = (sm % np2) * TC + sm / np2;
    // same thread writes +G then -G: the -G value wins at G=0 as in
    // src/BasisMapping.cpp:691-699
    A[lp] = {T(a - d), T(b + cc)};
    A[lm] = {T(a + d), T(cc - b)};
  }
  __syncthreads();
  cpx<T>* res = lds_fft<T, +1>(A, B, plan, t.ncol, 1, TC, true, tw);
  cpx<T>* zp = zbuf + (long)pair * ncols * np2;
  const int tot = t.ncol * np2;
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int k = i / t.ncol, cl = i - k * t.ncol;
    const int gcol = cl < t.np ? t.col0 + cl : t.colm + (cl - t.np);
    zp[(long)k * ncols + gcol] = res[k * TC + cl];
  }
}

// ----------------------------------------------------------------- x-bwd
template <typename T>
__global__ __launch_bounds__(256) void k_xbwd(
    const cpx<T>* __restrict__ zbuf, cpx<T>* __restrict__ xbuf, int np0, int np2, int ncols,
    int nact, int RB, int ngroups, const int32_t* __restrict__ row_start,
    const int32_t* __restrict__ rc_ivec, const int32_t* __restrict__ rc_x, FftPlan plan,
    const cpx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + RB * np0;
  const int k = blockIdx.x / ngroups;
  const int r0 = (blockIdx.x - k * ngroups) * RB;
  const int nr = min(RB, nact - r0);
  const int pair = blockIdx.y;
  zero_lds(A, RB * np0);
  __syncthreads();
  const cpx<T>* zp = zbuf + (long)pair * ncols * np2 + (long)k * ncols;
  const int e0 = row_start[r0], e1 = row_start[r0 + nr];
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    // row of entry e: linear search over the few rows of this block
    int r = r0;
    while (row_start[r + 1] <= e) r++;
    A[(r - r0) * np0 + rc_x[e]] = zp[rc_ivec[e]];
  }
  __syncthreads();
  cpx<T>* res = lds_fft<T, +1>(A, B, plan, nr, np0, 1, false, tw);
  cpx<T>* xp = xbuf + ((long)pair * np2 + k) * (long)nact * np0 + (long)r0 * np0;
  for (int i = threadIdx.x; i < nr * np0; i += blockDim.x) xp[i] = res[i];
}

// ----------------------------------------------------------------- x-fwd
template <typename T>
__global__ __launch_bounds__(256) void k_xfwd(
    const cpx<T>* __restrict__ xbuf, cpx<T>* __restrict__ zbuf, int np0, int np2, int ncols,
    int nact, int RB, int ngroups, const int32_t* __restrict__ row_start,
    const int32_t* __restrict__ rc_ivec, const int32_t* __restrict__ rc_x, FftPlan plan,
    const cpx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + RB * np0;
  const int k = blockIdx.x / ngroups;
  const int r0 = (blockIdx.x - k * ngroups) * RB;
  const int nr = min(RB, nact - r0);
  const int pair = blockIdx.y;
  const cpx<T>* xp = xbuf + ((long)pair * np2 + k) * (long)nact * np0 + (long)r0 * np0;
  for (int i = threadIdx.x; i < nr * np0; i += blockDim.x) A[i] = xp[i];
  __syncthreads();
  cpx<T>* res = lds_fft<T, -1>(A, B, plan, nr, np0, 1, false, tw);
  cpx<T>* zp = zbuf + (long)pair * ncols * np2 + (long)k * ncols;
  const int e0 = row_start[r0], e1 = row_start[r0 + nr];
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    int r = r0;
    while (row_start[r + 1] <= e) r++;
    zp[rc_ivec[e]] = res[(r - r0) * np0 + rc_x[e]];
  }
}

// ----------------------------------------------------------------- z-fwd
template <typename T, int MODE, int OUT>
__global__ __launch_bounds__(256) void k_zfwd(
    const cpx<T>* __restrict__ zbuf, double* __restrict__ out, long pstride, long off2,
    int last_has2, int nbatch, const int32_t* __restrict__ ip, const int32_t* __restrict__ im,
    const int32_t* __restrict__ glist, const qbg_tile* __restrict__ tiles, int np2, int ncols,
    int TC, FftPlan plan, const cpx<T>* __restrict__ tw, double scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + TC * np2;
  const qbg_tile t = tiles[blockIdx.x];
  const int pair = blockIdx.y;
  const cpx<T>* zp = zbuf + (long)pair * ncols * np2;
  const int tot = t.ncol * np2;
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int k = i / t.ncol, cl = i - k * t.ncol;
    const int gcol = cl < t.np ? t.col0 + cl : t.colm + (cl - t.np);
    A[k * TC + cl] = zp[(long)k * ncols + gcol];
  }
  __syncthreads();
  cpx<T>* res = lds_fft<T, -1>(A, B, plan, t.ncol, 1, TC, true, tw);
  const int slot0 = t.col0 * np2;
  double* o1 = out + 2 * (long)pair * pstride;
  double* o2 = o1 + 2 * off2;
  const bool has2 = (MODE == Z_PAIR) && (pair < nbatch - 1 || last_has2);
  for (int j = t.g0 + threadIdx.x; j < t.g1; j += blockDim.x) {
    const int ig = glist[j];
    const int sp = ip[ig] - slot0;
    const int lp = (sp % np2) * TC + sp / np2;
    const cpx<T> zpv = res[lp];
    const double a = (double)zpv.x * scale, b = (double)zpv.y * scale;
    if (MODE == Z_PAIR && has2) {
      const int sm = im[ig] - slot0;
      const int lm = (sm % np2) * TC + sm / np2;
      const cpx<T> zmv = res[lm];
      const double cc = (double)zmv.x * scale, d = (double)zmv.y * scale;
      const double r1 = 0.5 * (a + cc), i1 = 0.5 * (b - d);
      const double r2 = 0.5 * (b + d), i2 = 0.5 * (cc - a);
      if (OUT == OUT_ACCUM) {
        o1[2 * ig] += r1; o1[2 * ig + 1] += i1;
        o2[2 * ig] += r2; o2[2 * ig + 1] += i2;
      } else {
        o1[2 * ig] = r1; o1[2 * ig + 1] = i1;
        o2[2 * ig] = r2; o2[2 * ig + 1] = i2;
      }
    } else {
      // zvec_to_vector (src/BasisMapping.cpp:703-718)
      if (OUT == OUT_ACCUM) { o1[2 * ig] += a; o1[2 * ig + 1] += b; }
      else { o1[2 * ig] = a; o1[2 * ig + 1] = b; }
    }
  }
}

// ---------------------------------------------------------------- y pass
// Block (x-tile, plane k, chunk of PPB pairs).  LDS holds np1 x XT values
// with x fastest so that global rows are read/written contiguously.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_ypass(
    cpx<T>* __restrict__ xbuf, int np0, int np1, int np2, int nact, int XT,
    const int32_t* __restrict__ act_row, int nbatch, int PPB, FftPlan plan,
    const cpx<T>* __restrict__ tw, const double* __restrict__ vin,
    const double* __restrict__ fac, double* __restrict__ rho, double* __restrict__ fgrid,
    double scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* B = A + XT * np1;
  const int x0 = blockIdx.x * XT;
  const int nx = min(XT, np0 - x0);
  const int k = blockIdx.y;
  const long plane = (long)k * np0 * np1;
  const int nel = np1 * nx;
  constexpr int MAXE = 16;   // elements per thread (np1*XT <= 16*blockDim)
  double acc[MAXE];
#pragma unroll
  for (int m = 0; m < MAXE; m++) acc[m] = 0.0;

  if (MODE == Y_FWD_FULL || MODE == Y_FWD_REAL) {
    for (int i = threadIdx.x; i < nel; i += blockDim.x) {
      const int j = i / nx, xl = i - j * nx;
      const long gi = plane + (long)j * np0 + x0 + xl;
      if (MODE == Y_FWD_FULL) A[j * XT + xl] = {T(fgrid[2 * gi]), T(fgrid[2 * gi + 1])};
      else A[j * XT + xl] = {T(scale * vin[gi]), T(0)};
    }
    __syncthreads();
    cpx<T>* res = lds_fft<T, -1>(A, B, plan, nx, 1, XT, true, tw);
    cpx<T>* xp = xbuf + (long)k * nact * np0;
    for (int i = threadIdx.x; i < nact * nx; i += blockDim.x) {
      const int r = i / nx, xl = i - r * nx;
      xp[(long)r * np0 + x0 + xl] = res[act_row[r] * XT + xl];
    }
    return;
  }

  const int p0 = blockIdx.z * PPB;
  const int p1 = min(nbatch, p0 + PPB);
  for (int pair = p0; pair < p1; pair++) {
    zero_lds(A, XT * np1);
    __syncthreads();
    cpx<T>* xp = xbuf + ((long)pair * np2 + k) * (long)nact * np0;
    for (int i = threadIdx.x; i < nact * nx; i += blockDim.x) {
      const int r = i / nx, xl = i - r * nx;
      A[act_row[r] * XT + xl] = xp[(long)r * np0 + x0 + xl];
    }
    __syncthreads();
    cpx<T>* res = lds_fft<T, +1>(A, B, plan, nx, 1, XT, true, tw);
    cpx<T>* oth = (res == A) ? B : A;
    if (MODE == Y_VMUL) {
      for (int i = threadIdx.x; i < nel; i += blockDim.x) {
        const int j = i / nx, xl = i - j * nx;
        const T v = (T)vin[plane + (long)j * np0 + x0 + xl];
        cpx<T> z = res[j * XT + xl];
        res[j * XT + xl] = {z.x * v, z.y * v};
      }
      __syncthreads();
      cpx<T>* r2 = lds_fft<T, -1>(res, oth, plan, nx, 1, XT, true, tw);
      for (int i = threadIdx.x; i < nact * nx; i += blockDim.x) {
        const int r = i / nx, xl = i - r * nx;
        xp[(long)r * np0 + x0 + xl] = r2[act_row[r] * XT + xl];
      }
      __syncthreads();
    } else if (MODE == Y_DENSITY) {
      const double f1 = fac[2 * pair], f2 = fac[2 * pair + 1];
#pragma unroll
      for (int m = 0; m < MAXE; m++) {
        const int i = threadIdx.x + m * blockDim.x;
        if (i < nel) {
          const int j = i / nx, xl = i - j * nx;
          const cpx<T> z = res[j * XT + xl];
          const double p1v = (double)z.x, p2v = (double)z.y;
          acc[m] += f1 * p1v * p1v + f2 * p2v * p2v;
        }
      }
      __syncthreads();
    } else if (MODE == Y_BWD_FULL) {
      for (int i = threadIdx.x; i < nel; i += blockDim.x) {
        const int j = i / nx, xl = i - j * nx;
        const long gi = plane + (long)j * np0 + x0 + xl;
        const cpx<T> z = res[j * XT + xl];
        fgrid[2 * gi] = (double)z.x;
        fgrid[2 * gi + 1] = (double)z.y;
      }
      __syncthreads();
    } else if (MODE == Y_BWD_REAL_ADD) {
      // v_r = vxc_r + Re(vft->backward(vlocal_g)) (src/EnergyFunctional.cpp:380);
      // without vin: rhor = Re(vft->backward(rhog)) * scale (ChargeDensity::update_rhor)
      for (int i = threadIdx.x; i < nel; i += blockDim.x) {
        const int j = i / nx, xl = i - j * nx;
        const long gi = plane + (long)j * np0 + x0 + xl;
        fgrid[gi] = (vin ? vin[gi] : 0.0) + (double)res[j * XT + xl].x * scale;
      }
      __syncthreads();
    }
  }
  if (MODE == Y_DENSITY) {
#pragma unroll
    for (int m = 0; m < MAXE; m++) {
      const int i = threadIdx.x + m * blockDim.x;
      if (i < nel) {
        const int j = i / nx, xl = i - j * nx;
        rho[plane + (long)j * np0 + x0 + xl] += acc[m];
      }
    }
  }
}

// ---------------------------------------------------------------- launch
static inline size_t csz(int prec) { return prec == QBG_PREC_FP64 ? 16 : 8; }

template <typename T>
static const cpx<T>* tw_of(qbg_grid* g, int axis);
template <>
const cpx<double>* tw_of<double>(qbg_grid* g, int axis) {
  return reinterpret_cast<const cpx<double>*>(g->d_tw64[axis]);
}
template <>
const cpx<float>* tw_of<float>(qbg_grid* g, int axis) {
  return reinterpret_cast<const cpx<float>*>(g->d_tw32[axis]);
}

template <typename T>
static int zbwd_t(qbg_grid* g, int mode, const double* c, long pstride, long off2,
                  int last_has2, int nbatch, hipStream_t s) {
  const size_t lds = 2 * sizeof(cpx<T>) * g->zcols * g->np2;
  dim3 grid(g->ntiles, nbatch);
  if (mode == Z_PAIR)
    hipLaunchKernelGGL((k_zbwd<T, Z_PAIR>), grid, dim3(256), lds, s, c, pstride, off2,
                       last_has2, nbatch, g->d_ip, g->d_im, g->d_glist, g->d_tiles, g->np2, g->ncols,
                       g->zcols, g->plan2, tw_of<T>(g, 2), (cpx<T>*)g->d_zbuf);
  else
    hipLaunchKernelGGL((k_zbwd<T, Z_SINGLE>), grid, dim3(256), lds, s, c, pstride, off2,
                       last_has2, nbatch, g->d_ip, g->d_im, g->d_glist, g->d_tiles, g->np2, g->ncols,
                       g->zcols, g->plan2, tw_of<T>(g, 2), (cpx<T>*)g->d_zbuf);
  return hipGetLastError() == hipSuccess ? QBG_OK : QBG_E_HIP;
}


template <typename T>
static int xpass_t(qbg_grid* g, bool fwd, int nbatch, hipStream_t s) {
  const int ngroups = (g->nact + g->xrows - 1) / g->xrows;
  const size_t lds = 2 * sizeof(cpx<T>) * g->xrows * g->np0;
  dim3 grid(g->np2 * ngroups, nbatch);
  if (!fwd)
    hipLaunchKernelGGL((k_xbwd<T>), grid, dim3(256), lds, s, (const cpx<T>*)g->d_zbuf,
                       (cpx<T>*)g->xws, g->np0, g->np2, g->ncols, g->nact, g->xrows, ngroups,
                       g->d_row_start, g->d_rc_ivec, g->d_rc_x, g->plan0, tw_of<T>(g, 0));
  else
    hipLaunchKernelGGL((k_xfwd<T>), grid, dim3(256), lds, s, (const cpx<T>*)g->xws,
                       (cpx<T>*)g->d_zbuf, g->np0, g->np2, g->ncols, g->nact, g->xrows, ngroups,
                       g->d_row_start, g->d_rc_ivec, g->d_rc_x, g->plan0, tw_of<T>(g, 0));
  return hipGetLastError() == hipSuccess ? QBG_OK : QBG_E_HIP;
}

template <typename T, int MODE>
static void ylaunch(qbg_grid* g, int nbatch, int ppb, const double* v, const double* fac,
                    double* rho, double* fgrid, double scale, hipStream_t s) {
  const size_t lds = 2 * sizeof(cpx<T>) * g->ytile * g->np1;
  dim3 grid((g->np0 + g->ytile - 1) / g->ytile, g->np2, (nbatch + ppb - 1) / ppb);
  hipLaunchKernelGGL((k_ypass<T, MODE>), grid, dim3(256), lds, s, (cpx<T>*)g->xws, g->np0,
                     g->np1, g->np2, g->nact, g->ytile, g->d_act_row, nbatch, ppb, g->plan1,
                     tw_of<T>(g, 1), v, fac, rho, fgrid, scale);
}

template <typename T>
static int ypass_t(qbg_grid* g, int mode, int nbatch, const double* v, const double* fac,
                   double* rho, double* fgrid, double scale, hipStream_t s) {
  // pairs per block: all pairs for the density (one accumulator pass, no
  // atomics); a few for the potential multiply (reuse of the v tile)
  int ppb = 1;
  if (mode == Y_DENSITY) ppb = nbatch;
  else if (mode == Y_VMUL) ppb = 4;
  switch (mode) {
    case Y_VMUL: ylaunch<T, Y_VMUL>(g, nbatch, ppb, v, fac, rho, fgrid, scale, s); break;
    case Y_DENSITY: ylaunch<T, Y_DENSITY>(g, nbatch, ppb, v, fac, rho, fgrid, scale, s); break;
    case Y_BWD_FULL: ylaunch<T, Y_BWD_FULL>(g, nbatch, ppb, v, fac, rho, fgrid, scale, s); break;
    case Y_FWD_FULL: ylaunch<T, Y_FWD_FULL>(g, 1, 1, v, fac, rho, fgrid, scale, s); break;
    case Y_BWD_REAL_ADD: ylaunch<T, Y_BWD_REAL_ADD>(g, nbatch, ppb, v, fac, rho, fgrid, scale, s); break;
    case Y_FWD_REAL: ylaunch<T, Y_FWD_REAL>(g, 1, 1, v, fac, rho, fgrid, scale, s); break;
    default: return QBG_E_INVAL;
  }
  return hipGetLastError() == hipSuccess ? QBG_OK : QBG_E_HIP;
}

template <typename T>
static int zfwd_t(qbg_grid* g, int mode, int outmode, double* out, long pstride, long off2,
                  int last_has2, int nbatch, hipStream_t s) {
  const size_t lds = 2 * sizeof(cpx<T>) * g->zcols * g->np2;
  dim3 grid(g->ntiles, nbatch);
  const double scale = 1.0 / ((double)g->np0 * g->np1 * g->np2);
#define ZF(M, O)                                                                          \
  hipLaunchKernelGGL((k_zfwd<T, M, O>), grid, dim3(256), lds, s, (const cpx<T>*)g->d_zbuf, \
                     out, pstride, off2, last_has2, nbatch, g->d_ip, g->d_im, g->d_glist,   \
                     g->d_tiles, g->np2, g->ncols, g->zcols, g->plan2, tw_of<T>(g, 2), scale)
  if (mode == Z_PAIR) {
    if (outmode == OUT_ACCUM) ZF(Z_PAIR, OUT_ACCUM); else ZF(Z_PAIR, OUT_STORE);
  } else {
    if (outmode == OUT_ACCUM) ZF(Z_SINGLE, OUT_ACCUM); else ZF(Z_SINGLE, OUT_STORE);
  }
#undef ZF
  return hipGetLastError() == hipSuccess ? QBG_OK : QBG_E_HIP;
}

// exported launchers -------------------------------------------------------
// Each one checks, before launching, that the workspace holds the batch and
// that the arrays its mode dereferences are present.
static bool ws_ok(const qbg_grid* g, int nbatch) {
  return g && g->d_zbuf && g->xws && nbatch >= 1 && nbatch <= g->ws_pairs;
}

static int launched() { return hipGetLastError() == hipSuccess ? QBG_OK : QBG_E_HIP; }

int zbwd(qbg_grid* g, int prec, int mode, const double* c, long pstride, long off2,
         int last_has2, int nbatch, hipStream_t s) {
  if (!ws_ok(g, nbatch) || !c) return QBG_E_STATE;
  if (g->f4[2]) {
    g->f4[2]->zbwd(g, prec, mode == Z_SINGLE, c, pstride, off2, last_has2, nbatch, s);
    return launched();
  }
  return prec == QBG_PREC_FP64 ? zbwd_t<double>(g, mode, c, pstride, off2, last_has2, nbatch, s)
                               : zbwd_t<float>(g, mode, c, pstride, off2, last_has2, nbatch, s);
}
static inline int pix(int prec) { return prec == QBG_PREC_FP64 ? 0 : 1; }
static inline const F4Ops* xops(const qbg_grid* g, int prec) {
  const F4Ops* o = g->fx[pix(prec)];
  return o ? o : g->f4[0];
}
static inline const F4Ops* yops(const qbg_grid* g, int prec) {
  const F4Ops* o = g->fy[pix(prec)];
  return o ? o : g->f4[1];
}
int xbwd(qbg_grid* g, int prec, int nbatch, hipStream_t s) {
  if (!ws_ok(g, nbatch)) return QBG_E_STATE;
  if (g->f4[0]) {
    xops(g, prec)->xpass(g, prec, false, nbatch, s);
    return launched();
  }
  return prec == QBG_PREC_FP64 ? xpass_t<double>(g, false, nbatch, s)
                               : xpass_t<float>(g, false, nbatch, s);
}
int xfwd(qbg_grid* g, int prec, int nbatch, hipStream_t s) {
  if (!ws_ok(g, nbatch)) return QBG_E_STATE;
  if (g->f4[0]) {
    xops(g, prec)->xpass(g, prec, true, nbatch, s);
    return launched();
  }
  return prec == QBG_PREC_FP64 ? xpass_t<double>(g, true, nbatch, s)
                               : xpass_t<float>(g, true, nbatch, s);
}
int ypass(qbg_grid* g, int prec, int mode, int nbatch, const double* v, const double* fac,
          double* rho, double* fgrid, double scale, hipStream_t s) {
  if (!ws_ok(g, nbatch)) return QBG_E_STATE;
  const bool bad = (mode == Y_VMUL && !v) || (mode == Y_DENSITY && (!fac || !rho)) ||
                   ((mode == Y_BWD_FULL || mode == Y_FWD_FULL) && !fgrid) ||
                   (mode == Y_BWD_REAL_ADD && !fgrid) || (mode == Y_FWD_REAL && !v);
  if (bad) return QBG_E_INVAL;
  if (g->f4[1]) {
    const int yp = g->yppb_t[pix(prec)] > 0 ? g->yppb_t[pix(prec)] : g->yppb;
    const int ppb = mode == Y_DENSITY ? nbatch : (mode == Y_VMUL ? std::min(yp, nbatch) : 1);
    yops(g, prec)->ypass(g, prec, mode, nbatch, ppb, v, fac, rho, fgrid, scale, s);
    return launched();
  }
  return prec == QBG_PREC_FP64 ? ypass_t<double>(g, mode, nbatch, v, fac, rho, fgrid, scale, s)
                               : ypass_t<float>(g, mode, nbatch, v, fac, rho, fgrid, scale, s);
}
bool zfwd_fin_ok(const qbg_grid* g) { return g && g->f4[2]; }

int zfwd(qbg_grid* g, int prec, int mode, int outmode, double* out, long pstride, long off2,
         int last_has2, int nbatch, hipStream_t s, const ZFin* fin) {
  if (!ws_ok(g, nbatch) || !out) return QBG_E_STATE;
  if (fin && (!g->f4[2] || mode != Z_PAIR || !fin->c || !fin->y || !fin->g2))
    return QBG_E_STATE;
  if (g->f4[2]) {
    g->f4[2]->zfwd(g, prec, mode == Z_SINGLE, outmode == OUT_ACCUM, out, pstride, off2,
                   last_has2, nbatch, s, fin ? *fin : ZFin{});
    return launched();
  }
  return prec == QBG_PREC_FP64
             ? zfwd_t<double>(g, mode, outmode, out, pstride, off2, last_has2, nbatch, s)
             : zfwd_t<float>(g, mode, outmode, out, pstride, off2, last_has2, nbatch, s);
}

}  // namespace qbg

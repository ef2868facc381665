// fft.cpp -- oracle restatement of FourierTransform / BasisMapping transforms.
// TEST INFRASTRUCTURE ONLY (see qbo.h).
//
// The 1D transforms are an own Stockham mixed-radix FFT (the reference's
// FFT library is interchangeable: FFTW2/FFTW3/ESSL/cfftm all compute the
// same unscaled DFT).  The 3D driver restates the sparse schedule and the
// scatter/gather semantics of the reference exactly.
#include "qbo.h"
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>

namespace qbo {

namespace {

struct Plan {
  int n;
  std::vector<int> radix;
  std::vector<cplx> w;  // w[k] = exp(+2 pi i k / n)
};

std::map<int, Plan*>& plan_cache() {
  static std::map<int, Plan*> m;
  return m;
}
std::mutex plan_mutex;

const Plan* get_plan(int n) {
  std::lock_guard<std::mutex> lock(plan_mutex);
  auto& m = plan_cache();
  auto it = m.find(n);
  if (it != m.end()) return it->second;
  Plan* p = new Plan;
  p->n = n;
  int r = n;
  const int rads[] = {4, 2, 3, 5, 7, 11, 13};
  for (int f : rads) {
    while (r % f == 0) {
      p->radix.push_back(f);
      r /= f;
    }
  }
  if (r != 1) {
    // generic remaining prime factors
    for (int f = 17; r > 1; f += 2)
      while (r % f == 0) { p->radix.push_back(f); r /= f; }
  }
  p->w.resize(n);
  for (int k = 0; k < n; k++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * k / n;
    p->w[k] = cplx((double)cosl(a), (double)sinl(a));
  }
  m[n] = p;
  return p;
}

// small DFT of size R: v <- DFT(v), sign +1 synthesis, -1 analysis
inline void small_dft(cplx* v, int R, int sign, const Plan* p) {
  if (R == 2) {
    cplx a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
    return;
  }
  if (R == 4) {
    cplx a0 = v[0] + v[2], a1 = v[0] - v[2];
    cplx b0 = v[1] + v[3], b1 = v[1] - v[3];
    // multiply b1 by sign*i
    cplx jb1 = sign > 0 ? cplx(-b1.imag(), b1.real()) : cplx(b1.imag(), -b1.real());
    v[0] = a0 + b0; v[2] = a0 - b0;
    v[1] = a1 + jb1; v[3] = a1 - jb1;
    return;
  }
  cplx out[32];
  const int n = p->n;
  const int step = n / R;
  for (int k = 0; k < R; k++) {
    cplx s = 0.0;
    for (int j = 0; j < R; j++) {
      int e = (int)(((long)j * k) % R) * step;
      cplx w = p->w[e];
      if (sign < 0) w = std::conj(w);
      s += v[j] * w;
    }
    out[k] = s;
  }
  for (int k = 0; k < R; k++) v[k] = out[k];
}

}  // namespace

// Stockham autosort, decimation in time: at each stage with radix R and
// Ns = product of earlier radices, butterfly j reads x[j + r*n/R] (r<R),
// twiddles by w^(r*(j%Ns)*n/(Ns*R)), and writes y[(j/Ns)*Ns*R + j%Ns + r*Ns].
void fft1d(cplx* x, int n, int stride, int sign) {
  if (n == 1) return;
  const Plan* p = get_plan(n);
  thread_local std::vector<cplx> a, b;
  if ((int)a.size() < n) { a.resize(n); b.resize(n); }
  for (int i = 0; i < n; i++) a[i] = x[(long)i * stride];
  cplx* src = a.data();
  cplx* dst = b.data();
  int Ns = 1;
  for (int R : p->radix) {
    const int nR = n / R;
    for (int j = 0; j < nR; j++) {
      cplx v[32];
      const int jm = j % Ns;
      for (int r = 0; r < R; r++) {
        cplx t = src[j + r * nR];
        const long e = ((long)r * jm * (n / (Ns * R))) % n;
        cplx w = p->w[e];
        if (sign < 0) w = std::conj(w);
        v[r] = t * w;
      }
      small_dft(v, R, sign, p);
      const int base = (j / Ns) * Ns * R + jm;
      for (int r = 0; r < R; r++) dst[base + r * Ns] = v[r];
    }
    std::swap(src, dst);
    Ns *= R;
  }
  for (int i = 0; i < n; i++) x[(long)i * stride] = src[i];
}

void FFT3::init(const Basis& b, int n0, int n1, int n2) {
  basis = &b;
  map.build(b, n0, n1, n2);
}

// z transforms, transpose_bwd, fxy_inv:
// src/FourierTransform.cpp:222-239, :764-827, BasisMapping.cpp:445-467,
// FourierTransform.cpp:640-668 (NOLIB order: x on active rows, then y)
void FFT3::bwd_grid(std::vector<cplx>& zvec, cplx* f) const {
  const int np0 = map.np0, np1 = map.np1, np2 = map.np2, nvec = map.nvec;
  const long np01 = (long)np0 * np1;
#pragma omp parallel for schedule(static)
  for (int iv = 0; iv < nvec; iv++) fft1d(&zvec[(long)iv * np2], np2, 1, +1);
  std::memset((void*)f, 0, sizeof(cplx) * np01 * np2);
#pragma omp parallel for schedule(static)
  for (int iv = 0; iv < nvec; iv++) {
    const long dest = map.zvec_to_val[iv];
    for (int k = 0; k < np2; k++) f[dest + k * np01] = zvec[(long)iv * np2 + k];
  }
  const int nt = map.ntrans0;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < np2; k++) {
    cplx* plane = f + k * np01;
    for (int j = 0; j < nt; j++) fft1d(plane + (long)j * np0, np0, 1, +1);
    for (int j = np1 - nt; j < np1; j++) fft1d(plane + (long)j * np0, np0, 1, +1);
    for (int i = 0; i < np0; i++) fft1d(plane + i, np1, np0, +1);
  }
}

// fxy, transpose_fwd, fz: src/FourierTransform.cpp:202-220, :420-446,
// BasisMapping.cpp:543-561, FourierTransform.cpp:744-752 (scale 1/np012)
void FFT3::fwd_grid(cplx* f, std::vector<cplx>& zvec) const {
  const int np0 = map.np0, np1 = map.np1, np2 = map.np2, nvec = map.nvec;
  const long np01 = (long)np0 * np1;
  const int nt = map.ntrans0;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < np2; k++) {
    cplx* plane = f + k * np01;
    for (int i = 0; i < np0; i++) fft1d(plane + i, np1, np0, -1);
    for (int j = 0; j < nt; j++) fft1d(plane + (long)j * np0, np0, 1, -1);
    for (int j = np1 - nt; j < np1; j++) fft1d(plane + (long)j * np0, np0, 1, -1);
  }
  zvec.assign((size_t)nvec * np2, cplx(0.0));
  const double scale = 1.0 / ((double)np01 * np2);
#pragma omp parallel for schedule(static)
  for (int iv = 0; iv < nvec; iv++) {
    const long src = map.zvec_to_val[iv];
    cplx* z = &zvec[(long)iv * np2];
    for (int k = 0; k < np2; k++) z[k] = f[src + k * np01];
    fft1d(z, np2, 1, -1);
    for (int k = 0; k < np2; k++) z[k] *= scale;
  }
}

// vector_to_zvec, src/BasisMapping.cpp:628-664 (real basis: G and -G;
// complex basis: G only)
void FFT3::backward(const cplx* c, cplx* f) const {
  std::vector<cplx> zvec((size_t)map.nvec * map.np2, cplx(0.0));
  const int ng = basis->size;
  if (basis->real) {
    for (int ig = 0; ig < ng; ig++) {
      zvec[map.ip[ig]] = c[ig];
      zvec[map.im[ig]] = std::conj(c[ig]);
    }
  } else {
    for (int ig = 0; ig < ng; ig++) zvec[map.ip[ig]] = c[ig];
  }
  bwd_grid(zvec, f);
}

// doublevector_to_zvec, src/BasisMapping.cpp:667-700
void FFT3::backward(const cplx* c1, const cplx* c2, cplx* f) const {
  std::vector<cplx> zvec((size_t)map.nvec * map.np2, cplx(0.0));
  const int ng = basis->size;
  for (int ig = 0; ig < ng; ig++) {
    const double a = c1[ig].real(), b = c1[ig].imag();
    const double c = c2[ig].real(), d = c2[ig].imag();
    zvec[map.ip[ig]] = cplx(a - d, b + c);
    zvec[map.im[ig]] = cplx(a + d, c - b);
  }
  bwd_grid(zvec, f);
}

// zvec_to_vector, src/BasisMapping.cpp:703-718
void FFT3::forward(cplx* f, cplx* c) const {
  std::vector<cplx> zvec;
  fwd_grid(f, zvec);
  const int ng = basis->size;
  for (int ig = 0; ig < ng; ig++) c[ig] = zvec[map.ip[ig]];
}

// zvec_to_doublevector, src/BasisMapping.cpp:721-750
void FFT3::forward(cplx* f, cplx* c1, cplx* c2) const {
  std::vector<cplx> zvec;
  fwd_grid(f, zvec);
  const int ng = basis->size;
  for (int ig = 0; ig < ng; ig++) {
    const cplx zp = zvec[map.ip[ig]], zm = zvec[map.im[ig]];
    const double a = zp.real(), b = zp.imag(), c = zm.real(), d = zm.imag();
    c1[ig] = cplx(0.5 * (a + c), 0.5 * (b - d));
    c2[ig] = cplx(0.5 * (b + d), 0.5 * (c - a));
  }
}

}  // namespace qbo

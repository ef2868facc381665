// stepper.cpp -- CPU restatement of one SCF iteration's wavefunction and
// density updates (TEST INFRASTRUCTURE ONLY, see qbo.h):
//   SD / PSD / PSDA wavefunction steppers with the diagonal preconditioner,
//   SlaterDet::ortho_align with DoubleMatrix::polar, SlaterDet::dot, and the
//   charge mixing of BOSampleStepper with the AndersonMixer.
// Coefficients are band-major complex arrays c[n*ngw + g] (the real view of
// band n is 2*ngw doubles), the Gamma-point convention of a real basis: the
// stored half sphere counts twice, except the G = 0 element.
#include "qbo.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace qbo {

namespace {

// a[i + j*n] = alpha sum_k p_i[k] q_j[k] + gamma p_i[0] q_j[0]  (real views
// of length m; DoubleMatrix::gemm('t','n') followed by ger on row 0)
void tn(const double* p, const double* q, long m, int n, double alpha, double gamma, double* a) {
#pragma omp parallel for schedule(static) collapse(2) if ((long)n * n * m > (1L << 22))
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      const double* x = p + (long)i * m;
      const double* y = q + (long)j * m;
      double s = 0.0;
      for (long k = 0; k < m; k++) s += x[k] * y[k];
      a[i + (long)j * n] = alpha * s + gamma * x[0] * y[0];
    }
}

// c (m x n) = a (m x kk) b (kk x n), column major
void nn(const double* a, long m, const double* b, int kk, int n, double* c) {
#pragma omp parallel for schedule(static) if ((long)n * kk * m > (1L << 22))
  for (int j = 0; j < n; j++) {
    double* cj = c + (long)j * m;
    for (long r = 0; r < m; r++) cj[r] = 0.0;
    for (int l = 0; l < kk; l++) {
      const double f = b[l + (long)j * kk];
      const double* al = a + (long)l * m;
      for (long r = 0; r < m; r++) cj[r] += al[r] * f;
    }
  }
}

// lower Cholesky of the n x n column-major s in place; 0 or failing pivot+1
int potrf_lower(double* s, int n) {
  for (int j = 0; j < n; j++) {
    double d = s[j + (long)j * n];
    for (int k = 0; k < j; k++) d -= s[j + (long)k * n] * s[j + (long)k * n];
    if (!(d > 0.0)) return j + 1;
    d = std::sqrt(d);
    s[j + (long)j * n] = d;
    for (int i = j + 1; i < n; i++) {
      double t = s[i + (long)j * n];
      for (int k = 0; k < j; k++) t -= s[i + (long)k * n] * s[j + (long)k * n];
      s[i + (long)j * n] = t / d;
    }
  }
  return 0;
}

// solve L X = B (trans = 0) or L^T X = B (trans = 1) in place, L lower
// (dtrtrs('l', 'n'/'t', 'n'), src/Matrix.cpp:1959-1989)
void trtrs_lower(const double* l, int n, double* b, int nrhs, int trans) {
#pragma omp parallel for schedule(static) if ((long)n * n * nrhs > (1L << 22))
  for (int c = 0; c < nrhs; c++) {
    double* x = b + (long)c * n;
    if (!trans) {
      for (int i = 0; i < n; i++) {
        double s = x[i];
        for (int k = 0; k < i; k++) s -= l[i + (long)k * n] * x[k];
        x[i] = s / l[i + (long)i * n];
      }
    } else {
      for (int i = n - 1; i >= 0; i--) {
        double s = x[i];
        for (int k = i + 1; k < n; k++) s -= l[k + (long)i * n] * x[k];
        x[i] = s / l[i + (long)i * n];
      }
    }
  }
}

// inverse of the general n x n column-major a in place, Gauss-Jordan with
// partial pivoting; false if singular (DoubleMatrix::inverse,
// src/Matrix.cpp: dgetrf + dgetri)
bool inverse(double* a, int n) {
  std::vector<double> m((size_t)n * n);
  // row-major working copy
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) m[(size_t)i * n + j] = a[i + (long)j * n];
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::fabs(m[(size_t)i * n + k]) > std::fabs(m[(size_t)p * n + k])) p = i;
    if (m[(size_t)p * n + k] == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; j++) {
        std::swap(m[(size_t)p * n + j], m[(size_t)k * n + j]);
        std::swap(inv[(size_t)p * n + j], inv[(size_t)k * n + j]);
      }
    const double d = 1.0 / m[(size_t)k * n + k];
    for (int j = 0; j < n; j++) { m[(size_t)k * n + j] *= d; inv[(size_t)k * n + j] *= d; }
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      const double f = m[(size_t)i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) {
        m[(size_t)i * n + j] -= f * m[(size_t)k * n + j];
        inv[(size_t)i * n + j] -= f * inv[(size_t)k * n + j];
      }
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) a[i + (long)j * n] = inv[(size_t)i * n + j];
  return true;
}

// eigenvalues w (ascending) and eigenvectors (columns of v) of the symmetric
// n x n column-major a, cyclic Jacobi (stands in for dsyev in AndersonMixer)
void syev(const double* a_in, int n, double* w, double* v) {
  std::vector<double> a(a_in, a_in + (size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) v[i + j * n] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += a[i + j * n] * a[i + j * n];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[p + q * n];
        if (apq == 0.0) continue;
        const double th = (a[q + q * n] - a[p + p * n]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = a[k + p * n], akq = a[k + q * n];
          a[k + p * n] = c * akp - s * akq;
          a[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = a[p + k * n], aqk = a[q + k * n];
          a[p + k * n] = c * apk - s * aqk;
          a[q + k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = v[k + p * n], vkq = v[k + q * n];
          v[k + p * n] = c * vkp - s * vkq;
          v[k + q * n] = s * vkp + c * vkq;
        }
      }
  }
  // sort ascending
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x + x * n] < a[y + y * n]; });
  std::vector<double> vv(v, v + (size_t)n * n);
  for (int k = 0; k < n; k++) {
    w[k] = a[ord[k] + ord[k] * n];
    for (int i = 0; i < n; i++) v[i + k * n] = vv[i + ord[k] * n];
  }
}

}  // namespace

// SlaterDet::dot, real basis (src/SlaterDet.cpp:1122-1148):
// 2 sum c.d - sum_n Re c_n(0) Re d_n(0)
double sd_dot(const cplx* c, const cplx* d, int ngw, int nst) {
  const double* p = reinterpret_cast<const double*>(c);
  const double* q = reinterpret_cast<const double*>(d);
  const long m = 2L * ngw;
  double s = 0.0, s0 = 0.0;
  for (int n = 0; n < nst; n++) {
    const double* x = p + n * m;
    const double* y = q + n * m;
    double t = 0.0;
    for (long k = 0; k < m; k++) t += x[k] * y[k];
    s += t;
    s0 += x[0] * y[0];
  }
  return 2.0 * s - s0;
}

// Preconditioner::update (src/Preconditioner.cpp:88-143): per-band kinetic
// energy with the reference's factors: buf_n = 2 (sum |c|^2 - |c(0)|^2 / 2),
// buf_e = 2 sum |c|^2 g2, ekin_n = 0.5 buf_n > 0 ? 0.5 buf_e / buf_n : 0
void prec_ekin(const cplx* c, const double* g2, int ngw, int nst, double* ekin) {
  for (int n = 0; n < nst; n++) {
    double sn = 0.0, se = 0.0;
    for (int g = 0; g < ngw; g++) {
      const double p2 = std::norm(c[(long)n * ngw + g]);
      sn += p2;
      se += p2 * g2[g];
    }
    sn -= 0.5 * std::norm(c[(long)n * ngw]);
    const double bn = 2.0 * sn, be = 2.0 * se;
    ekin[n] = 0.5 * bn > 0.0 ? 0.5 * be / bn : 0.0;
  }
}

// Preconditioner::diag (src/Preconditioner.cpp:48-85), basic adaptive form
// (ecutprec == 0) or fixed ecutprec; no confinement potential (fstress = 0)
double prec_diag(double ekin_n, double g2, double ecutprec) {
  const double e = 0.5 * g2;
  if (ecutprec == 0.0) {
    if (ekin_n == 0.0) ekin_n = 1.0;
    return e < ekin_n ? 0.5 / ekin_n : 0.5 / e;
  }
  return e < ecutprec ? 0.5 / ecutprec : 0.5 / e;
}

// residual of the PSD/PSDA steppers (src/PSDWavefunctionStepper.cpp:47-65):
// a = 2 c^T d - c(0) d(0)^T; d <- d - c a.  a (nst x nst) returned if non-null
void residual(const cplx* c, cplx* d, int ngw, int nst, double* a_out) {
  const long m = 2L * ngw;
  const double* p = reinterpret_cast<const double*>(c);
  double* q = reinterpret_cast<double*>(d);
  std::vector<double> a((size_t)nst * nst), ca((size_t)m * nst);
  tn(p, q, m, nst, 2.0, -1.0, a.data());
  nn(p, m, a.data(), nst, nst, ca.data());
  for (size_t i = 0; i < ca.size(); i++) q[i] -= ca[i];
  if (a_out) std::memcpy(a_out, a.data(), 8 * a.size());
}

// DoubleMatrix::polar (src/Matrix.cpp:2446-2517): the orthogonal polar factor
// of x (n x n, column major) by Bjorck-Bowie steps while ||I - x^T x||_F < 1,
// Higham steps x <- (x + x^-T)/2 otherwise.  Returns the number of
// iterations, or -1 if an inverse failed.
int polar(double* x, int n, double tol, int maxiter) {
  const size_t nn2 = (size_t)n * n;
  std::vector<double> q(nn2), t(nn2), qt(nn2), xp(nn2);
  double qnrm2 = 1e300;
  int iter = 0;
  while (iter < maxiter && qnrm2 > tol) {
    // q = I - x^T x
    tn(x, x, n, n, -1.0, 0.0, q.data());
    for (int i = 0; i < n; i++) q[i + (long)i * n] += 1.0;
    double s = 0.0;
    for (size_t i = 0; i < nn2; i++) s += q[i] * q[i];
    qnrm2 = std::sqrt(s);
    if (qnrm2 < 1.0) {
      // x <- x (I + q/2 (I + 3q/4))
      for (size_t i = 0; i < nn2; i++) t[i] = 0.75 * q[i];
      for (int i = 0; i < n; i++) t[i + (long)i * n] += 1.0;
      nn(q.data(), n, t.data(), n, n, qt.data());
      for (size_t i = 0; i < nn2; i++) t[i] = 0.5 * qt[i];
      for (int i = 0; i < n; i++) t[i + (long)i * n] += 1.0;
      nn(x, n, t.data(), n, n, xp.data());
      std::memcpy(x, xp.data(), 8 * nn2);
    } else {
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) t[i + (long)j * n] = x[j + (long)i * n];
      if (!inverse(t.data(), n)) return -1;
      for (size_t i = 0; i < nn2; i++) x[i] = 0.5 * (x[i] + t[i]);
    }
    iter++;
  }
  return iter;
}

// SlaterDet::ortho_align, real basis (src/SlaterDet.cpp:768-876): c becomes
// orthonormal and aligned with cl: c <- c L^-T Q, L L^T = 2 c^T c - c0 c0^T,
// Q the polar factor of L^-1 (2 c^T cl - c0 cl0^T).  0 or the failing pivot+1.
int ortho_align(cplx* c, const cplx* cl, int ngw, int nst) {
  cleanup(c, ngw, nst);
  const long m = 2L * ngw;
  double* p = reinterpret_cast<double*>(c);
  const double* pl = reinterpret_cast<const double*>(cl);
  std::vector<double> l((size_t)nst * nst), x((size_t)nst * nst), ct(p, p + m * nst);
  tn(p, p, m, nst, 2.0, -1.0, l.data());
  const int info = potrf_lower(l.data(), nst);
  if (info) return info;
  tn(p, pl, m, nst, 2.0, -1.0, x.data());
  trtrs_lower(l.data(), nst, x.data(), nst, 0);
  if (polar(x.data(), nst, 1.e-6, 20) < 0) return -1;
  trtrs_lower(l.data(), nst, x.data(), nst, 1);
  nn(ct.data(), m, x.data(), nst, nst, p);
  return 0;
}

// One update of the wavefunction steppers.  kind 0 SD (param = dt2bye;
// src/SDWavefunctionStepper.cpp:36-51), 1 PSD (param = ecutprec;
// src/PSDWavefunctionStepper.cpp:41-119), 2 PSDA (param = ecutprec;
// src/PSDAWavefunctionStepper.cpp:45-247, which orthogonalizes by
// ortho_align).  d holds H c on entry and the (preconditioned) descent
// direction on exit.  PSDA keeps c_last, d_last and the extrapolate flag
// between calls.  Returns 0 or a Cholesky failure (pivot+1).
int wf_update(int kind, double param, const double* g2, cplx* c, cplx* d, int ngw, int nst,
              cplx* c_last, cplx* d_last, int* extrapolate) {
  const long m = 2L * ngw;
  double* p = reinterpret_cast<double*>(c);
  double* q = reinterpret_cast<double*>(d);
  if (kind == 0) {
    for (long i = 0; i < m * nst; i++) p[i] -= param * q[i];
    return gram(c, ngw, nst);
  }
  residual(c, d, ngw, nst, nullptr);
  std::vector<double> ekin(nst);
  prec_ekin(c, g2, ngw, nst, ekin.data());
  if (kind == 1) {
    for (int n = 0; n < nst; n++)
      for (int g = 0; g < ngw; g++) {
        const double f = prec_diag(ekin[n], g2[g], param);
        const long i = n * m + 2L * g;
        const double dr = f * q[i], di = f * q[i + 1];
        p[i] -= dr;
        p[i + 1] -= di;
      }
    return gram(c, ngw, nst);
  }
  // PSDA
  double* pl = reinterpret_cast<double*>(c_last);
  double* ql = reinterpret_cast<double*>(d_last);
  for (int n = 0; n < nst; n++)
    for (int g = 0; g < ngw; g++) {
      const double f = prec_diag(ekin[n], g2[g], param);
      const long i = n * m + 2L * g;
      q[i] = -f * q[i];
      q[i + 1] = -f * q[i + 1];
    }
  if (*extrapolate) {
    double a = 0.0, b = 0.0;
    for (long i = 0; i < m * nst; i++) {
      const double f = q[i], df = f - ql[i];
      a += f * df;
      b += df * df;
    }
    a *= 2.0;
    b *= 2.0;
    for (int n = 0; n < nst; n++) {
      const long i = n * m;
      const double f0 = q[i], f1 = q[i + 1];
      const double d0 = f0 - ql[i], d1 = f1 - ql[i + 1];
      a -= f0 * d0 + f1 * d1;
      b -= d0 * d0 + d1 * d1;
    }
    double theta = 0.0;
    if (b != 0.0) theta = -a / b;
    if (theta < -1.0) theta = 0.0;
    theta = std::min(2.0, theta);
    for (long i = 0; i < m * nst; i++) {
      const double x = p[i], xl = pl[i], xbar = x + theta * (x - xl);
      const double f = q[i], fl = ql[i], fbar = f + theta * (f - fl);
      p[i] = xbar + fbar;
      pl[i] = x;
      ql[i] = f;
    }
  } else {
    for (long i = 0; i < m * nst; i++) {
      const double x = p[i], f = q[i];
      p[i] = x + f;
      pl[i] = x;
      ql[i] = f;
    }
  }
  const int info = ortho_align(c, c_last, ngw, nst);
  *extrapolate = 1;
  return info;
}

// ------------------------------------------------------------- AndersonMixer
// src/AndersonMixer.cpp:25-270 (diag_ = true, eig_ratio_ = 0)
void Anderson::restart() { n_ = -1; k_ = -1; }

void Anderson::update(const double* x, const double* f, double* xbar, double* fbar) {
  k_ = (k_ + 1) % (nmax + 1);
  if (n_ < nmax) n_++;
  std::memcpy(&xs[(size_t)k_ * m], x, 8 * m);
  std::memcpy(&fs[(size_t)k_ * m], f, 8 * m);
  std::vector<double> theta(std::max(n_, 1), 0.0);
  const double* fk = &fs[(size_t)k_ * m];
  if (n_ > 0) {
    std::vector<double> a((size_t)n_ * n_, 0.0), b(n_), w(n_), u((size_t)n_ * n_);
    for (int i = 0; i < n_; i++) {
      const double* fi = &fs[(size_t)((k_ - i + nmax) % (nmax + 1)) * m];
      for (int j = 0; j <= i; j++) {
        const double* fj = &fs[(size_t)((k_ - j + nmax) % (nmax + 1)) * m];
        double s = 0.0;
        for (long l = 0; l < m; l++) s += (fk[l] - fi[l]) * (fk[l] - fj[l]);
        a[i + j * n_] = s;
        a[j + i * n_] = s;
      }
      double bs = 0.0;
      for (long l = 0; l < m; l++) bs += (fk[l] - fi[l]) * fk[l];
      b[i] = bs;
    }
    syev(a.data(), n_, w.data(), u.data());
    const double eig_ratio = 0.0;
    for (int k = 0; k < n_; k++) {
      if (w[k] > eig_ratio * w[n_ - 1]) {
        double ub = 0.0;
        for (int j = 0; j < n_; j++) ub += u[j + k * n_] * b[j];
        for (int i = 0; i < n_; i++) theta[i] += (ub / w[k]) * u[i + k * n_];
      }
    }
  }
  const double* xk = &xs[(size_t)k_ * m];
  for (long l = 0; l < m; l++) { fbar[l] = fk[l]; xbar[l] = xk[l]; }
  for (int i = 0; i < n_; i++) {
    const double* fi = &fs[(size_t)((k_ - i + nmax) % (nmax + 1)) * m];
    const double* xi = &xs[(size_t)((k_ - i + nmax) % (nmax + 1)) * m];
    for (long l = 0; l < m; l++) {
      fbar[l] -= theta[i] * (fk[l] - fi[l]);
      xbar[l] -= theta[i] * (xk[l] - xi[l]);
    }
  }
}

// ------------------------------------------------------------- charge mixing
// BOSampleStepper (src/BOSampleStepper.cpp:319-406 weights, :810-879 mixing)
void Mixer::init(const System& s, double alpha_, double rc_kerker, int ndim, double rc1) {
  alpha = alpha_;
  const int ng = s.vb.size;
  wkerker.assign(ng, 1.0);
  wls.assign(ng, 1.0);
  const double* g2 = s.vb.g2.data();
  if (rc1 != 0.0) {
    const double q1 = 2.0 * M_PI / rc1;
    for (int i = 0; i < ng; i++) wls[i] = g2[i] != 0.0 ? std::sqrt(g2[i] / (g2[i] + q1 * q1)) : 1.0;
  }
  if (rc_kerker > 0.0) {
    const double q0 = 2 * M_PI / rc_kerker, q02 = q0 * q0;
    for (int i = 0; i < ng; i++) wkerker[i] = g2[i] != 0.0 ? g2[i] / (g2[i] + q02) : 1.0;
  }
  anderson = ndim > 0;
  // both spins in one vector (src/BOSampleStepper.cpp:318-327)
  am.m = 2L * s.nspin * ng;
  am.nmax = ndim;
  am.xs.assign((size_t)(ndim + 1) * am.m, 0.0);
  am.fs.assign((size_t)(ndim + 1) * am.m, 0.0);
  am.restart();
  rhog_in.assign((size_t)s.nspin * ng, cplx(0.0));
  itscf = 0;
}

void Mixer::restart() { am.restart(); itscf = 0; }

void Mixer::update(System& s) {
  const int ng = s.vb.size, ns = s.nspin;
  const size_t n = (size_t)ns * ng;
  // the mixed vector: rhog of each spin, spin-major
  std::vector<cplx> rg(n);
  if (ns == 1) {
    std::copy(s.rhog.begin(), s.rhog.begin() + ng, rg.begin());
  } else {
    s.rhog_of(s.rhor.data(), rg.data());
    s.rhog_of(s.rhor_dn.data(), rg.data() + ng);
  }
  if (itscf == 0) {
    rhog_in = rg;
  } else {
    std::vector<cplx> drhog(n);
    for (size_t i = 0; i < n; i++) drhog[i] = rg[i] - rhog_in[i];
    if (anderson) {
      for (size_t i = 0; i < n; i++) drhog[i] /= wls[i % ng];
      std::vector<cplx> rhobar(n), drhobar(n);
      am.update(reinterpret_cast<const double*>(rhog_in.data()),
                reinterpret_cast<const double*>(drhog.data()),
                reinterpret_cast<double*>(rhobar.data()), reinterpret_cast<double*>(drhobar.data()));
      for (size_t i = 0; i < n; i++) {
        drhobar[i] *= wls[i % ng];
        rhog_in[i] = rhobar[i] + alpha * drhobar[i] * wkerker[i % ng];
      }
    } else {
      for (size_t i = 0; i < n; i++) rhog_in[i] += alpha * drhog[i] * wkerker[i % ng];
    }
    if (ns == 1) {
      for (int i = 0; i < ng; i++) s.rhog[i] = rhog_in[i];
      s.update_rhor();
    } else {
      // rhor of each spin; rhog = rhog[0] + rhog[1] (src/EnergyFunctional.cpp:256-261)
      s.rhor_of(rhog_in.data(), s.rhor.data());
      s.rhor_of(rhog_in.data() + ng, s.rhor_dn.data());
      for (int i = 0; i < ng; i++) s.rhog[i] = rhog_in[i] + rhog_in[i + ng];
    }
  }
  itscf++;
}

// ChargeDensity::update_rhor (src/ChargeDensity.cpp:201-221):
// rhor = Re(vft.backward(rhog)) / omega
void System::rhog_of(const double* r, cplx* g) {
  const long N = vft.np012();
  std::vector<cplx> t(N);
  for (long i = 0; i < N; i++) t[i] = cplx(cell.volume * r[i], 0.0);
  vft.forward(t.data(), g);
}

void System::rhor_of(const cplx* g, double* r) {
  std::vector<cplx> f(vft.np012());
  vft.backward(g, f.data());
  const double oi = 1.0 / cell.volume;
  for (size_t i = 0; i < f.size(); i++) r[i] = f[i].real() * oi;
}

void System::update_rhor() {
  std::vector<cplx> f(vft.np012());
  vft.backward(rhog.data(), f.data());
  const double oi = 1.0 / cell.volume;
  for (size_t i = 0; i < rhor.size(); i++) rhor[i] = f[i].real() * oi;
}

}  // namespace qbo

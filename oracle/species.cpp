// species.cpp -- oracle restatement of the SLPP species tables, splines,
// sine/cosine transforms and the LDA functional.
// TEST INFRASTRUCTURE ONLY (see qbo.h).
#include "qbo.h"
#include <cmath>
#include <cassert>
#include <stdexcept>

namespace qbo {

// src/Species.cpp:34-48
static double simpsn(int n, const double* t) {
  const double c0 = 17.0 / 48.0, c1 = 59.0 / 48.0, c2 = 43.0 / 48.0,
               c3 = 49.0 / 48.0;
  double sum = c0 * (t[0] + t[n - 1]) + c1 * (t[1] + t[n - 2]) +
               c2 * (t[2] + t[n - 3]) + c3 * (t[3] + t[n - 4]);
  for (int i = 4; i < n - 4; i++) sum += t[i];
  return sum;
}

// Table of sin/cos(pi m / n), m in [0, 2n), in extended precision.
static void trig_table(int n, std::vector<double>& s, std::vector<double>& c) {
  s.resize(2 * n);
  c.resize(2 * n);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int m = 0; m < 2 * n; m++) {
    long double a = pi * m / n;
    s[m] = (double)sinl(a);
    c[m] = (double)cosl(a);
  }
}

// src/sinft.cpp:50-105: y[i] <- sum_{j=1}^{n-1} y[j] sin(pi i j / n),
// evaluated here as the direct sum it equals (odd extension, 2n-point DFT).
void sinft(int n, double* y) {
  std::vector<double> s, c, out(n);
  trig_table(n, s, c);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double acc = 0.0;
    long m = 0;
    for (int j = 1; j < n; j++) {
      m += i; if (m >= 2 * n) m -= 2 * n;
      acc += y[j] * s[m];
    }
    out[i] = acc;
  }
  for (int i = 0; i < n; i++) y[i] = out[i];
}

// src/sinft.cpp:108-174: y[i] <- y0/2 + (-1)^i y[n]/2 + sum_{j=1}^{n-1}
// y[j] cos(pi i j / n), i < n (even extension, 2n-point DFT; y has n+1
// entries and y[n] is left as is).
void cosft1(int n, double* y) {
  std::vector<double> s, c, out(n);
  trig_table(n, s, c);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double acc = 0.5 * y[0] + 0.5 * ((i % 2 == 0) ? y[n] : -y[n]);
    long m = 0;
    for (int j = 1; j < n; j++) {
      m += i; if (m >= 2 * n) m -= 2 * n;
      acc += y[j] * c[m];
    }
    out[i] = acc;
  }
  for (int i = 0; i < n; i++) y[i] = out[i];
}

// src/spline.cpp:21-42
static void tridsolve(int n, double* d, double* e, double* f, double* x) {
  for (int i = 1; i < n; i++) {
    f[i] /= d[i - 1];
    d[i] -= f[i] * e[i - 1];
  }
  for (int i = 1; i < n; i++) x[i] -= f[i] * x[i - 1];
  x[n - 1] /= d[n - 1];
  for (int i = n - 2; i >= 0; i--) x[i] = (x[i] - e[i] * x[i + 1]) / d[i];
}

// src/spline.cpp:44-108
void spline(int n, const double* x, const double* y, double yp_left,
            double yp_right, int bcnat_left, int bcnat_right, double* y2) {
  const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
  std::vector<double> d(n), e(n), f(n);
  if (bcnat_left == 0) {
    const double h = x[1] - x[0];
    d[0] = third * h; e[0] = sixth * h; f[0] = 0.0;
    y2[0] = (y[1] - y[0]) / h - yp_left;
  } else {
    d[0] = 1.0; e[0] = 0.0; f[0] = 0.0; y2[0] = 0.0;
  }
  if (bcnat_right == 0) {
    const double h = x[n - 1] - x[n - 2];
    d[n - 1] = third * h; e[n - 1] = 0.0; f[n - 1] = sixth * h;
    y2[n - 1] = yp_right - (y[n - 1] - y[n - 2]) / h;
  } else {
    d[n - 1] = 1.0; e[n - 1] = 0.0; f[n - 1] = 0.0; y2[n - 1] = 0.0;
  }
  for (int i = 1; i < n - 1; i++) {
    const double hp = x[i + 1] - x[i], hm = x[i] - x[i - 1];
    d[i] = third * (hp + hm);
    e[i] = sixth * hp;
    f[i] = sixth * hm;
    y2[i] = (y[i + 1] - y[i]) / hp - (y[i] - y[i - 1]) / hm;
  }
  tridsolve(n, d.data(), e.data(), f.data(), y2);
}

static inline void bracket(int n, const double* xa, double x, int& kl, int& kh) {
  kl = 0; kh = n - 1;
  while (kh - kl > 1) {
    const int k = (kh + kl) / 2;
    if (xa[k] > x) kh = k; else kl = k;
  }
}

// src/spline.cpp:110-136
void splint(int n, const double* xa, const double* ya, const double* y2a,
            double x, double* y) {
  int kl, kh;
  bracket(n, xa, x, kl, kh);
  const double h = xa[kh] - xa[kl];
  const double a = (xa[kh] - x) / h, b = (x - xa[kl]) / h;
  *y = a * ya[kl] + b * ya[kh] +
       h * h * (1.0 / 6.0) * ((a * a * a - a) * y2a[kl] + (b * b * b - b) * y2a[kh]);
}

// src/spline.cpp:138-168
void splintd(int n, const double* xa, const double* ya, const double* y2a,
             double x, double* y, double* dy) {
  int kl, kh;
  bracket(n, xa, x, kl, kh);
  const double h = xa[kh] - xa[kl];
  const double a = (xa[kh] - x) / h, b = (x - xa[kl]) / h;
  *y = a * ya[kl] + b * ya[kh] +
       h * h * (1.0 / 6.0) * ((a * a * a - a) * y2a[kl] + (b * b * b - b) * y2a[kh]);
  *dy = (ya[kh] - ya[kl]) / h +
        h * (((1.0 / 6.0) - 0.5 * a * a) * y2a[kl] + (0.5 * b * b - (1.0 / 6.0)) * y2a[kh]);
}

// src/Species.cpp:510-1053 (SLPP; one projector per l, l <= 3)
void Species::initialize() {
  const double fpi = 4.0 * M_PI;
  const int np = (int)vlocal.size();
  if (lmax > 3) throw std::runtime_error("oracle species: l > 3");
  nop = lmax + 1;
  nlm = 0;
  for (int l = 0; l <= lmax; l++) nlm += 2 * l + 1;
  if (quad()) {
    // projectors per nonlocal channel and quadrature point (the projector
    // count npr of src/NonLocalPotential.cpp:93-100)
    nlm = 0;
    for (int l = 0; l <= lmax; l++)
      if (l != llocal) {
        if (l > 2) throw std::runtime_error("semilocal quadrature: nonlocal l > 2");
        nlm += (2 * l + 1) * nquad;
      }
  }
  ndft = 1;
  while (ndft * deltar < 40.0) ndft *= 2;
  rps.resize(ndft);
  for (int i = 0; i < ndft; i++) rps[i] = i * deltar;
  std::vector<std::vector<double>> phi(nop, std::vector<double>(ndft, 0.0));
  if (!quad())
    for (int iop = 0; iop < nop; iop++)
      for (int i = 0; i < np; i++) phi[iop][i] = proj[iop][i];
  std::vector<double> vlocr(ndft), fint(ndft);
  for (int i = 0; i < np; i++) vlocr[i] = vlocal[i];
  for (int i = np; i < ndft; i++) vlocr[i] = -zval / rps[i];
  const double ztot = zval;
  // :648-655 subtract the smeared-charge long range part
  vlocr[0] += (ztot / rcps) * M_2_SQRTPI;
  for (int i = 1; i < ndft; i++) vlocr[i] += (ztot / rps[i]) * std::erf(rps[i] / rcps);
  for (int i = 0; i < ndft; i++) vlocr[i] *= fpi * rps[i] * deltar;
  for (int i = 0; i < ndft; i++) fint[i] = vlocr[i] * rps[i];
  const double v0 = simpsn(ndft, fint.data());
  vlocg = vlocr;
  sinft(ndft, vlocg.data());
  gspl.resize(ndft);
  gspl[0] = 0.0;
  vlocg[0] = v0;
  const double fac = M_PI / (ndft * deltar);
  for (int i = 1; i < ndft; i++) {
    gspl[i] = i * fac;
    vlocg[i] /= gspl[i];
  }
  vlocg2.resize(ndft);
  spline(ndft, gspl.data(), vlocg.data(), 0.0, 0.0, 0, 1, vlocg2.data());
  // core density (Species::initialize_nlcc, :1056-1100, and :698-717):
  // nlcc(G) = 4 pi / G sum_r sin(Gr) r dr nlcc(r), the G = 0 value the
  // Simpson integral of 4 pi r^2 nlcc(r)
  nlccg.clear();
  nlccg2.clear();
  zcore = 0.0;
  if (has_nlcc()) {
    nlccg.assign(ndft, 0.0);
    for (int i = 0; i < (int)nlcc.size() && i < ndft; i++)
      nlccg[i] = fpi * rps[i] * deltar * nlcc[i];
    for (int i = 0; i < ndft; i++) fint[i] = nlccg[i] * rps[i];
    zcore = simpsn(ndft, fint.data());
    sinft(ndft, nlccg.data());
    nlccg[0] = zcore;
    for (int i = 1; i < ndft; i++) nlccg[i] /= gspl[i];
    nlccg2.resize(ndft);
    spline(ndft, gspl.data(), nlccg.data(), 0.0, 0.0, 0, 1, nlccg2.data());
  }

  vnlg.assign(nop, std::vector<double>(ndft + 1, 0.0));
  vnlg2.assign(nop, std::vector<double>(ndft + 1, 0.0));
  if (quad()) {
    // semilocal potentials extended by -Z/r and splined with v'(0) = 0 and
    // v'(r_max) = Z/r_max^2 (src/Species.cpp:162-184, dvpsr :1120-1131);
    // trapezoidal points and weights (src/NonLocalPotential.cpp:119-131)
    const double h = rquad / (nquad + 1);
    rq.resize(nquad);
    for (int q = 0; q < nquad; q++) rq[q] = (q + 1) * h;
    wql.assign(lmax + 1, std::vector<double>(nquad, 0.0));
    std::vector<std::vector<double>> vs(lmax + 1, std::vector<double>(ndft)),
        vs2(lmax + 1, std::vector<double>(ndft));
    const double dvdr = zval / (rps[ndft - 1] * rps[ndft - 1]);
    for (int l = 0; l <= lmax; l++) {
      for (int i = 0; i < ndft; i++)
        vs[l][i] = i < (int)vsl[l].size() ? vsl[l][i] : -zval / rps[i];
      spline(ndft, rps.data(), vs[l].data(), 0.0, dvdr, 0, 0, vs2[l].data());
    }
    auto vpsr = [&](int l, double r) {
      if (r > rps[ndft - 1]) return 0.0;
      double v, dv;
      splintd(ndft, rps.data(), vs[l].data(), vs2[l].data(), r, &v, &dv);
      return v;
    };
    for (int l = 0; l <= lmax; l++) {
      if (l == llocal) continue;
      for (int q = 0; q < nquad; q++) wql[l][q] = (vpsr(l, rq[q]) - vpsr(llocal, rq[q])) * h;
    }
    return;
  }
  for (int iop = 0; iop < nop; iop++) {
    std::vector<double> vnlr(ndft);
    for (int i = 0; i < ndft; i++) vnlr[i] = fpi * deltar * phi[iop][i];
    std::vector<double>& v = vnlg[iop];
    if (iop == 0) {  // :782-811
      for (int i = 0; i < ndft; i++) fint[i] = vnlr[i] * rps[i] * rps[i];
      const double w0 = simpsn(ndft, fint.data());
      for (int i = 0; i < ndft; i++) v[i] = vnlr[i] * rps[i];
      sinft(ndft, v.data());
      v[0] = w0;
      for (int i = 1; i < ndft; i++) v[i] /= gspl[i];
      spline(ndft, gspl.data(), v.data(), 0.0, 0.0, 0, 1, vnlg2[iop].data());
    } else if (iop == 2) {  // :872-940, j_2 in three transforms; 0 at G = 0
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = vnlr[i] / rps[i];
      sinft(ndft, v.data());
      fint[0] = 0.0;
      for (int i = 1; i < ndft; i++) fint[i] = 3.0 * v[i] / (gspl[i] * gspl[i] * gspl[i]);
      for (int i = 0; i < ndft; i++) v[i] = vnlr[i] * rps[i];
      sinft(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] += -v[i] / gspl[i];
      for (int i = 0; i < ndft; i++) v[i] = vnlr[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] += -3.0 * v[i] / (gspl[i] * gspl[i]);
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = fint[i];
      spline(ndft, gspl.data(), v.data(), 0.0, 0.0, 0, 1, vnlg2[iop].data());
    } else if (iop == 3) {  // :948-1046, j_3 in four transforms; 0 at G = 0
      // 15/G^4 sum sin(Gr) w/r^2 - 6/G^2 sum sin(Gr) w + 1/G sum cos(Gr) r w
      // - 15/G^3 sum cos(Gr) w/r
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = vnlr[i] / (rps[i] * rps[i]);
      sinft(ndft, v.data());
      fint[0] = 0.0;
      for (int i = 1; i < ndft; i++) {
        const double g = gspl[i];
        fint[i] = 15.0 * v[i] / (g * g * g * g);
      }
      for (int i = 0; i < ndft; i++) v[i] = vnlr[i];
      sinft(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] += -6.0 * v[i] / (gspl[i] * gspl[i]);
      for (int i = 0; i < ndft; i++) v[i] = vnlr[i] * rps[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] += v[i] / gspl[i];
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = vnlr[i] / rps[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      for (int i = 1; i < ndft; i++) {
        const double g = gspl[i];
        fint[i] += -15.0 * v[i] / (g * g * g);
      }
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = fint[i];
      spline(ndft, gspl.data(), v.data(), 0.0, 0.0, 0, 1, vnlg2[iop].data());
    } else {  // l == 1, :812-855
      const double w0 = 0.0;
      for (int i = 0; i < ndft; i++) v[i] = vnlr[i];
      sinft(ndft, v.data());
      fint[0] = w0;
      for (int i = 1; i < ndft; i++) fint[i] = v[i] / (gspl[i] * gspl[i]);
      for (int i = 0; i < ndft; i++) v[i] = vnlr[i] * rps[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      v[0] = w0;
      for (int i = 1; i < ndft; i++) v[i] = fint[i] - v[i] / gspl[i];
      spline(ndft, gspl.data(), v.data(), 0.0, 0.0, 1, 1, vnlg2[iop].data());
    }
  }
}

void Species::dvlocg(double g, double& v, double& dv) const {
  if (g > gspl[ndft - 1]) { v = 0.0; dv = 0.0; return; }
  splintd(ndft, gspl.data(), vlocg.data(), vlocg2.data(), g, &v, &dv);
}

void Species::dvnlg(int iop, double g, double& v, double& dv) const {
  if (g > gspl[ndft - 1]) { v = 0.0; dv = 0.0; return; }
  splintd(ndft, gspl.data(), vnlg[iop].data(), vnlg2[iop].data(), g, &v, &dv);
}

// 4 pi j_l(G r_q) r_q and its G derivative, the expressions of
// src/NonLocalPotential.cpp:317-420 (l = 0, with the G = 0 limit 4 pi r),
// :560-580 (l = 1) and :838-855 (l = 2)
void Species::quad_radial(int l, int q, double g, double& v, double& dv) const {
  const double fpi = 4.0 * M_PI, r = rq[q], z = g * r;
  v = 0.0;
  dv = 0.0;
  if (l == 0) {
    if (g == 0.0) { v = fpi * r; return; }
    const double gi = 1.0 / g, s = std::sin(z), c = std::cos(z);
    v = fpi * s * gi;
    dv = fpi * gi * (r * c - s * gi);
    return;
  }
  if (z == 0.0) return;
  const double zi = 1.0 / z, c = std::cos(z), s = std::sin(z);
  if (l == 1) {
    const double j1 = (s * zi - c) * zi;
    const double dj1 = (2.0 * z * c + (z * z - 2.0) * s) * zi * zi * zi;
    v = fpi * j1 * r;
    dv = fpi * dj1 * r * r;
  } else {
    const double j2 = ((3.0 * zi * zi - 1.0) * s - 3.0 * zi * c) * zi;
    const double z2 = z * z;
    const double dj2 = ((4.0 * z2 - 9.0) * s + z * (9.0 - z2) * c) / (z2 * z2);
    v = fpi * j2 * r;
    dv = fpi * dj2 * r * r;
  }
}

void Species::drhocoreg(double g, double& rho, double& drho) const {
  if (!has_nlcc() || g > gspl[ndft - 1]) { rho = 0.0; drho = 0.0; return; }
  splintd(ndft, gspl.data(), nlccg.data(), nlccg2.data(), g, &rho, &drho);
}

double Species::rhopsg(double g) const {
  const double arg = 0.25 * rcps * rcps * g * g;
  return -(double)zval * std::exp(-arg);
}

double Species::eself() const {
  const double z = zval;
  return z * z / (std::sqrt(2.0 * M_PI) * rcps);
}

// src/Species.cpp:1417-1464: nonzero only on the (l,m) diagonal
double Species::dmatrix(int ipr, int jpr) const {
  if (quad()) {
    if (ipr != jpr) return 0.0;
    const int q = ipr % nquad, ilm = ipr / nquad;
    int j = 0;
    for (int l = 0; l <= lmax; l++) {
      if (l == llocal) continue;
      if (ilm < j + 2 * l + 1) return wql[l][q];
      j += 2 * l + 1;
    }
    return 0.0;
  }
  auto lm = [&](int ipr_, int& l, int& m) {
    int j = 0;
    for (int ll = 0; ll <= lmax; ll++)
      for (int mm = -ll; mm <= ll; mm++) {
        if (j == ipr_) { l = ll; m = mm; return; }
        j++;
      }
    l = -1; m = 0;
  };
  int li, mi, lj, mj;
  lm(ipr, li, mi);
  lm(jpr, lj, mj);
  if (li != lj || mi != mj) return 0.0;
  return dl[li];
}

// src/LDAFunctional.cpp:92-151
void lda_unpolarized(double rh, double& ee, double& vv) {
  const double c1 = 0.6203504908994001;
  const double c3 = -0.610887057711;
  const double A = 0.0311, B = -0.048, b1 = 1.0529, b2 = 0.3334, G = -0.1423;
  const double D = G / (1.0 + b1 + b2) - B;
  const double C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)));
  ee = 0.0;
  vv = 0.0;
  if (rh > 0.0) {
    const double ro13 = std::cbrt(rh);
    const double rs = c1 / ro13;
    double ex, vx, ec, vc;
    vx = c3 / rs;
    ex = 0.75 * vx;
    if (rs < 1.0) {
      const double logrs = std::log(rs);
      ec = A * logrs + B + C * rs * logrs + D * rs;
      vc = A * logrs + (B - A / 3.0) + (2.0 / 3.0) * C * rs * logrs +
           ((2.0 * D - C) / 3.0) * rs;
    } else {
      const double sqrtrs = std::sqrt(rs);
      const double den = 1.0 + b1 * sqrtrs + b2 * rs;
      ec = G / den;
      vc = ec * (1.0 + (7.0 / 6.0) * b1 * sqrtrs + (4.0 / 3.0) * b2 * rs) / den;
    }
    ee = ex + ec;
    vv = vx + vc;
  }
}

// Perdew-Zunger fit of the fully polarized electron gas (exchange constant
// 2^(1/3) times the paramagnetic one, the polarized correlation parameters;
// C and D from continuity at rs = 1), src/LDAFunctional.cpp:153-213
void lda_ferro(double rh, double& ee, double& vv) {
  const double c1 = 0.6203504908994001;
  const double cx = -0.769669463118;
  const double A = 0.01555, B = -0.0269, b1 = 1.3981, b2 = 0.2611, G = -0.0843;
  const double D = G / (1.0 + b1 + b2) - B;
  const double C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)));
  ee = 0.0;
  vv = 0.0;
  if (!(rh > 0.0)) return;
  const double rs = c1 / std::cbrt(rh);
  const double vx = cx / rs, ex = 0.75 * vx;
  double ec, vc;
  if (rs < 1.0) {
    const double l = std::log(rs);
    ec = A * l + B + C * rs * l + D * rs;
    vc = A * l + (B - A / 3.0) + (2.0 / 3.0) * C * rs * l + ((2.0 * D - C) / 3.0) * rs;
  } else {
    const double q = std::sqrt(rs);
    const double den = 1.0 + b1 * q + b2 * rs;
    ec = G / den;
    vc = ec * (1.0 + (7.0 / 6.0) * b1 * q + (4.0 / 3.0) * b2 * rs) / den;
  }
  ee = ex + ec;
  vv = vx + vc;
}

// spin-polarized LDA (src/LDAFunctional.cpp:42-89): energy per electron and
// the two spin potentials by von Barth-Hedin interpolation between the
// paramagnetic and ferromagnetic fits, f(z) = ((1+z)^(4/3) + (1-z)^(4/3) - 2)
// / (2^(4/3) - 2); zero where the total density is not positive
void lsda(double rup, double rdn, double& ee, double& vup, double& vdn) {
  ee = vup = vdn = 0.0;
  const double r = rup + rdn;
  if (!(r > 0.0)) return;
  const double kf = 1.0 / (2.0 * std::cbrt(2.0) - 2.0);
  const double z = (rup - rdn) / r;
  const double a = 1.0 + z, b = 1.0 - z;
  const double a3 = std::cbrt(a), b3 = std::cbrt(b);
  const double f = kf * (a3 * a + b3 * b - 2.0);
  const double df = (4.0 / 3.0) * kf * (a3 - b3);
  double eu, vu, ep, vp;
  lda_unpolarized(r, eu, vu);
  lda_ferro(r, ep, vp);
  const double de = ep - eu;
  ee = eu + f * de;
  const double v = vu + f * (vp - vu);
  vup = v + de * (1.0 - z) * df;
  vdn = v + de * (-1.0 - z) * df;
}

// Species::initialize_ncpp, Kleinman-Bylander branch (src/Species.cpp:71-330,
// nquad = 0), restated as the equivalent one-projector-per-l species: local
// potential vps[llocal]; nonlocal l: projector dv_l phi_l with D_l = w_l,
// w_l = 1 / (deltar * simpsn(dv_l (phi_l r)^2)), dv_l = vps_l - vps_llocal,
// integrated on the 40-bohr transform mesh (zero beyond the input mesh).
void ncpp_species(Species& q, int lmax, int llocal, int nmesh, const double* vps,
                  const double* phi) {
  if (lmax < 0 || lmax > 3 || llocal < 0 || llocal > lmax)
    throw std::runtime_error("ncpp: bad lmax/llocal");
  int ndft = 1;
  while (ndft * q.deltar < 40.0) ndft *= 2;
  int lkb = 0;
  for (int l = 0; l <= lmax; l++)
    if (l != llocal) {
      if (l > 2) throw std::runtime_error("ncpp: nonlocal channel with l > 2");
      lkb = std::max(lkb, l);
    }
  const double* vl = vps + (size_t)llocal * nmesh;
  q.vlocal.assign(vl, vl + nmesh);
  q.lmax = lkb;
  q.proj.assign(lkb + 1, std::vector<double>(nmesh, 0.0));
  q.dl.assign(lkb + 1, 0.0);
  for (int l = 0; l <= lkb; l++) {
    if (l == llocal) continue;
    const double* v = vps + (size_t)l * nmesh;
    const double* f = phi + (size_t)l * nmesh;
    std::vector<double> w(ndft, 0.0);
    for (int i = 0; i < nmesh && i < ndft; i++) {
      const double fr = f[i] * (q.deltar * i);
      w[i] = (v[i] - vl[i]) * fr * fr;
    }
    const double integral = simpsn(ndft, w.data());
    if (integral == 0.0) throw std::runtime_error("ncpp: zero projector norm");
    q.dl[l] = 1.0 / (q.deltar * integral);
    for (int i = 0; i < nmesh; i++) q.proj[l][i] = (v[i] - vl[i]) * f[i];
  }
}

}  // namespace qbo

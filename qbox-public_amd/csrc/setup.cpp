// setup.cpp -- host setup of a system (see setup.h).
#include "setup.h"
#include <algorithm>
#include <cmath>
#include <complex>

namespace qbg {
namespace host {

static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// b_i = 2 pi (A^-T)_i with the reference's operation order
// (src/UnitCell.cpp:40-66): amt0 = (a1/vol) x a2, etc.
void Cell::set(const double* a9) {
  for (int i = 0; i < 9; i++) a[i / 3][i % 3] = a9[i];
  double t[3];
  cross3(a[1], a[2], t);
  volume = std::fabs(dot3(a[0], t));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) b[i][j] = 0.0;
  if (volume > 0.0) {
    const double f = 1.0 / volume;
    for (int i = 0; i < 3; i++) {
      const double* p = a[(i + 1) % 3];
      const double* q = a[(i + 2) % 3];
      double fp[3] = {f * p[0], f * p[1], f * p[2]};
      double r[3];
      cross3(fp, q, r);
      for (int j = 0; j < 3; j++) b[i][j] = 2.0 * M_PI * r[j];
    }
  }
  // the 13 neighbour vectors of fold_in_ws (src/UnitCell.cpp:92-115)
  static const int s[13][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0},
                               {0, 1, 1}, {0, 1, -1}, {1, 0, 1}, {-1, 0, 1}, {1, 1, 1},
                               {1, -1, -1}, {1, 1, -1}, {1, -1, 1}};
  for (int i = 0; i < 13; i++) {
    for (int j = 0; j < 3; j++) {
      // sums left to right in the order a0, a1, a2 except an_[7], an_[8]
      // which start from a2
      double v;
      if (i == 7) v = a[2][j] + a[0][j];
      else if (i == 8) v = a[2][j] - a[0][j];
      else v = (s[i][0] * a[0][j] + s[i][1] * a[1][j]) + s[i][2] * a[2][j];
      an[i][j] = v;
    }
    an2h[i] = 0.5 * dot3(an[i], an[i]);
  }
}

void Cell::fold_in_ws(double v[3]) const {
  const double eps = 1.e-10;
  bool done = false;
  for (int iter = 0; !done && iter < 10; iter++) {
    done = true;
    for (int i = 0; i < 13; i++) {
      const double sp = dot3(v, an[i]);
      if (sp > an2h[i] + eps) {
        done = false;
        do { for (int j = 0; j < 3; j++) v[j] -= an[i][j]; } while (dot3(v, an[i]) > an2h[i] + eps);
      } else if (sp < -an2h[i] - eps) {
        done = false;
        do { for (int j = 0; j < 3; j++) v[j] += an[i][j]; } while (dot3(v, an[i]) < -an2h[i] - eps);
      }
    }
  }
}

bool factorizable(int n) {
  for (int f : {11, 7, 5, 3, 3})
    if (n % f == 0) n /= f;
  while (n % 2 == 0) n /= 2;
  return n == 1;
}

// Rods of the Gamma half sphere: (0,0) first, then (0,k>0), then (h>0,k);
// stable by size with rod (0,0) exchanged into slot 0 afterwards.
void Basis::build(const Cell& def, const Cell& cell, double ecut, const double* kp,
                  bool force_complex) {
  for (int i = 0; i < 3; i++) kpoint[i] = kp ? kp[i] : 0.0;
  real = !force_complex && kpoint[0] == 0.0 && kpoint[1] == 0.0 && kpoint[2] == 0.0;
  const double two_ecut = 2.0 * ecut;
  const double fac = std::sqrt(two_ecut) / (2.0 * M_PI);
  int nmax[3];
  for (int i = 0; i < 3; i++) nmax[i] = (int)(1.5 + fac * std::sqrt(dot3(def.a[i], def.a[i])));
  const double* b0 = def.b[0];
  const double* b1 = def.b[1];
  const double* b2 = def.b[2];
  struct Rod { int h, k, lmin, size; };
  std::vector<Rod> rods;
  const double b2inv2 = 1.0 / dot3(b2, b2);
  const int lend0 = (int)std::sqrt(two_ecut * b2inv2);
  int used[3][2] = {{0, 0}, {0, 0}, {0, lend0}};   // [axis][min,max]
  if (real) rods.push_back({0, 0, 0, lend0 + 1});
  else
    for (int i = 0; i < 3; i++) { used[i][0] = nmax[i]; used[i][1] = -nmax[i]; }
  // a rod (h, k) of the full sphere around -k (:446-481)
  auto add_rod_k = [&](int h, int k) {
    int lo = nmax[2], hi = -nmax[2];
    for (int l = -nmax[2]; l <= nmax[2]; l++) {
      double g2 = 0.0;
      for (int j = 0; j < 3; j++) {
        const double t = ((kpoint[0] + h) * b0[j] + (kpoint[1] + k) * b1[j]) + (kpoint[2] + l) * b2[j];
        g2 += t * t;
      }
      if (g2 < two_ecut) { lo = std::min(lo, l); hi = std::max(hi, l); }
    }
    if (lo > hi) return;
    rods.push_back({h, k, lo, hi - lo + 1});
    used[0][0] = std::min(used[0][0], h); used[0][1] = std::max(used[0][1], h);
    used[1][0] = std::min(used[1][0], k); used[1][1] = std::max(used[1][1], k);
    used[2][0] = std::min(used[2][0], lo); used[2][1] = std::max(used[2][1], hi);
  };
  auto add_rod = [&](int h, int k) {
    int lo = nmax[2], hi = -nmax[2];
    for (int l = -nmax[2]; l <= nmax[2]; l++) {
      double g2 = 0.0;
      for (int j = 0; j < 3; j++) {
        const double t = (h == 0 ? k * b1[j] : h * b0[j] + k * b1[j]) + l * b2[j];
        g2 += t * t;
      }
      if (g2 < two_ecut) { lo = std::min(lo, l); hi = std::max(hi, l); }
    }
    if (lo > hi) return;
    rods.push_back({h, k, lo, hi - lo + 1});
    used[0][0] = std::min(used[0][0], h); used[0][1] = std::max(used[0][1], h);
    used[1][0] = std::min(used[1][0], k); used[1][1] = std::max(used[1][1], k);
    used[2][0] = std::min(used[2][0], lo); used[2][1] = std::max(used[2][1], hi);
  };
  if (real) {
    for (int k = 1; k <= nmax[1]; k++) add_rod(0, k);
    for (int h = 1; h <= nmax[0]; h++)
      for (int k = -nmax[1]; k <= nmax[1]; k++) add_rod(h, k);
  } else {
    for (int h = -nmax[0]; h <= nmax[0]; h++)
      for (int k = -nmax[1]; k <= nmax[1]; k++) add_rod_k(h, k);
  }
  for (int i = 0; i < 3; i++) { idxmin[i] = used[i][0]; idxmax[i] = used[i][1]; }
  for (int i = 0; i < 3; i++) {
    int n = 2 * nmax[i];
    while (!factorizable(n)) n += 2;
    np[i] = n;
  }
  std::stable_sort(rods.begin(), rods.end(), [](const Rod& x, const Rod& y) { return x.size < y.size; });
  bool rod0 = false;
  for (size_t i = 0; i < rods.size(); i++)
    if (rods[i].h == 0 && rods[i].k == 0) { std::swap(rods[i], rods[0]); rod0 = true; break; }
  if (!rod0) throw std::invalid_argument("basis without the rod (0,0): k outside the zone?");
  nrods = (int)rods.size();
  size = 0;
  rod_h.clear(); rod_k.clear(); rod_lmin.clear(); rod_size.clear(); idx.clear();
  for (const Rod& r : rods) {
    rod_h.push_back(r.h); rod_k.push_back(r.k); rod_lmin.push_back(r.lmin); rod_size.push_back(r.size);
    for (int l = 0; l < r.size; l++) {
      idx.push_back(r.h); idx.push_back(r.k); idx.push_back(r.lmin + l);
    }
    size += r.size;
  }
  // update_g with the current cell (src/Basis.cpp:646-690)
  g.resize(size); g2.resize(size); gi.resize(size); g2i.resize(size); gx.resize(3 * size);
  kpg.resize(size); kpg2.resize(size); kpgi.resize(size); kpgx.resize(3 * size);
  for (int i = 0; i < size; i++) {
    double t[3], u[3];
    for (int j = 0; j < 3; j++) {
      t[j] = (idx[3 * i] * cell.b[0][j] + idx[3 * i + 1] * cell.b[1][j]) + idx[3 * i + 2] * cell.b[2][j];
      u[j] = ((kpoint[0] + idx[3 * i]) * cell.b[0][j] + (kpoint[1] + idx[3 * i + 1]) * cell.b[1][j]) +
             (kpoint[2] + idx[3 * i + 2]) * cell.b[2][j];
    }
    for (int j = 0; j < 3; j++) { gx[j * size + i] = t[j]; kpgx[j * size + i] = u[j]; }
    g2[i] = dot3(t, t);
    g[i] = std::sqrt(g2[i]);
    gi[i] = g[i] > 0.0 ? 1.0 / g[i] : 0.0;
    g2i[i] = gi[i] * gi[i];
    kpg2[i] = dot3(u, u);
    kpg[i] = std::sqrt(kpg2[i]);
    kpgi[i] = kpg[i] > 0.0 ? 1.0 / kpg[i] : 0.0;
  }
}

bool Maps::build(const Basis& b, int np0, int np1, int np2) {
  if (!b.real) return false;   // the band grids run on real bases only
  if (!(b.idxmax[0] < np0 / 2 && b.idxmin[0] >= -np0 / 2 && b.idxmax[1] < np1 / 2 &&
        b.idxmin[1] >= -np1 / 2 && b.idxmax[2] < np2 / 2 && b.idxmin[2] >= -np2 / 2))
    return false;
  nvec = 2 * b.nrods - 1;
  ip.resize(b.size);
  im.resize(b.size);
  zvec_to_val.assign(1, 0);
  int ig = 0;
  for (int r = 0; r < b.nrods; r++) {
    for (int i = 0; i < b.rod_size[r]; i++, ig++) {
      const int l = b.rod_lmin[r] + i;
      const int zp = ((l % np2) + np2) % np2, zm = ((-l % np2) + np2) % np2;
      if (r == 0) { ip[ig] = zp; im[ig] = zm; }
      else { ip[ig] = (2 * r - 1) * np2 + zp; im[ig] = 2 * r * np2 + zm; }
    }
    if (r > 0) {
      const int hp = ((b.rod_h[r] % np0) + np0) % np0, kp = ((b.rod_k[r] % np1) + np1) % np1;
      const int hm = (np0 - hp) % np0, km = (np1 - kp) % np1;
      zvec_to_val.push_back(hp + np0 * kp);
      zvec_to_val.push_back(hm + np0 * km);
    }
  }
  return true;
}

// ---------------------------------------------------------------- species
// radix-2 complex FFT, e^{-2 pi i jk/n} (analysis), n a power of two
static void fft_pow2(std::vector<std::complex<double>>& z) {
  const int n = (int)z.size();
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(z[i], z[j]);
  }
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int len = 2; len <= n; len <<= 1) {
    std::vector<std::complex<double>> w(len / 2);
    for (int k = 0; k < len / 2; k++) {
      const long double a = -2.0L * pi * k / len;
      w[k] = {(double)cosl(a), (double)sinl(a)};
    }
    for (int i = 0; i < n; i += len)
      for (int k = 0; k < len / 2; k++) {
        const std::complex<double> u = z[i + k], v = z[i + k + len / 2] * w[k];
        z[i + k] = u + v;
        z[i + k + len / 2] = u - v;
      }
  }
}

// src/sinft.cpp:50-105: odd extension, 2n-point forward DFT, -Im/2
static void sinft(int n, double* y) {
  std::vector<std::complex<double>> z(2 * n, 0.0);
  for (int i = 1; i < n; i++) { z[i] = y[i]; z[2 * n - i] = -y[i]; }
  fft_pow2(z);
  for (int i = 0; i < n; i++) y[i] = -0.5 * z[i].imag();
}

// src/sinft.cpp:108-174: even extension of n+1 points, Re/2
static void cosft1(int n, double* y) {
  std::vector<std::complex<double>> z(2 * n, 0.0);
  z[0] = y[0];
  for (int i = 1; i <= n; i++) { z[i] = y[i]; z[2 * n - i] = y[i]; }
  fft_pow2(z);
  for (int i = 0; i < n; i++) y[i] = 0.5 * z[i].real();
}

static double simpsn(int n, const double* t) {   // src/Species.cpp:34-48
  double s = (17.0 / 48.0) * (t[0] + t[n - 1]) + (59.0 / 48.0) * (t[1] + t[n - 2]) +
             (43.0 / 48.0) * (t[2] + t[n - 3]) + (49.0 / 48.0) * (t[3] + t[n - 4]);
  for (int i = 4; i < n - 4; i++) s += t[i];
  return s;
}

// cubic spline second derivatives (src/spline.cpp:21-108)
static void spline(int n, const double* x, const double* y, bool nat_left, double* y2) {
  std::vector<double> d(n), e(n), f(n);
  if (!nat_left) {
    const double h = x[1] - x[0];
    d[0] = h / 3.0; e[0] = h / 6.0; f[0] = 0.0; y2[0] = (y[1] - y[0]) / h;
  } else {
    d[0] = 1.0; e[0] = 0.0; f[0] = 0.0; y2[0] = 0.0;
  }
  d[n - 1] = 1.0; e[n - 1] = 0.0; f[n - 1] = 0.0; y2[n - 1] = 0.0;   // natural right end
  for (int i = 1; i < n - 1; i++) {
    const double hp = x[i + 1] - x[i], hm = x[i] - x[i - 1];
    d[i] = (1.0 / 3.0) * (hp + hm);
    e[i] = (1.0 / 6.0) * hp;
    f[i] = (1.0 / 6.0) * hm;
    y2[i] = (y[i + 1] - y[i]) / hp - (y[i] - y[i - 1]) / hm;
  }
  for (int i = 1; i < n; i++) { f[i] /= d[i - 1]; d[i] -= f[i] * e[i - 1]; }
  for (int i = 1; i < n; i++) y2[i] -= f[i] * y2[i - 1];
  y2[n - 1] /= d[n - 1];
  for (int i = n - 2; i >= 0; i--) y2[i] = (y2[i] - e[i] * y2[i + 1]) / d[i];
}

// cubic spline with given first derivatives at both ends (src/spline.cpp:44-108
// with bcnat_left = bcnat_right = 0)
static void spline_clamped(int n, const double* x, const double* y, double yp1, double ypn,
                           double* y2) {
  std::vector<double> d(n), e(n), f(n);
  const double h0 = x[1] - x[0], hn = x[n - 1] - x[n - 2];
  d[0] = h0 / 3.0; e[0] = h0 / 6.0; f[0] = 0.0; y2[0] = (y[1] - y[0]) / h0 - yp1;
  d[n - 1] = hn / 3.0; e[n - 1] = 0.0; f[n - 1] = hn / 6.0;
  y2[n - 1] = ypn - (y[n - 1] - y[n - 2]) / hn;
  for (int i = 1; i < n - 1; i++) {
    const double hp = x[i + 1] - x[i], hm = x[i] - x[i - 1];
    d[i] = (1.0 / 3.0) * (hp + hm);
    e[i] = (1.0 / 6.0) * hp;
    f[i] = (1.0 / 6.0) * hm;
    y2[i] = (y[i + 1] - y[i]) / hp - (y[i] - y[i - 1]) / hm;
  }
  for (int i = 1; i < n; i++) { f[i] /= d[i - 1]; d[i] -= f[i] * e[i - 1]; }
  for (int i = 1; i < n; i++) y2[i] -= f[i] * y2[i - 1];
  y2[n - 1] /= d[n - 1];
  for (int i = n - 2; i >= 0; i--) y2[i] = (y2[i] - e[i] * y2[i + 1]) / d[i];
}

static void splintd(int n, const double* xa, const double* ya, const double* y2a, double x,
                    double& y, double& dy) {
  int kl = 0, kh = n - 1;
  while (kh - kl > 1) {
    const int k = (kh + kl) / 2;
    if (xa[k] > x) kh = k; else kl = k;
  }
  const double h = xa[kh] - xa[kl];
  const double a = (xa[kh] - x) / h, b = (x - xa[kl]) / h;
  y = a * ya[kl] + b * ya[kh] + h * h * (1.0 / 6.0) * ((a * a * a - a) * y2a[kl] + (b * b * b - b) * y2a[kh]);
  dy = (ya[kh] - ya[kl]) / h + h * (((1.0 / 6.0) - 0.5 * a * a) * y2a[kl] + (0.5 * b * b - (1.0 / 6.0)) * y2a[kh]);
}

bool ncpp_to_kb(int lmax, int llocal, int nmesh, double deltar, const double* vps,
                const double* phi, std::vector<double>& vloc, std::vector<double>& proj,
                std::vector<double>& dl, int& lmax_kb) {
  if (lmax < 0 || lmax > 3 || llocal < 0 || llocal > lmax || nmesh < 8 || !(deltar > 0.0))
    return false;
  int ndft = 1;
  while (ndft * deltar < 40.0) ndft *= 2;
  const double* vl = vps + (size_t)llocal * nmesh;
  lmax_kb = 0;
  for (int l = 0; l <= lmax; l++)
    if (l != llocal) {
      if (l > 2) return false;
      lmax_kb = std::max(lmax_kb, l);
    }
  vloc.assign(vl, vl + nmesh);
  proj.assign((size_t)(lmax_kb + 1) * nmesh, 0.0);
  dl.assign(lmax_kb + 1, 0.0);
  std::vector<double> fint(ndft, 0.0);
  for (int l = 0; l <= lmax_kb; l++) {
    if (l == llocal) continue;
    const double* v = vps + (size_t)l * nmesh;
    const double* f = phi + (size_t)l * nmesh;
    // weight 1 / (deltar * simpson((v_l - v_loc) (phi_l r)^2)); both vanish
    // beyond the mesh (src/Species.cpp:272-289)
    std::fill(fint.begin(), fint.end(), 0.0);
    for (int i = 0; i < std::min(nmesh, ndft); i++) {
      const double t = f[i] * (i * deltar);
      fint[i] = (v[i] - vl[i]) * t * t;
    }
    const double tmp = simpsn(ndft, fint.data());
    if (!(tmp != 0.0)) return false;
    dl[l] = 1.0 / (deltar * tmp);
    for (int i = 0; i < nmesh; i++) proj[(size_t)l * nmesh + i] = (v[i] - vl[i]) * f[i];
  }
  return true;
}

bool Species::init() {
  if (lmax < 0 || lmax > 3 || nmesh < 8 || !(deltar > 0.0)) return false;
  const double fpi = 4.0 * M_PI;
  nlm = (lmax + 1) * (lmax + 1);
  if (quad()) {
    if (llocal < 0 || llocal > lmax || !(rquad > 0.0) ||
        (int)vsl.size() != (lmax + 1) * nmesh)
      return false;
    nlm = 0;
    for (int l = 0; l <= lmax; l++)
      if (l != llocal) {
        if (l > 2) return false;   // the reference has no l = 3 quadrature projectors
        nlm += (2 * l + 1) * nquad;
      }
    vloc.assign(vsl.begin() + (size_t)llocal * nmesh, vsl.begin() + (size_t)(llocal + 1) * nmesh);
  }
  ndft = 1;
  while (ndft * deltar < 40.0) ndft *= 2;
  std::vector<double> r(ndft), vr(ndft), fint(ndft);
  for (int i = 0; i < ndft; i++) r[i] = i * deltar;
  for (int i = 0; i < ndft; i++) vr[i] = i < nmesh ? vloc[i] : -zval / r[i];
  // local part minus the Gaussian pseudocharge potential (src/Species.cpp:648-655)
  const double z = zval;
  vr[0] += (z / rcps) * M_2_SQRTPI;
  for (int i = 1; i < ndft; i++) vr[i] += (z / r[i]) * std::erf(r[i] / rcps);
  for (int i = 0; i < ndft; i++) vr[i] *= fpi * r[i] * deltar;
  for (int i = 0; i < ndft; i++) fint[i] = vr[i] * r[i];
  const double v0 = simpsn(ndft, fint.data());
  vlocg = vr;
  sinft(ndft, vlocg.data());
  gspl.resize(ndft);
  const double dg = M_PI / (ndft * deltar);
  gspl[0] = 0.0;
  vlocg[0] = v0;
  for (int i = 1; i < ndft; i++) { gspl[i] = i * dg; vlocg[i] /= gspl[i]; }
  vlocg2.resize(ndft);
  spline(ndft, gspl.data(), vlocg.data(), false, vlocg2.data());
  // core density (Species::initialize_nlcc, src/Species.cpp:1056-1100, and
  // :698-717): (4 pi / G) sum_r sin(Gr) r dr nlcc(r); G = 0: the Simpson
  // integral of 4 pi r^2 nlcc(r)
  nlccg.clear();
  nlccg2.clear();
  zcore = 0.0;
  if (has_nlcc()) {
    nlccg.assign(ndft, 0.0);
    for (int i = 0; i < ndft && i < (int)nlcc.size(); i++) nlccg[i] = fpi * r[i] * deltar * nlcc[i];
    for (int i = 0; i < ndft; i++) fint[i] = nlccg[i] * r[i];
    zcore = simpsn(ndft, fint.data());
    sinft(ndft, nlccg.data());
    nlccg[0] = zcore;
    for (int i = 1; i < ndft; i++) nlccg[i] /= gspl[i];
    nlccg2.resize(ndft);
    spline(ndft, gspl.data(), nlccg.data(), false, nlccg2.data());
  }
  vnlg.assign(lmax + 1, std::vector<double>(ndft + 1, 0.0));
  vnlg2.assign(lmax + 1, std::vector<double>(ndft + 1, 0.0));
  if (quad()) {
    // semilocal potentials extended by -Z/r, splined with v'(0) = 0 and
    // v'(r_max) = Z/r_max^2 (src/Species.cpp:162-184, Species::dvpsr
    // :1120-1131); trapezoid rule with interior points
    // (src/NonLocalPotential.cpp:119-131)
    const double h = rquad / (nquad + 1);
    rq.resize(nquad);
    for (int q = 0; q < nquad; q++) rq[q] = (q + 1) * h;
    std::vector<std::vector<double>> vs(lmax + 1, std::vector<double>(ndft)),
        vs2(lmax + 1, std::vector<double>(ndft));
    const double dvdr = zval / (r[ndft - 1] * r[ndft - 1]);
    for (int l = 0; l <= lmax; l++) {
      for (int i = 0; i < ndft; i++) vs[l][i] = i < nmesh ? vsl[(size_t)l * nmesh + i] : -zval / r[i];
      spline_clamped(ndft, r.data(), vs[l].data(), 0.0, dvdr, vs2[l].data());
    }
    auto vpsr = [&](int l, double x) {
      if (x > r[ndft - 1]) return 0.0;
      double v, dv;
      splintd(ndft, r.data(), vs[l].data(), vs2[l].data(), x, v, dv);
      return v;
    };
    wql.assign(lmax + 1, std::vector<double>(nquad, 0.0));
    for (int l = 0; l <= lmax; l++)
      if (l != llocal)
        for (int q = 0; q < nquad; q++) wql[l][q] = (vpsr(l, rq[q]) - vpsr(llocal, rq[q])) * h;
    return true;
  }
  for (int l = 0; l <= lmax; l++) {
    std::vector<double> w(ndft);
    for (int i = 0; i < ndft; i++) w[i] = i < nmesh ? fpi * deltar * proj[(size_t)l * nmesh + i] : 0.0;
    std::vector<double>& v = vnlg[l];
    if (l == 0) {   // (1/G) sum_r sin(Gr) r w ; G=0 by Simpson
      for (int i = 0; i < ndft; i++) fint[i] = w[i] * r[i] * r[i];
      const double w0 = simpsn(ndft, fint.data());
      for (int i = 0; i < ndft; i++) v[i] = w[i] * r[i];
      sinft(ndft, v.data());
      v[0] = w0;
      for (int i = 1; i < ndft; i++) v[i] /= gspl[i];
      spline(ndft, gspl.data(), v.data(), false, vnlg2[l].data());
    } else if (l == 1) {   // (1/G^2) sum sin(Gr) w - (1/G) sum cos(Gr) r w
      for (int i = 0; i < ndft; i++) v[i] = w[i];
      sinft(ndft, v.data());
      fint[0] = 0.0;
      for (int i = 1; i < ndft; i++) fint[i] = v[i] / (gspl[i] * gspl[i]);
      for (int i = 0; i < ndft; i++) v[i] = w[i] * r[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = fint[i] - v[i] / gspl[i];
      spline(ndft, gspl.data(), v.data(), true, vnlg2[l].data());
    } else if (l == 3) {
      // j_3 (src/Species.cpp:948-1046): 15/G^4 sum sin(Gr) w/r^2 - 6/G^2 sum sin(Gr) w
      // + 1/G sum cos(Gr) r w - 15/G^3 sum cos(Gr) w/r; zero at G = 0
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = w[i] / (r[i] * r[i]);
      sinft(ndft, v.data());
      fint[0] = 0.0;
      for (int i = 1; i < ndft; i++) {
        const double g = gspl[i];
        fint[i] = 15.0 * v[i] / (g * g * g * g);
      }
      for (int i = 0; i < ndft; i++) v[i] = w[i];
      sinft(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] -= 6.0 * v[i] / (gspl[i] * gspl[i]);
      for (int i = 0; i < ndft; i++) v[i] = w[i] * r[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] += v[i] / gspl[i];
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = w[i] / r[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      for (int i = 1; i < ndft; i++) {
        const double g = gspl[i];
        fint[i] -= 15.0 * v[i] / (g * g * g);
      }
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = fint[i];
      spline(ndft, gspl.data(), v.data(), false, vnlg2[l].data());
    } else {   // j_2: (3/G^3) sum sin(Gr) w/r - (1/G) sum sin(Gr) r w - (3/G^2) sum cos(Gr) w
               // (src/Species.cpp:872-940); zero at G = 0
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = w[i] / r[i];
      sinft(ndft, v.data());
      fint[0] = 0.0;
      for (int i = 1; i < ndft; i++) fint[i] = 3.0 * v[i] / (gspl[i] * gspl[i] * gspl[i]);
      for (int i = 0; i < ndft; i++) v[i] = w[i] * r[i];
      sinft(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] -= v[i] / gspl[i];
      for (int i = 0; i < ndft; i++) v[i] = w[i];
      v[ndft] = 0.0;
      cosft1(ndft, v.data());
      for (int i = 1; i < ndft; i++) fint[i] -= 3.0 * v[i] / (gspl[i] * gspl[i]);
      v[0] = 0.0;
      for (int i = 1; i < ndft; i++) v[i] = fint[i];
      // zero slope at G = 0, natural at the end
      spline(ndft, gspl.data(), v.data(), false, vnlg2[l].data());
    }
  }
  return true;
}

void Species::dvlocg(double g, double& v, double& dv) const {
  if (g > gspl[ndft - 1]) { v = dv = 0.0; return; }
  splintd(ndft, gspl.data(), vlocg.data(), vlocg2.data(), g, v, dv);
}

void Species::dvnlg(int l, double g, double& v, double& dv) const {
  if (g > gspl[ndft - 1]) { v = dv = 0.0; return; }
  splintd(ndft, gspl.data(), vnlg[l].data(), vnlg2[l].data(), g, v, dv);
}

void Species::quad_radial(int l, int q, double g, double& v, double& dv) const {
  const double fpi = 4.0 * M_PI, r = rq[q], z = g * r;
  v = dv = 0.0;
  if (l == 0) {   // 4 pi sin(Gr)/G, the G = 0 limit 4 pi r
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

void Species::projectors(std::vector<int>& pl, std::vector<int>& ph, std::vector<int>& pq,
                         std::vector<double>& pd) const {
  pl.assign(nlm, 0);
  ph.assign(nlm, 0);
  pq.assign(nlm, -1);
  pd.assign(nlm, 0.0);
  if (!quad()) {
    for (int ipr = 0; ipr < nlm; ipr++) {
      const int l = ipr == 0 ? 0 : (ipr < 4 ? 1 : (ipr < 9 ? 2 : 3));
      pl[ipr] = l;
      ph[ipr] = ipr;
      pd[ipr] = dl[l];   // D diagonal in (l, m) (src/Species.cpp:1450-1464)
    }
    return;
  }
  int ilm = 0;
  for (int l = 0; l <= lmax; l++) {
    if (l == llocal) continue;
    for (int m = 0; m < 2 * l + 1; m++)
      for (int q = 0; q < nquad; q++) {
        const int ipr = q + nquad * (ilm + m);
        pl[ipr] = l;
        ph[ipr] = l * l + m;
        pq[ipr] = q;
        pd[ipr] = wql[l][q];
      }
    ilm += 2 * l + 1;
  }
}

void Species::drhocoreg(double g, double& rho, double& drho) const {
  if (!has_nlcc() || g > gspl[ndft - 1]) { rho = drho = 0.0; return; }
  splintd(ndft, gspl.data(), nlccg.data(), nlccg2.data(), g, rho, drho);
}

double Species::rhopsg(double g) const {
  const double arg = 0.25 * rcps * rcps * g * g;
  return -(double)zval * std::exp(-arg);
}

double Species::eself() const { return (double)zval * zval / (std::sqrt(2.0 * M_PI) * rcps); }

}  // namespace host
}  // namespace qbg

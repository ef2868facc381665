// basis.cpp -- oracle restatement of UnitCell, Basis and BasisMapping.
// TEST INFRASTRUCTURE ONLY (see qbo.h).
#include "qbo.h"
#include <cmath>
#include <cassert>
#include <algorithm>
#include <stdexcept>

namespace qbo {

static inline void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// src/UnitCell.cpp:25-124
void Cell::set(const double* a9) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = a9[3 * i + j];
  double c12[3], c20[3], c01[3];
  cross(a[1], a[2], c12);
  volume = std::fabs(dot3(a[0], c12));
  if (volume > 0.0) {
    const double fac = 1.0 / volume;
    // amt0 = (fac*a1)^a2 etc. (operator precedence of the reference)
    double fa1[3] = {fac * a[1][0], fac * a[1][1], fac * a[1][2]};
    double fa2[3] = {fac * a[2][0], fac * a[2][1], fac * a[2][2]};
    double fa0[3] = {fac * a[0][0], fac * a[0][1], fac * a[0][2]};
    double amt0[3], amt1[3], amt2[3];
    cross(fa1, a[2], amt0);
    cross(fa2, a[0], amt1);
    cross(fa0, a[1], amt2);
    for (int j = 0; j < 3; j++) {
      b[0][j] = 2.0 * M_PI * amt0[j];
      b[1][j] = 2.0 * M_PI * amt1[j];
      b[2][j] = 2.0 * M_PI * amt2[j];
    }
    (void)c20; (void)c01;
  } else {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) b[i][j] = 0.0;
  }
  const int comb[13][3] = {{1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},
                           {1, -1, 0}, {0, 1, 1},  {0, 1, -1}, {1, 0, 1},
                           {-1, 0, 1}, {1, 1, 1},  {1, -1, -1}, {1, 1, -1},
                           {1, -1, 1}};
  // an_[i] built as sums/differences of a_ in the reference order
  for (int i = 0; i < 13; i++) {
    for (int j = 0; j < 3; j++) an[i][j] = 0.0;
  }
  for (int j = 0; j < 3; j++) {
    an[0][j] = a[0][j];
    an[1][j] = a[1][j];
    an[2][j] = a[2][j];
    an[3][j] = a[0][j] + a[1][j];
    an[4][j] = a[0][j] - a[1][j];
    an[5][j] = a[1][j] + a[2][j];
    an[6][j] = a[1][j] - a[2][j];
    an[7][j] = a[2][j] + a[0][j];
    an[8][j] = a[2][j] - a[0][j];
    an[9][j] = (a[0][j] + a[1][j]) + a[2][j];
    an[10][j] = (a[0][j] - a[1][j]) - a[2][j];
    an[11][j] = (a[0][j] + a[1][j]) - a[2][j];
    an[12][j] = (a[0][j] - a[1][j]) + a[2][j];
  }
  (void)comb;
  for (int i = 0; i < 13; i++) an2h[i] = 0.5 * dot3(an[i], an[i]);
}

// src/UnitCell.cpp:155-186
void Cell::fold_in_ws(double v[3]) const {
  const double epsilon = 1.e-10;
  bool done = false;
  const int maxiter = 10;
  int iter = 0;
  while (!done && iter < maxiter) {
    done = true;
    for (int i = 0; i < 13; i++) {
      const double sp = dot3(v, an[i]);
      if (sp > an2h[i] + epsilon) {
        done = false;
        do {
          for (int j = 0; j < 3; j++) v[j] -= an[i][j];
        } while (dot3(v, an[i]) > an2h[i] + epsilon);
      } else if (sp < -an2h[i] - epsilon) {
        done = false;
        do {
          for (int j = 0; j < 3; j++) v[j] += an[i][j];
        } while (dot3(v, an[i]) < -an2h[i] - epsilon);
      }
    }
    iter++;
  }
  if (iter >= maxiter) throw std::runtime_error("fold_in_ws did not converge");
}

// src/Basis.cpp:105-133
bool factorizable(int n) {
  if (n % 11 == 0) n /= 11;
  if (n % 7 == 0) n /= 7;
  if (n % 5 == 0) n /= 5;
  if (n % 3 == 0) n /= 3;
  if (n % 3 == 0) n /= 3;
  while (n % 2 == 0) n /= 2;
  return n == 1;
}

// norm2 of (h*b0 + k*b1) + l*b2, with the reference's evaluation order:
// each term is a D3vector scaled componentwise, sums left to right.
static inline double norm2_hkl(const double* b0, const double* b1,
                               const double* b2, double h, double k, double l,
                               bool use_b0) {
  double v[3];
  for (int j = 0; j < 3; j++) {
    double t = k * b1[j];
    if (use_b0) t = h * b0[j] + t;
    v[j] = t + l * b2[j];
  }
  return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}

// norm2 of (kx+h)*b0 + (ky+k)*b1 + (kz+l)*b2 in the reference's order
// (src/Basis.cpp:460)
static inline double norm2_kpg(const double* b0, const double* b1, const double* b2,
                               const double* kp, int h, int k, int l) {
  double v[3];
  for (int j = 0; j < 3; j++)
    v[j] = ((kp[0] + h) * b0[j] + (kp[1] + k) * b1[j]) + (kp[2] + l) * b2[j];
  return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}

// src/Basis.cpp:255-643 (one task)
void Basis::resize(const Cell& c, const Cell* refcell, double ecut_ha, const double* kp,
                   bool force_complex) {
  cell = c;
  ecut = ecut_ha;
  for (int i = 0; i < 3; i++) kpoint[i] = kp ? kp[i] : 0.0;
  real = !force_complex && kpoint[0] == 0.0 && kpoint[1] == 0.0 && kpoint[2] == 0.0;
  const Cell& defcell = refcell ? *refcell : c;
  const double two_ecut = 2.0 * ecut;
  const double twopi = 2.0 * M_PI;
  const double* b0 = defcell.b[0];
  const double* b1 = defcell.b[1];
  const double* b2 = defcell.b[2];
  const double normb2 = dot3(b2, b2);
  const double b2inv2 = 1.0 / normb2;
  const double fac = std::sqrt(two_ecut) / twopi;
  auto len = [](const double* x) { return std::sqrt(dot3(x, x)); };
  const int hmax = (int)(1.5 + fac * len(defcell.a[0]));
  const int kmax = (int)(1.5 + fac * len(defcell.a[1]));
  const int lmax = (int)(1.5 + fac * len(defcell.a[2]));
  const int kmin = -kmax, lmin = -lmax;

  // rods in insertion order, then a stable sort by size reproduces
  // multiset<Rod> ordering (size-only comparison, ties after equals)
  struct R { int h, k, lmin, size; };
  std::vector<R> rods;
  int hmax_used = 0, hmin_used = 0, kmax_used = 0, kmin_used = 0;
  int lmax_used = 0, lmin_used = 0;
  size = 0;
  if (!real) {
    // k != 0: every rod (h, k) of the sphere |k + G| < sqrt(2 ecut) (:446-481)
    const int hmin = -hmax;
    hmax_used = hmin; hmin_used = hmax; kmax_used = kmin; kmin_used = kmax;
    lmax_used = lmin; lmin_used = lmax;
    for (int h = hmin; h <= hmax; h++)
      for (int k = kmin; k <= kmax; k++) {
        int lstart = lmax, lend = lmin;
        bool found = false;
        for (int l = lmin; l <= lmax; l++)
          if (norm2_kpg(b0, b1, b2, kpoint, h, k, l) < two_ecut) {
            lstart = std::min(l, lstart);
            lend = std::max(l, lend);
            found = true;
          }
        if (!found) continue;
        const int rs = lend - lstart + 1;
        size += rs;
        rods.push_back({h, k, lstart, rs});
        hmax_used = std::max(h, hmax_used); hmin_used = std::min(h, hmin_used);
        kmax_used = std::max(k, kmax_used); kmin_used = std::min(k, kmin_used);
        lmax_used = std::max(lend, lmax_used); lmin_used = std::min(lstart, lmin_used);
      }
  }
  int lend0 = (int)(std::sqrt(two_ecut * b2inv2));
  if (real) {
    rods.push_back({0, 0, 0, lend0 + 1});
    lmax_used = lend0;
    size = lend0 + 1;
  }
  auto scan = [&](int h, int k, bool use_b0) {
    int lstart = lmax, lend = lmin;
    bool found = false;
    for (int l = lmin; l <= lmax; l++) {
      const double two_e = norm2_hkl(b0, b1, b2, h, k, l, use_b0);
      if (two_e < two_ecut) {
        lstart = std::min(l, lstart);
        lend = std::max(l, lend);
        found = true;
      }
    }
    if (found) {
      const int rs = lend - lstart + 1;
      size += rs;
      rods.push_back({h, k, lstart, rs});
      if (use_b0) {
        hmax_used = std::max(h, hmax_used);
        hmin_used = std::min(h, hmin_used);
      }
      kmax_used = std::max(k, kmax_used);
      kmin_used = std::min(k, kmin_used);
      lmax_used = std::max(lend, lmax_used);
      lmin_used = std::min(lstart, lmin_used);
    }
  };
  if (real) {
    for (int k = 1; k <= kmax; k++) scan(0, k, false);     // :366-390
    for (int h = 1; h <= hmax; h++)                         // :391-420
      for (int k = kmin; k <= kmax; k++) scan(h, k, true);
  }
  nrods = (int)rods.size();
  idxmax[0] = hmax_used; idxmin[0] = hmin_used;
  idxmax[1] = kmax_used; idxmin[1] = kmin_used;
  idxmax[2] = lmax_used; idxmin[2] = lmin_used;

  int n;
  n = 2 * hmax; while (!factorizable(n)) n += 2; np[0] = n;
  n = 2 * kmax; while (!factorizable(n)) n += 2; np[1] = n;
  n = 2 * lmax; while (!factorizable(n)) n += 2; np[2] = n;

  std::stable_sort(rods.begin(), rods.end(),
                   [](const R& x, const R& y) { return x.size < y.size; });
  // :593-597 rod(0,0) swapped into position 0
  int r0 = -1;
  for (int i = 0; i < nrods; i++)
    if (rods[i].h == 0 && rods[i].k == 0) { r0 = i; break; }
  if (r0 < 0) throw std::runtime_error("basis without the rod (0,0)");
  std::swap(rods[r0], rods[0]);

  rod_h.resize(nrods); rod_k.resize(nrods); rod_lmin.resize(nrods);
  rod_size.resize(nrods); rod_first.resize(nrods);
  for (int i = 0; i < nrods; i++) {
    rod_h[i] = rods[i].h; rod_k[i] = rods[i].k;
    rod_lmin[i] = rods[i].lmin; rod_size[i] = rods[i].size;
    rod_first[i] = i == 0 ? 0 : rod_first[i - 1] + rod_size[i - 1];
  }
  idx.resize(3 * size);
  int i = 0;
  for (int irod = 0; irod < nrods; irod++)
    for (int l = 0; l < rod_size[irod]; l++) {
      idx[3 * i] = rod_h[irod];
      idx[3 * i + 1] = rod_k[irod];
      idx[3 * i + 2] = rod_lmin[irod] + l;
      i++;
    }
  update_g();
}

// src/Basis.cpp:646-690
void Basis::update_g() {
  g.resize(size); g2.resize(size); gi.resize(size); g2i.resize(size);
  gx.resize(3 * size);
  kpg.resize(size); kpg2.resize(size); kpgi.resize(size); kpgx.resize(3 * size);
  for (int i = 0; i < size; i++) {
    double gt[3], kt[3];
    for (int j = 0; j < 3; j++) {
      gt[j] = (idx[3 * i] * cell.b[0][j] + idx[3 * i + 1] * cell.b[1][j]) +
              idx[3 * i + 2] * cell.b[2][j];
      kt[j] = ((kpoint[0] + idx[3 * i]) * cell.b[0][j] +
               (kpoint[1] + idx[3 * i + 1]) * cell.b[1][j]) +
              (kpoint[2] + idx[3 * i + 2]) * cell.b[2][j];
    }
    for (int j = 0; j < 3; j++) {
      gx[j * size + i] = gt[j];
      kpgx[j * size + i] = kt[j];
    }
    g2[i] = dot3(gt, gt);
    g[i] = std::sqrt(g2[i]);
    gi[i] = g[i] > 0.0 ? 1.0 / g[i] : 0.0;
    g2i[i] = gi[i] * gi[i];
    kpg2[i] = dot3(kt, kt);
    kpg[i] = std::sqrt(kpg2[i]);
    kpgi[i] = kpg[i] > 0.0 ? 1.0 / kpg[i] : 0.0;
  }
}

// src/BasisMapping.cpp:38-228, single task, real basis
void Mapping::build(const Basis& b, int n0, int n1, int n2) {
  np0 = n0; np1 = n1; np2 = n2;
  // fits_in_grid, src/Basis.cpp:139-144
  if (!(b.idxmax[0] < np0 / 2 && b.idxmin[0] >= -np0 / 2 &&
        b.idxmax[1] < np1 / 2 && b.idxmin[1] >= -np1 / 2 &&
        b.idxmax[2] < np2 / 2 && b.idxmin[2] >= -np2 / 2))
    throw std::runtime_error("basis does not fit in grid");
  ntrans0 = std::max(std::abs(b.idxmax[1]), std::abs(b.idxmin[1])) + 1;
  if (!b.real) {
    // complex basis: one column per rod, ip only (:90-93, :166-188, :208-220)
    nvec = b.nrods;
    ip.assign(b.size, 0);
    im.assign(b.size, -1);
    int ig = 0;
    for (int irod = 0; irod < b.nrods; irod++)
      for (int i = 0; i < b.rod_size[irod]; i++, ig++) {
        int iz = i + b.rod_lmin[irod];
        if (iz < 0) iz += np2;
        ip[ig] = irod * np2 + iz;
      }
    zvec_to_val.clear();
    for (int irod = 0; irod < b.nrods; irod++) {
      int h = b.rod_h[irod], k = b.rod_k[irod];
      if (h < 0) h += np0;
      if (k < 0) k += np1;
      zvec_to_val.push_back(h + np0 * k);
    }
    return;
  }
  nvec = 2 * b.nrods - 1;
  ip.assign(b.size, 0);
  im.assign(b.size, 0);
  int ig = 0;
  ip[0] = 0; im[0] = 0; ig++;
  for (int l = 1; l < b.rod_size[0]; l++) {
    ip[ig] = l;
    im[ig] = np2 - l;
    ig++;
  }
  for (int irod = 1; irod < b.nrods; irod++) {
    for (int i = 0; i < b.rod_size[irod]; i++) {
      const int l = i + b.rod_lmin[irod];
      int izp = l, izm = -l;
      if (izp < 0) izp += np2;
      if (izm < 0) izm += np2;
      ip[ig] = (2 * irod - 1) * np2 + izp;
      im[ig] = (2 * irod) * np2 + izm;
      ig++;
    }
  }
  assert(ig == b.size);
  zvec_to_val.clear();
  zvec_to_val.push_back(0);
  for (int irod = 1; irod < b.nrods; irod++) {
    int hp = b.rod_h[irod], kp = b.rod_k[irod];
    if (hp < 0) hp += np0;
    if (kp < 0) kp += np1;
    int hm = -hp, km = -kp;
    if (hm < 0) hm += np0;
    if (km < 0) km += np1;
    zvec_to_val.push_back(hp + np0 * kp);
    zvec_to_val.push_back(hm + np0 * km);
  }
}

}  // namespace qbo

// ref_shim.cpp -- C entry points over the reference's own spline, LDA and
// UnitCell sources (compiled from /root/reference/src by `make ref` into
// _ref/libqbref.so).  Test infrastructure: used only to pin the oracle.
#include "spline.h"
#include "LDAFunctional.h"
#include "UnitCell.h"
#include "Base64Transcoder.h"
#include <sstream>
#include <cstring>
#include <vector>

extern "C" {

void ref_spline(int n, double* x, double* y, double yp1, double ypn, int bl,
                int br, double* y2) {
  spline(n, x, y, yp1, ypn, bl, br, y2);
}

void ref_splintd(int n, double* xa, double* ya, double* y2a, double x,
                 double* y, double* dy) {
  splintd(n, xa, ya, y2a, x, y, dy);
}

void ref_lda(int n, const double* rho, double* e, double* v) {
  std::vector<std::vector<double> > rhoe(1, std::vector<double>(rho, rho + n));
  LDAFunctional f(rhoe);
  f.setxc();
  for (int i = 0; i < n; i++) { e[i] = f.exc[i]; v[i] = f.vxc1[i]; }
}

void ref_lsda(int n, const double* rup, const double* rdn, double* e, double* vup,
              double* vdn) {
  std::vector<std::vector<double> > rhoe(2);
  rhoe[0].assign(rup, rup + n);
  rhoe[1].assign(rdn, rdn + n);
  LDAFunctional f(rhoe);
  f.setxc();
  for (int i = 0; i < n; i++) { e[i] = f.exc[i]; vup[i] = f.vxc1_up[i]; vdn[i] = f.vxc1_dn[i]; }
}

// base64 of nbytes bytes as Qbox writes it in a grid_function (encode, then
// print with 72-character lines); returns the length written to out (cap)
long ref_b64_print(long nbytes, const unsigned char* in, char* out, long cap) {
  Base64Transcoder x;
  const size_t n = x.nchars((size_t)nbytes);
  std::vector<char> b(n + 1);
  x.encode((size_t)nbytes, in, b.data());
  std::ostringstream os;
  x.print(n, b.data(), os);
  const std::string s = os.str();
  if ((long)s.size() > cap) return -1;
  std::memcpy(out, s.data(), s.size());
  return (long)s.size();
}

// decode base64 text (whitespace removed by the caller) into nbytes bytes
long ref_b64_decode(long nchars, const char* in, unsigned char* out, long nbytes) {
  Base64Transcoder x;
  (void)nbytes;
  return (long)x.decode((size_t)nchars, in, out);
}

void ref_cell(const double* a9, double* b9, double* vol) {
  UnitCell c(D3vector(a9[0], a9[1], a9[2]), D3vector(a9[3], a9[4], a9[5]),
             D3vector(a9[6], a9[7], a9[8]));
  for (int i = 0; i < 3; i++) {
    b9[3 * i] = c.b(i).x; b9[3 * i + 1] = c.b(i).y; b9[3 * i + 2] = c.b(i).z;
  }
  *vol = c.volume();
}

void ref_fold_in_ws(const double* a9, double* v) {
  UnitCell c(D3vector(a9[0], a9[1], a9[2]), D3vector(a9[3], a9[4], a9[5]),
             D3vector(a9[6], a9[7], a9[8]));
  D3vector w(v[0], v[1], v[2]);
  c.fold_in_ws(w);
  v[0] = w.x; v[1] = w.y; v[2] = w.z;
}

}  // extern "C"

// fft4_reg.cpp -- registry of the compiled four-step line lengths.
#include "internal.h"

namespace qbg {

const F4Ops* f4_ops_a(int n);
const F4Ops* f4_ops_b(int n);
const F4Ops* f4_ops_c(int n);
const F4Ops* f4_ops_d(int n);

const F4Ops* f4_ops(int n) {
  if (const F4Ops* o = f4_ops_a(n)) return o;
  if (const F4Ops* o = f4_ops_b(n)) return o;
  if (const F4Ops* o = f4_ops_c(n)) return o;
  if (const F4Ops* o = f4_ops_d(n)) return o;
  return nullptr;
}

}  // namespace qbg

// fft4_reg.cpp -- registry of the compiled four-step line lengths: the
// default split of each length (fft4_a..d) and the alternatives the grid's
// tuner may pick instead (fft4_e, fft4_f).
#include "internal.h"

namespace qbg {

const F4Ops* f4_table_a(int* count);
const F4Ops* f4_table_b(int* count);
const F4Ops* f4_table_c(int* count);
const F4Ops* f4_table_d(int* count);
const F4Ops* f4_table_e(int* count);
const F4Ops* f4_table_f(int* count);

namespace {
// defaults first, so the first match of a length is its default split
const F4Ops* (*const kTables[])(int*) = {f4_table_a, f4_table_b, f4_table_c,
                                          f4_table_d, f4_table_e, f4_table_f};
}  // namespace

const F4Ops* f4_ops(int n) {
  const F4Ops* out = nullptr;
  return f4_candidates(n, &out, 1) ? out : nullptr;
}

int f4_candidates(int n, const F4Ops** out, int cap) {
  int k = 0;
  for (auto table : kTables) {
    int cnt = 0;
    const F4Ops* t = table(&cnt);
    for (int i = 0; i < cnt; i++)
      if (t[i].n == n && k < cap) out[k++] = &t[i];
  }
  return k;
}

}  // namespace qbg

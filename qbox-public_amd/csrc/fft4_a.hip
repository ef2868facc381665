// fft4_a.hip -- four-step pass kernels instantiated for line lengths
// 22, 24, 28, 64, 90 (see fft4_kernels.h).
#include "fft4_kernels.h"

namespace qbg {

#define QBG_F4(A, B)                                                          \
  {A * B, A, B, f4::Launch<A, B>::L, &f4::Launch<A, B>::zbwd, &f4::Launch<A, B>::zfwd, \
   &f4::Launch<A, B>::xpass, &f4::Launch<A, B>::ypass}

static const F4Ops ops_a[] = {
    QBG_F4(11, 2),   // 22
    QBG_F4(8, 3),   // 24
    QBG_F4(7, 4),   // 28
    QBG_F4(8, 8),   // 64
    QBG_F4(10, 9),   // 90
};

const F4Ops* f4_table_a(int* count) {
  *count = (int)(sizeof(ops_a) / sizeof(ops_a[0]));
  return ops_a;
}

}  // namespace qbg

// fft4_b.hip -- four-step pass kernels instantiated for line lengths
// 96, 112, 126, 128 (see fft4_kernels.h).
#include "fft4_kernels.h"

namespace qbg {

#define QBG_F4(A, B)                                                          \
  {A * B, A, B, f4::Launch<A, B>::L, &f4::Launch<A, B>::zbwd, &f4::Launch<A, B>::zfwd, \
   &f4::Launch<A, B>::xpass, &f4::Launch<A, B>::ypass}

static const F4Ops ops_b[] = {
    QBG_F4(12, 8),   // 96
    QBG_F4(16, 7),   // 112
    QBG_F4(14, 9),   // 126
    QBG_F4(16, 8),   // 128
};

const F4Ops* f4_table_b(int* count) {
  *count = (int)(sizeof(ops_b) / sizeof(ops_b[0]));
  return ops_b;
}

}  // namespace qbg

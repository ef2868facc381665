// fft4_c.hip -- four-step pass kernels instantiated for line lengths
// 144, 160, 168 (see fft4_kernels.h).
#include "fft4_kernels.h"

namespace qbg {

#define QBG_F4(A, B)                                                          \
  {A * B, A, B, f4::Launch<A, B>::L, &f4::Launch<A, B>::zbwd, &f4::Launch<A, B>::zfwd, \
   &f4::Launch<A, B>::xpass, &f4::Launch<A, B>::ypass}

static const F4Ops ops_c[] = {
    QBG_F4(16, 9),   // 144
    QBG_F4(16, 10),   // 160
    QBG_F4(21, 8),   // 168
};

const F4Ops* f4_table_c(int* count) {
  *count = (int)(sizeof(ops_c) / sizeof(ops_c[0]));
  return ops_c;
}

}  // namespace qbg

// fft4_d.hip -- four-step pass kernels instantiated for line lengths
// 198, 210, 224 (see fft4_kernels.h).
#include "fft4_kernels.h"

namespace qbg {

#define QBG_F4(A, B)                                                          \
  {A * B, A, B, f4::Launch<A, B>::L, &f4::Launch<A, B>::zbwd, &f4::Launch<A, B>::zfwd, \
   &f4::Launch<A, B>::xpass, &f4::Launch<A, B>::ypass}

static const F4Ops ops_d[] = {
    QBG_F4(18, 11),   // 198
    QBG_F4(15, 14),   // 210
    QBG_F4(16, 14),   // 224
};

const F4Ops* f4_table_d(int* count) {
  *count = (int)(sizeof(ops_d) / sizeof(ops_d[0]));
  return ops_d;
}

}  // namespace qbg

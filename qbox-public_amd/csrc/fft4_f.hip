// fft4_f.hip -- alternative line splits of the lengths 160, 168, 198, 210
// and 224 (see fft4_kernels.h); the grid's tuner times them against the
// default splits and keeps the faster per axis and precision.
#include "fft4_kernels.h"

namespace qbg {

#define QBG_F4(A, B)                                                          \
  {A * B, A, B, f4::Launch<A, B>::L, &f4::Launch<A, B>::zbwd, &f4::Launch<A, B>::zfwd, \
   &f4::Launch<A, B>::xpass, &f4::Launch<A, B>::ypass}

static const F4Ops ops_f[] = {
    QBG_F4(20, 8),    // 160
    QBG_F4(14, 12),   // 168
    QBG_F4(22, 9),    // 198
    QBG_F4(14, 15),   // 210
    QBG_F4(32, 7),    // 224
};

const F4Ops* f4_table_f(int* count) {
  *count = (int)(sizeof(ops_f) / sizeof(ops_f[0]));
  return ops_f;
}

}  // namespace qbg

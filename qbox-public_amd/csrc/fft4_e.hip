// fft4_e.hip -- alternative line splits of the lengths 90, 96, 112, 126, 128
// and 144 (see fft4_kernels.h); the grid's tuner times them against the
// default splits and keeps the faster per axis and precision.
#include "fft4_kernels.h"

namespace qbg {

#define QBG_F4(A, B)                                                          \
  {A * B, A, B, f4::Launch<A, B>::L, &f4::Launch<A, B>::zbwd, &f4::Launch<A, B>::zfwd, \
   &f4::Launch<A, B>::xpass, &f4::Launch<A, B>::ypass}

static const F4Ops ops_e[] = {
    QBG_F4(18, 5),    // 90
    QBG_F4(16, 6),    // 96
    QBG_F4(14, 8),    // 112
    QBG_F4(18, 7),    // 126
    QBG_F4(32, 4),    // 128
    QBG_F4(12, 12),   // 144
};

const F4Ops* f4_table_e(int* count) {
  *count = (int)(sizeof(ops_e) / sizeof(ops_e[0]));
  return ops_e;
}

}  // namespace qbg

// SU(3) two-row (12-real) link layout: packed-spinor dslash instantiations
// for every storage precision (FP32 / FP64 arithmetic as the native ones).
#include "inst_common.h"
using namespace qmrhs;

#define QM_R12(R, CS, A)                                                                \
  QM_REG(3, R, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE12, false, false, CS, CS, CS, A) \
  QM_REG(3, R, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE12, true, false, CS, CS, CS, A)

QM_R12(1, c32, float)
QM_R12(2, c32, float)
QM_R12(4, c32, float)
QM_R12(1, c64, double)
QM_R12(2, c64, double)
QM_R12(1, c16, float)
QM_R12(2, c16, float)
QM_R12(4, c16, float)
QM_R12(1, cb16, float)
QM_R12(2, cb16, float)
QM_R12(4, cb16, float)

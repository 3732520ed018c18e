// SU(3) eight-real link layout: packed-spinor dslash instantiations for every
// storage precision. The links are rebuilt while they are staged in LDS, so
// only LDS-staging tiles exist.
#include "inst_common.h"
using namespace qmrhs;

#define QM_R8(R, CS, A) QM_REG(3, R, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE8, true, false, CS, CS, CS, A)

QM_R8(1, c32, float)
QM_R8(2, c32, float)
QM_R8(4, c32, float)
QM_R8(1, c64, double)
QM_R8(2, c64, double)
QM_R8(1, c16, float)
QM_R8(2, c16, float)
QM_R8(4, c16, float)
QM_R8(1, cb16, float)
QM_R8(2, cb16, float)
QM_R8(4, cb16, float)

// Native-layout dslash compiled for exactly 16 right-hand sides, SU(3), links
// staged in LDS: FP16 / BF16 storage, and the solver's 16-bit-in FP32-out
// operator. The dispatcher prefers these over the run-time-k kernels.
#include "inst_common.h"
using namespace qmrhs;
#define QM_K16(R, CS, CO) QM_REG(3, R, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CS, CO, float, 16)
QM_K16(1, c16, c16)
QM_K16(2, c16, c16)
QM_K16(4, c16, c16)
QM_K16(1, cb16, cb16)
QM_K16(2, cb16, cb16)
QM_K16(4, cb16, cb16)
QM_K16(1, c16, c32)
QM_K16(2, c16, c32)
QM_K16(4, c16, c32)
QM_K16(1, cb16, c32)
QM_K16(2, cb16, c32)
QM_K16(4, cb16, c32)

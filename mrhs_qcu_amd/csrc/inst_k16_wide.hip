// Native-layout dslash compiled for exactly 16 right-hand sides (the
// benchmark and solver shape), SU(3), links staged in LDS: FP32 and FP64.
// The dispatcher prefers these over the run-time-k kernels of the same tile.
#include "inst_common.h"
using namespace qmrhs;
#define QM_K16(R, CS, CO, A) QM_REG(3, R, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CS, CO, A, 16)
QM_K16(1, c32, c32, float)
QM_K16(2, c32, c32, float)
QM_K16(4, c32, c32, float)
QM_K16(1, c64, c64, double)
QM_K16(2, c64, c64, double)

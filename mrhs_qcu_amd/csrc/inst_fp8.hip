// FP8 (OCP E4M3) spinor storage with binary16 links and FP32 arithmetic,
// SU(3): FP8 output, and FP32 output for the solver's last operator of an
// iteration; links staged in LDS; the run-time RHS count and 16 RHS.
#include "inst_common.h"
using namespace qmrhs;
#define QM_F8(R, CO, KF) QM_REG(3, R, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, c8, c16, CO, float, KF)
QM_F8(2, c8, 0)
QM_F8(4, c8, 0)
QM_F8(2, c32, 0)
QM_F8(4, c32, 0)
QM_F8(2, c8, 16)
QM_F8(4, c8, 16)
QM_F8(2, c32, 16)
QM_F8(4, c32, 16)

// Hypercube-mapped dslash instantiations (dslash_hcube.h): packed layout,
// native links staged in LDS, k / R lanes per site in {1, 2, 4, 8}.
#include "inst_common.h"
using namespace qmrhs;
#define QM_HC(N, R, CS, CO, A, KF) static ::qmrhs::RegHc<N, R, CS, CS, CO, A, KF> QM_CAT(qm_reg_, __COUNTER__);
QM_HC(3, 2, c32, c32, float, 16)
QM_HC(2, 1, c32, c32, float, 0)
QM_HC(2, 2, c32, c32, float, 0)
QM_HC(3, 1, c32, c32, float, 0)
QM_HC(3, 2, c32, c32, float, 0)
QM_HC(4, 1, c32, c32, float, 0)
QM_HC(4, 2, c32, c32, float, 0)
QM_HC(3, 1, c64, c64, double, 0)
// 16 right-hand sides: FP32 (two per lane, and one per lane over two
// wavefronts), FP64 (one per lane over two wavefronts), FP16 / BF16 and the
// solver's 16-bit-in FP32-out operator.
QM_HC(3, 1, c32, c32, float, 16)
QM_HC(3, 1, c64, c64, double, 16)
QM_HC(3, 1, c16, c16, float, 16)
QM_HC(3, 2, c16, c16, float, 16)
QM_HC(3, 4, c16, c16, float, 16)
QM_HC(3, 1, cb16, cb16, float, 16)
QM_HC(3, 2, cb16, cb16, float, 16)
QM_HC(3, 1, c16, c32, float, 16)
QM_HC(3, 2, c16, c32, float, 16)
QM_HC(3, 4, c16, c32, float, 16)
// Run-time RHS count, 16-bit storage.
QM_HC(3, 1, c16, c16, float, 0)
QM_HC(3, 2, c16, c16, float, 0)
QM_HC(3, 1, cb16, cb16, float, 0)
QM_HC(3, 2, cb16, cb16, float, 0)
QM_HC(3, 1, cb16, c32, float, 16)
QM_HC(3, 2, cb16, c32, float, 16)
// User AoS layout through per-RHS pointers (the qcu.h startDslash path), one
// RHS per lane.
#define QM_HC_AOS(N, C, A) static ::qmrhs::RegHc<N, 1, C, C, C, A, 0, ::qmrhs::SL_AOS_PTRS> QM_CAT(qm_reg_, __COUNTER__);
QM_HC_AOS(3, c32, float)
QM_HC_AOS(3, c64, double)
QM_HC_AOS(3, c16, float)
// FP8 (e4m3) spinors with binary16 links, and the FP8-sloppy solver's
// FP8-in FP32-out operator.
#define QM_HC_FP8(R, CO) static ::qmrhs::RegHc<3, R, ::qmrhs::c8, ::qmrhs::c16, CO, float, 16> QM_CAT(qm_reg_, __COUNTER__);
QM_HC_FP8(2, c8)
QM_HC_FP8(4, c8)
QM_HC_FP8(2, c32)
QM_HC_FP8(4, c32)

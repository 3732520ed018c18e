// Sloppy-storage (FP16 / BF16) input with FP32 output: the last operator
// application of each mixed-precision CG iteration.
#include "inst_common.h"
using namespace qmrhs;
QM_MIXED_TILES(2, c16)
QM_MIXED_TILES(3, c16)
QM_MIXED_TILES(4, c16)
QM_MIXED_TILES(5, c16)
QM_MIXED_R4(2, c16)
QM_MIXED_R4(3, c16)
QM_MIXED_BASIC(6, c16)
QM_MIXED_BASIC(7, c16)
QM_MIXED_BASIC(8, c16)
QM_MIXED_TILES(2, cb16)
QM_MIXED_TILES(3, cb16)
QM_MIXED_TILES(4, cb16)
QM_MIXED_TILES(5, cb16)
QM_MIXED_R4(2, cb16)
QM_MIXED_R4(3, cb16)
QM_MIXED_BASIC(6, cb16)
QM_MIXED_BASIC(7, cb16)
QM_MIXED_BASIC(8, cb16)

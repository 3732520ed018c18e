// BF16-storage (FP32 accumulate) native-layout dslash instantiations.
#include "inst_common.h"
using namespace qmrhs;
QM_PACKED_TILES(2, cb16, cb16, cb16, float)
QM_PACKED_TILES(3, cb16, cb16, cb16, float)
QM_PACKED_TILES(4, cb16, cb16, cb16, float)
QM_PACKED_TILES(5, cb16, cb16, cb16, float)
QM_PACKED_R4(2, cb16, cb16, cb16, float)
QM_PACKED_R4(3, cb16, cb16, cb16, float)
QM_PACKED_BASIC(6, cb16, cb16, cb16, float)
QM_PACKED_BASIC(7, cb16, cb16, cb16, float)
QM_PACKED_BASIC(8, cb16, cb16, cb16, float)

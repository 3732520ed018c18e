// FP16-storage (FP32 accumulate) native-layout dslash instantiations.
#include "inst_common.h"
using namespace qmrhs;
QM_PACKED_TILES(2, c16, c16, c16, float)
QM_PACKED_TILES(3, c16, c16, c16, float)
QM_PACKED_TILES(4, c16, c16, c16, float)
QM_PACKED_TILES(5, c16, c16, c16, float)
QM_PACKED_R4(2, c16, c16, c16, float)
QM_PACKED_R4(3, c16, c16, c16, float)
QM_PACKED_BASIC(6, c16, c16, c16, float)
QM_PACKED_BASIC(7, c16, c16, c16, float)
QM_PACKED_BASIC(8, c16, c16, c16, float)

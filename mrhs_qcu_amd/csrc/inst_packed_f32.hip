// FP32 native-layout dslash instantiations.
#include "inst_common.h"
using namespace qmrhs;
QM_PACKED_TILES(2, c32, c32, c32, float)
QM_PACKED_TILES(3, c32, c32, c32, float)
QM_PACKED_TILES(4, c32, c32, c32, float)
QM_PACKED_TILES(5, c32, c32, c32, float)
QM_PACKED_R4(2, c32, c32, c32, float)
QM_PACKED_R4(3, c32, c32, c32, float)
QM_PACKED_BASIC(6, c32, c32, c32, float)
QM_PACKED_BASIC(7, c32, c32, c32, float)
QM_PACKED_BASIC(8, c32, c32, c32, float)

// FP64 native-layout dslash instantiations.
#include "inst_common.h"
using namespace qmrhs;
QM_PACKED_TILES(2, c64, c64, c64, double)
QM_PACKED_TILES(3, c64, c64, c64, double)
QM_PACKED_TILES(4, c64, c64, c64, double)
QM_PACKED_TILES(5, c64, c64, c64, double)
QM_PACKED_BASIC(6, c64, c64, c64, double)
QM_PACKED_BASIC(7, c64, c64, c64, double)
QM_PACKED_BASIC(8, c64, c64, c64, double)

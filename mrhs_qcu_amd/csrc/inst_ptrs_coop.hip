// User AoS pointer dslash read and written by site groups (SL_AOS_COOP):
// the qcu.h and qcuMrhsDslashUser path for power-of-two RHS counts.
#include "inst_common.h"
using namespace qmrhs;
QM_PTRS_COOP(2, c32, float)
QM_PTRS_COOP(3, c32, float)
QM_PTRS_COOP(4, c32, float)
QM_PTRS_COOP(2, c16, float)
QM_PTRS_COOP(3, c16, float)
QM_PTRS_COOP(4, c16, float)

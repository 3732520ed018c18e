// Layout shifts between the user AoS layout and the reference SoA layout
// (interface of the reference's include/storage/qcu_storage.h:5-6).
//
// The two reference signatures keep their meaning: complex<double>, N = 3.
// The overloads with nColors / bytesPerComplex shift any colour count and any
// element width (8, 16 or 4 bytes per complex) with the same index maps.
#pragma once
#include "qcu_enum.h"
namespace qcu {
namespace storage {
void shiftVectorStorage(void *dstVec, void *srcVec, ShiftDirection shift_direction, int subLx, int Ly, int Lz, int Lt);
void shiftGaugeStorage(void *dst_vec, void *src_vec, ShiftDirection shift_direction, int Lx, int Ly, int Lz, int Lt);

void shiftVectorStorage(void *dstVec, void *srcVec, ShiftDirection shift_direction, int subLx, int Ly, int Lz, int Lt,
                        int nColors, int bytesPerComplex);
void shiftGaugeStorage(void *dst_vec, void *src_vec, ShiftDirection shift_direction, int Lx, int Ly, int Lz, int Lt,
                       int nColors, int bytesPerComplex);
}  // namespace storage
}  // namespace qcu

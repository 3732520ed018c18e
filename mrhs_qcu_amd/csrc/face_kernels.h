// Face exchange kernels of a domain-decomposed hopping term (host entry
// points, asynchronous on `st`). Fields are packed [S][4][N][k], the gauge is
// the native layout, faces are [F][2][N][k]; all in one storage precision.
#pragma once
#include <hip/hip_runtime.h>

#include "dslash_dispatch.h"

namespace qmrhs {

// Sites of one parity on one face of dimension mu (F).
long long face_sites(const Geom &g, int mu);
// Half spinors of the input field (parity 1 - parity) on both faces of mu.
hipError_t launch_pack_faces(void *to_bwd, void *to_fwd, const void *in, const void *gauge, int prec, int k,
                             int parity, int dagger, const Geom &geom, int mu, hipStream_t st);
// out (parity `parity`) += the boundary terms of mu built from received faces.
hipError_t launch_dslash_faces(void *out, const void *from_bwd, const void *from_fwd, const void *gauge, int prec,
                               int k, int parity, int dagger, const Geom &geom, int mu, hipStream_t st);

}  // namespace qmrhs

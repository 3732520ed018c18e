// Clover term kernels (host entry points, asynchronous on `st`).
#pragma once
#include <hip/hip_runtime.h>

#include "qcu_device.h"

namespace qmrhs {

// sigma_{mu nu} for the 6 planes (mu < nu, lexicographic), as its two
// diagonal 2x2 spin blocks (spins {0,1} and {2,3}).
struct SigmaBlocks {
  double re[6][2][2][2], im[6][2][2][2];
};

// A = 1 + coeff sum sigma F on both parities, FP64 [2][S][2][2N][2N]; gauge
// is the user layout [4][2][S][N][N] in FP64 or FP32 (gauge_prec). N <= 5.
hipError_t clover_compute(c64 *a, const void *gauge, int gauge_prec, const int *L, int N, double coeff,
                          const SigmaBlocks &sig, hipStream_t st);
// Inverse of each (2N x 2N) block (nblocks of them), FP64.
hipError_t clover_invert(c64 *inv, const c64 *a, long long nblocks, int N, hipStream_t st);
// out = A in on the sites of `parity` (packed spinors, clover field in prec).
hipError_t clover_apply(int prec, void *out, const void *in, const void *clov, int parity, int S, int k, int N,
                        hipStream_t st);

}  // namespace qmrhs

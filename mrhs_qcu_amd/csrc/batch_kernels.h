// Work-list launch of the batched hopping term: the workgroups of several
// packed-layout FP32 problems (different lattices, N = 2..5, nrhs) in one
// kernel. Each problem's workgroups form a contiguous range that starts at a
// multiple of the XCD count, so the problem's own XCD-aware workgroup order
// applies unchanged inside its range.
#pragma once

#include <hip/hip_runtime.h>

#include "dslash_kernel.h"

namespace qmrhs {

struct BatchItem {
  DslashArgs a;    // the problem's kernel arguments (as launch_dslash would pass them)
  uint32_t first;  // first workgroup of the problem's range (multiple of kNumXcd)
  int n;           // colour count: selects the compiled body
};

// Tile family of the work-list kernel per colour count (the rhs_per_lane and
// link staging its body is compiled with): N = 2..4 two RHS per lane with
// links staged in LDS, N = 5 one RHS per lane without.
inline bool batch_body_tile(int N, int *R, int *lds)
{
  if (N >= 2 && N <= 4) {
    *R = 2;
    *lds = 1;
    return true;
  }
  if (N == 5) {
    *R = 1;
    *lds = 0;
    return true;
  }
  return false;
}
constexpr int kBatchBlock = 128;  // lanes per workgroup of every problem in a work list

hipError_t launch_batch_worklist(const BatchItem *items_dev, int nitems, uint32_t nblocks, size_t shm,
                                 hipStream_t st);

}  // namespace qmrhs

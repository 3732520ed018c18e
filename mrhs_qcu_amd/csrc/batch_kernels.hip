// Work-list kernel of the batched hopping term (see batch_kernels.h).
#include "batch_kernels.h"

namespace qmrhs {
namespace {

__global__ void __launch_bounds__(256) batch_dslash_kernel(const BatchItem *__restrict__ items, int nitems)
{
  const uint32_t hw = blockIdx.x;
  int p = 0;
  while (p + 1 < nitems && hw >= items[p + 1].first) ++p;
  const DslashArgs a = items[p].a;
  const uint32_t local = hw - items[p].first;
  // Padding workgroups of a range leave before any barrier (the body's own
  // column-order padding likewise).
  if (local >= a.ngrid) return;
  switch (items[p].n) {
    case 2: dslash_body<2, 2, SL_PACKED, GL_NATIVE, true, false, c32, c32, c32, float>(a, local); break;
    case 3: dslash_body<3, 2, SL_PACKED, GL_NATIVE, true, false, c32, c32, c32, float>(a, local); break;
    case 4: dslash_body<4, 2, SL_PACKED, GL_NATIVE, true, false, c32, c32, c32, float>(a, local); break;
    default: dslash_body<5, 1, SL_PACKED, GL_NATIVE, false, false, c32, c32, c32, float>(a, local); break;
  }
}

}  // namespace

hipError_t launch_batch_worklist(const BatchItem *items_dev, int nitems, uint32_t nblocks, size_t shm,
                                 hipStream_t st)
{
  hipLaunchKernelGGL(batch_dslash_kernel, dim3(nblocks), dim3(kBatchBlock), shm, st, items_dev, nitems);
  return hipGetLastError();
}

}  // namespace qmrhs

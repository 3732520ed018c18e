// Pack of k right-hand-side pointers (interface of the reference's
// include/qcuDesc/mrhs_desc.h:5-10).
#pragma once

#include <cstddef>

#include <hip/hip_runtime.h>

namespace qcu {

template <size_t _Nvec>
struct MrhsInput {
  void *fermion[_Nvec];
  void *data() { return fermion; }
  __device__ __host__ void *operator[](size_t i) { return fermion[i]; }
};

}  // namespace qcu

// Execution streams of the library (interface of the reference's
// include/qcu_kernel_param.cuh:9-38, on HIP streams).
//
// One main stream carries every launch; the eight boundary streams (one per
// dimension and direction) are the face streams a domain-decomposed run
// overlaps with the interior launch. The holder owns its streams and cannot be
// copied, so a stream is destroyed exactly once.
#pragma once
#include <hip/hip_runtime.h>

#include "qcu_macro.cuh"

namespace qcu {

struct ExecutionStreams {
  hipStream_t stream;
  hipStream_t boundaryStreams[4 * 2];  // 4 dims, 2 dirs

  ExecutionStreams()
  {
    CHECK_HIP(hipStreamCreate(&stream));
    for (int i = 0; i < 4 * 2; i++) CHECK_HIP(hipStreamCreate(&boundaryStreams[i]));
  }
  ExecutionStreams(const ExecutionStreams &) = delete;
  ExecutionStreams &operator=(const ExecutionStreams &) = delete;

  ~ExecutionStreams()
  {
    // Destruction at process exit may run after the HIP runtime has been torn
    // down; errors are ignored rather than turned into a non-zero exit.
    (void)hipStreamDestroy(stream);
    for (int i = 0; i < 4 * 2; i++) (void)hipStreamDestroy(boundaryStreams[i]);
  }
};

}  // namespace qcu

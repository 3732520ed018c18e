// Streaming-bandwidth kernels: the ceiling against which the dslash's
// roofline fraction is read. Every access is a 16-byte vector load or store;
// each thread handles kUnroll chunks spaced one workgroup apart (consecutive
// lanes on consecutive 16-byte chunks), and the grid covers the buffer once,
// so every CU takes part and no chunk is touched twice per pass.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "api_internal.h"
#include "storage_api.h"

namespace qmrhs {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int kBlock = 256;
constexpr int kUnroll = 4;

// mode 0: read src (the sum of what was read is stored only if it equals a
//         value no finite input produces, so the loads cannot be dropped);
// mode 1: write dst; mode 2: copy src -> dst; mode 3: dst = dst * one in place;
// mode 4: copy src's 8-byte elements to dst at a 128-byte stride (one FP32
//         right-hand side into its slot of a 16-RHS packed field: every
//         128-byte line of dst receives 8 bytes).
template <int MODE>
__global__ void __launch_bounds__(kBlock) stream_kernel(f4 *dst, const f4 *src, size_t n16, float one)
{
  const size_t base = size_t(blockIdx.x) * kBlock * kUnroll + threadIdx.x;
  f4 v[kUnroll];
  if constexpr (MODE == 0 || MODE == 2 || MODE == 3 || MODE == 4) {
    const f4 *p = MODE == 3 ? dst : src;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = base + size_t(u) * kBlock;
      v[u] = i < n16 ? p[i] : f4{0.f, 0.f, 0.f, 0.f};
    }
  }
  if constexpr (MODE == 0) {
    f4 s = v[0];
#pragma unroll
    for (int u = 1; u < kUnroll; ++u) s += v[u];
    const float t = s.x + s.y + s.z + s.w;
    if (t == 3.0e38f * one) dst[0] = s;
  } else {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = base + size_t(u) * kBlock;
      if (i >= n16) break;
      if constexpr (MODE == 1) {
        dst[i] = f4{one, one, one, one};
      } else if constexpr (MODE == 2) {
        dst[i] = v[u];
      } else if constexpr (MODE == 4) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        f2 *d2 = reinterpret_cast<f2 *>(dst);  // element e -> d2[16 e]
        d2[32 * i] = f2{v[u].x, v[u].y};
        d2[32 * i + 16] = f2{v[u].z, v[u].w};
      } else {
        dst[i] = v[u] * one;
      }
    }
  }
}

constexpr unsigned kCopyGrid = 1024;

__global__ void __launch_bounds__(kBlock) copy_kernel(f4 *__restrict__ dst, const f4 *__restrict__ src, size_t n16)
{
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n16; i += size_t(gridDim.x) * kBlock) dst[i] = src[i];
}

int launch_stream(int mode, void *dst, const void *src, size_t n16, hipStream_t st)
{
  const size_t per_block = size_t(kBlock) * kUnroll;
  const size_t grid = (n16 + per_block - 1) / per_block;
  f4 *d = static_cast<f4 *>(dst);
  const f4 *s = static_cast<const f4 *>(src);
  switch (mode) {
    case 0: hipLaunchKernelGGL(stream_kernel<0>, dim3(unsigned(grid)), dim3(kBlock), 0, st, d, s, n16, 1.0f); break;
    case 1: hipLaunchKernelGGL(stream_kernel<1>, dim3(unsigned(grid)), dim3(kBlock), 0, st, d, s, n16, 1.0f); break;
    case 2: hipLaunchKernelGGL(stream_kernel<2>, dim3(unsigned(grid)), dim3(kBlock), 0, st, d, s, n16, 1.0f); break;
    case 3: hipLaunchKernelGGL(stream_kernel<3>, dim3(unsigned(grid)), dim3(kBlock), 0, st, d, s, n16, 1.0f); break;
    default: hipLaunchKernelGGL(stream_kernel<4>, dim3(unsigned(grid)), dim3(kBlock), 0, st, d, s, n16, 1.0f); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

int copy_device(void *dst, const void *src, size_t bytes, hipStream_t st, std::string *err)
{
  if (bytes == 0) return 0;
  if (bytes % 16 || (reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16) {
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return *err = "device copy failed", 1;
    return 0;
  }
  const size_t n16 = bytes / 16;
  const size_t need = (n16 + kBlock - 1) / kBlock;
  const unsigned grid = unsigned(need < kCopyGrid ? need : kCopyGrid);
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(kBlock), 0, st, static_cast<f4 *>(dst), static_cast<const f4 *>(src),
                     n16);
  if (hipGetLastError() != hipSuccess) return *err = "device copy launch failed", 1;
  return 0;
}

}  // namespace qmrhs

using namespace qmrhs;

extern "C" int qcuMrhsStreamBandwidth(int mode, void *dst, const void *src, size_t bytes, int reps, double *gbs,
                                      void *stream)
{
  if (mode < 0 || mode > 4)
    return api_fail("stream mode must be 0 (read), 1 (write), 2 (copy), 3 (read+write) or 4 (strided copy)");
  if (bytes < 16 || bytes % 16) return api_fail("stream bytes must be a positive multiple of 16");
  if (bytes / 16 / (size_t(kBlock) * kUnroll) >= (size_t(1) << 31)) return api_fail("stream buffer too large");
  if (reps < 1) return api_fail("reps < 1");
  const bool need_src = mode == 0 || mode == 2 || mode == 4, need_dst = mode != 0;
  if ((need_src && !src) || (need_dst && !dst) || (mode == 0 && !dst))
    return api_fail("null stream buffer (mode 0 needs 16 bytes at dst for its guard store)");
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16)
    return api_fail("stream buffers must be 16-byte aligned");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t n16 = bytes / 16;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return api_fail("hipEventCreate failed");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return api_fail("hipEventCreate failed");
  }
  int rc = launch_stream(mode, dst, src, n16, st);  // warm-up pass
  if (!rc) rc = hipEventRecord(e0, st) != hipSuccess;
  for (int r = 0; r < reps && !rc; ++r) rc = launch_stream(mode, dst, src, n16, st);
  if (!rc) rc = hipEventRecord(e1, st) != hipSuccess;
  if (!rc) rc = hipEventSynchronize(e1) != hipSuccess;
  float ms = 0.f;
  if (!rc) rc = hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return api_fail("stream kernel launch or timing failed");
  const double moved = double(bytes) * (mode >= 2 ? 2.0 : 1.0) * reps;  // mode 4: useful bytes
  if (gbs) *gbs = moved / (double(ms) * 1e-3) / 1e9;
  return 0;
}

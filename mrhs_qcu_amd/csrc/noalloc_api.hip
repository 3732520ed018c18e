// qcuMrhsDslashExNoAlloc: qcuMrhsDslashEx without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc). The per-workgroup norm partials and
// the reduction's stage buffer live in one file-scope device array, `scratch`,
// of kScratchBytes; every call on a GPU shares that GPU's copy.
//
// The array's name is a host-side stand-in, not a device address: its address
// on each device is taken once with hipGetSymbolAddress (with that device
// current) and only that address is handed to kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <string>

#include "api_internal.h"

namespace {

// 16 MiB: the partials of one launch are (workgroups + 256) x nrhs doubles,
// 2.6 MiB at 24^3x48 with 16 RHS and the tuned FP32 tile, 8.4 MiB at 32^3x64.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
double *scratch_ptr(std::string *err)
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *err = "no current HIP device";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

using namespace qmrhs;

extern "C" int qcuMrhsDslashExNoAlloc(void *out, int out_prec, const void *in, const void *x, int in_prec, double a,
                                      double b, const float *sc_in, const float *sc_x, const float *sc_out,
                                      double *norm2_out, const void *gauge, int nrhs, int parity, int dagger,
                                      const qcuMrhsGeom *gc, void *stream)
{
  std::string err;
  DslashCall c;
  if (!dslash_ex_call(&c, out, out_prec, in, x, in_prec, a, b, sc_in, sc_x, sc_out, gauge, nrhs, parity, dagger, gc,
                      stream, &err))
    return api_fail(err);
  double *buf = nullptr;
  if (norm2_out) {
    const size_t need = dslash_ex_scratch_bytes(c);
    if (need > kScratchBytes) {
      char msg[200];
      snprintf(msg, sizeof msg,
               "qcuMrhsDslashExNoAlloc: the norms need %zu bytes of scratch, more than its fixed %zu bytes "
               "(use qcuMrhsDslashEx, or fewer right-hand sides per call)",
               need, kScratchBytes);
      return api_fail(msg);
    }
    buf = scratch_ptr(&err);
    if (!buf) return api_fail(err);
  }
  if (dslash_ex_launch(c, norm2_out, buf, &err)) return api_fail(err);
  return 0;
}

// C ABI of the compressed SU(3) link layouts: two rows (12 reals per link
// instead of 18: qcuMrhsGauge12Bytes, qcuMrhsPackGauge12, qcuMrhsDslash12)
// and eight reals (qcuMrhsGauge8Bytes, qcuMrhsPackGauge8, qcuMrhsDslash8).
#include <hip/hip_runtime.h>

#include <string>

#include "api_internal.h"
#include "qcu_mrhs.h"
#include "storage_api.h"

namespace qmrhs {
namespace {

// Distance from SU(3) a link may have and still be stored compressed.
constexpr double kSu3Tol = 1e-5;
// Smallest |U01|^2 + |U02|^2 the eight-real rebuild divides by: storage
// errors grow by up to 1 / sqrt of it (1000x at the limit).
constexpr double kMinN8 = 1e-6;

// Device scratch of one synchronous call; the stream is drained before the
// memory is freed, so an early error return cannot free what a launched
// kernel still uses.
struct DeviceScratch {
  void *p = nullptr;
  hipStream_t st = nullptr;
  ~DeviceScratch()
  {
    if (p) {
      (void)hipStreamSynchronize(st);
      (void)hipFree(p);
    }
  }
};

}  // namespace
}  // namespace qmrhs

using namespace qmrhs;

extern "C" {

size_t qcuMrhsGauge12Bytes(const qcuMrhsGeom *gc, int prec)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err) || g.N != 3 || !prec_bytes(prec)) return 0;
  return size_t(sites_per_parity(g)) * 2 * 8 * 6 * prec_bytes(prec);
}

int qcuMrhsPackGauge12(void *dst, int dst_prec, const void *src, int src_prec, const qcuMrhsGeom *gc, int t_sign,
                       void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (g.N != 3) return api_fail("two-row links need nColors = 3");
  if (t_sign != 1 && t_sign != -1) return api_fail("t_sign must be +1 or -1");
  if (!dst || !src) return api_fail("null gauge pointer");
  if (!prec_bytes(dst_prec)) return api_fail("unsupported gauge precision");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t native_bytes = size_t(sites_per_parity(g)) * 2 * 8 * 9 * 16;
  DeviceScratch native, bad;
  native.st = bad.st = st;
  if (hipMalloc(&native.p, native_bytes) != hipSuccess || hipMalloc(&bad.p, sizeof(int)) != hipSuccess)
    return api_fail("out of device memory for the gauge compression");
  if (pack_gauge(native.p, PREC_DOUBLE, src, src_prec, 3, g.L, st, &err)) return api_fail(err);
  if (hipMemsetAsync(bad.p, 0, sizeof(int), st) != hipSuccess) return api_fail("memset failed");
  if (compress_gauge12(dst, dst_prec, native.p, g.L, t_sign, kSu3Tol, static_cast<int *>(bad.p), st, &err))
    return api_fail(err);
  int nbad = 0;
  if (hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return api_fail("gauge compression failed");
  if (nbad)
    return api_fail(std::to_string(nbad) +
                    " links are not SU(3) to 1e-5 (two rows cannot represent them; with antiperiodic time folded "
                    "into the links pass t_sign = -1)");
  return 0;
}

int qcuMrhsDslash12(void *out, const void *in, const void *gauge12, int prec, int nrhs, int parity, int dagger,
                    int t_sign, const qcuMrhsGeom *gc, const qcuMrhsTile *tc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (t_sign != 1 && t_sign != -1) return api_fail("t_sign must be +1 or -1");
  if (nrhs < 1) return api_fail("nrhs must be >= 1");
  DslashCall c;
  c.sl = SL_PACKED;
  c.gl = GL_NATIVE12;
  c.prec_s = c.prec_g = c.prec_o = prec;
  c.geom = g;
  c.tile = tc ? tile_from_c(*tc) : tile_for(g, prec, nrhs);
  c.k = nrhs;
  c.parity = parity;
  c.dagger = dagger;
  c.out = out;
  c.in = in;
  c.gauge = gauge12;
  c.t_flip = t_sign < 0 ? 1 : 0;
  c.stream = static_cast<hipStream_t>(stream);
  if (launch_dslash(c, &err)) return api_fail(err);
  return 0;
}


size_t qcuMrhsGauge8Bytes(const qcuMrhsGeom *gc, int prec)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err) || g.N != 3 || !prec_bytes(prec)) return 0;
  return size_t(sites_per_parity(g)) * 2 * 8 * 4 * prec_bytes(prec);
}

int qcuMrhsPackGauge8(void *dst, int dst_prec, const void *src, int src_prec, const qcuMrhsGeom *gc, int t_sign,
                      void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (g.N != 3) return api_fail("eight-real links need nColors = 3");
  if (t_sign != 1 && t_sign != -1) return api_fail("t_sign must be +1 or -1");
  if (!dst || !src) return api_fail("null gauge pointer");
  if (!prec_bytes(dst_prec)) return api_fail("unsupported gauge precision");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t native_bytes = size_t(sites_per_parity(g)) * 2 * 8 * 9 * 16;
  DeviceScratch native, bad;
  native.st = bad.st = st;
  if (hipMalloc(&native.p, native_bytes) != hipSuccess || hipMalloc(&bad.p, 2 * sizeof(int)) != hipSuccess)
    return api_fail("out of device memory for the gauge compression");
  if (pack_gauge(native.p, PREC_DOUBLE, src, src_prec, 3, g.L, st, &err)) return api_fail(err);
  if (hipMemsetAsync(bad.p, 0, 2 * sizeof(int), st) != hipSuccess) return api_fail("memset failed");
  if (compress_gauge8(dst, dst_prec, native.p, g.L, t_sign, kSu3Tol, kMinN8, static_cast<int *>(bad.p), st, &err))
    return api_fail(err);
  int nbad[2] = {0, 0};
  if (hipMemcpyAsync(nbad, bad.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return api_fail("gauge compression failed");
  if (nbad[0])
    return api_fail(std::to_string(nbad[0]) +
                    " links are not SU(3) to 1e-5 (eight reals cannot represent them; with antiperiodic time "
                    "folded into the links pass t_sign = -1)");
  if (nbad[1])
    return api_fail(std::to_string(nbad[1]) +
                    " links have |U01|^2 + |U02|^2 < 1e-6 (e.g. a unit gauge), which the eight-real rebuild "
                    "divides by; store this field as two rows or full links");
  return 0;
}

int qcuMrhsDslash8(void *out, const void *in, const void *gauge8, int prec, int nrhs, int parity, int dagger,
                   int t_sign, const qcuMrhsGeom *gc, const qcuMrhsTile *tc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (t_sign != 1 && t_sign != -1) return api_fail("t_sign must be +1 or -1");
  if (nrhs < 1) return api_fail("nrhs must be >= 1");
  DslashCall c;
  c.sl = SL_PACKED;
  c.gl = GL_NATIVE8;
  c.prec_s = c.prec_g = c.prec_o = prec;
  c.geom = g;
  if (tc) {
    c.tile = tile_from_c(*tc);
  } else {
    // The native tile for this shape, with the links staged in LDS (where
    // the eight-real links are rebuilt).
    c.tile = tile_for(g, prec, nrhs);
    if (!c.tile.lds) {
      c.tile.lds = 1;
      const int R = c.tile.R >= 1 && nrhs % c.tile.R == 0 ? c.tile.R : 1;
      c.tile.R = R;
      // Fewer lanes per workgroup when the full links of its sites would not fit.
      const size_t site_bytes = size_t(8) * 9 * prec_bytes(prec);
      if (size_t((c.tile.block - 1) / (nrhs / R) + 2) * site_bytes > size_t(kMaxDynLds)) c.tile.block = 64;
    }
  }
  c.k = nrhs;
  c.parity = parity;
  c.dagger = dagger;
  c.out = out;
  c.in = in;
  c.gauge = gauge8;
  c.t_flip = t_sign < 0 ? 1 : 0;
  c.stream = static_cast<hipStream_t>(stream);
  if (launch_dslash(c, &err)) return api_fail(err);
  return 0;
}

}  // extern "C"

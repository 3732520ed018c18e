// C ABI of the clover term: qcuMrhsCloverBytes, qcuMrhsCloverCompute,
// qcuMrhsCloverApply.
#include <hip/hip_runtime.h>

#include <complex>
#include <string>

#include "api_internal.h"
#include "clover.h"
#include "qcu_mrhs.h"
#include "storage_api.h"

namespace qmrhs {
namespace {

using cd = std::complex<double>;

// DeGrand-Rossi gamma matrices (the projectors of the hop table,
// include/kernels/wilson_dslash_4d.cuh:74-290, are 1 -+ gamma_mu of these).
void gamma_matrix(int mu, cd (&gm)[4][4])
{
  const cd I(0, 1), O(0, 0), P(1, 0);
  const cd tab[4][4][4] = {
      {{O, O, O, I}, {O, O, I, O}, {O, -I, O, O}, {-I, O, O, O}},
      {{O, O, O, -P}, {O, O, P, O}, {O, P, O, O}, {-P, O, O, O}},
      {{O, O, I, O}, {O, O, O, -I}, {-I, O, O, O}, {O, I, O, O}},
      {{O, O, P, O}, {O, O, O, P}, {P, O, O, O}, {O, P, O, O}},
  };
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) gm[i][j] = tab[mu][i][j];
}

// sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu]; its off-diagonal 2x2 spin blocks
// vanish in this basis (checked).
bool sigma_blocks(SigmaBlocks *sb)
{
  int plane = 0;
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++plane) {
      cd a[4][4], b[4][4], s[4][4];
      gamma_matrix(mu, a);
      gamma_matrix(nu, b);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          cd c(0, 0);
          for (int l = 0; l < 4; ++l) c += a[i][l] * b[l][j] - b[i][l] * a[l][j];
          s[i][j] = cd(0, 0.5) * c;
        }
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          if ((i / 2) != (j / 2) && std::abs(s[i][j]) > 0) return false;
      for (int blk = 0; blk < 2; ++blk)
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            sb->re[plane][blk][i][j] = s[2 * blk + i][2 * blk + j].real();
            sb->im[plane][blk][i][j] = s[2 * blk + i][2 * blk + j].imag();
          }
    }
  return true;
}

// Device scratch of one synchronous call; the stream is drained before the
// memory is freed, so an early error return cannot free what a launched
// kernel still uses.
struct DeviceBuf {
  void *p = nullptr;
  hipStream_t st = nullptr;
  ~DeviceBuf()
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

size_t qcuMrhsCloverBytes(const qcuMrhsGeom *gc, int prec)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err) || g.N > 5 || !prec_bytes(prec)) return 0;
  return size_t(sites_per_parity(g)) * 2 * 8 * g.N * g.N * prec_bytes(prec);
}

int qcuMrhsCloverCompute(void *clover, void *clover_inv, int prec, const void *gauge, int gauge_prec, double coeff,
                         const qcuMrhsGeom *gc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (g.N > 5) return api_fail("the clover term supports nColors 2..5");
  if (!prec_bytes(prec)) return api_fail("unsupported clover precision");
  if (gauge_prec != PREC_DOUBLE && gauge_prec != PREC_SINGLE) return api_fail("the gauge must be FP64 or FP32");
  if (!clover || !gauge) return api_fail("null clover or gauge pointer");
  SigmaBlocks sb;
  if (!sigma_blocks(&sb)) return api_fail("internal: sigma_{mu nu} not block diagonal");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long S = sites_per_parity(g);
  const long long n = 2 * S * 8LL * g.N * g.N;  // complex numbers per field
  DeviceBuf a64, inv64;
  a64.st = inv64.st = st;
  const bool direct = prec == PREC_DOUBLE;
  void *a = direct ? clover : nullptr;
  if (!direct) {
    if (hipMalloc(&a64.p, size_t(n) * 16) != hipSuccess) return api_fail("out of device memory");
    a = a64.p;
  }
  hipError_t e = clover_compute(static_cast<c64 *>(a), gauge, gauge_prec, g.L, g.N, coeff, sb, st);
  if (e != hipSuccess) return api_fail(std::string("clover launch failed: ") + hipGetErrorString(e));
  if (clover_inv) {
    void *inv = direct ? clover_inv : nullptr;
    if (!direct) {
      if (hipMalloc(&inv64.p, size_t(n) * 16) != hipSuccess) return api_fail("out of device memory");
      inv = inv64.p;
    }
    e = clover_invert(static_cast<c64 *>(inv), static_cast<const c64 *>(a), 2 * S * 2, g.N, st);
    if (e != hipSuccess) return api_fail(std::string("clover inversion failed: ") + hipGetErrorString(e));
    if (!direct && convert_complex(clover_inv, prec, inv, PREC_DOUBLE, n, st, &err)) return api_fail(err);
  }
  if (!direct && convert_complex(clover, prec, a, PREC_DOUBLE, n, st, &err)) return api_fail(err);
  if (hipStreamSynchronize(st) != hipSuccess) return api_fail("clover computation failed");
  return 0;
}

int qcuMrhsCloverApply(void *out, const void *in, const void *clover, int prec, int nrhs, int parity,
                       const qcuMrhsGeom *gc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (g.N > 5) return api_fail("the clover term supports nColors 2..5");
  if (!prec_bytes(prec)) return api_fail("unsupported precision");
  if (nrhs < 1) return api_fail("nrhs must be >= 1");
  if (parity != 0 && parity != 1) return api_fail("parity must be 0 or 1");
  if (!out || !in || !clover) return api_fail("null spinor or clover pointer");
  if (out == in) return api_fail("out and in must not alias");
  const hipError_t e = clover_apply(prec, out, in, clover, parity, int(sites_per_parity(g)), nrhs, g.N,
                                    static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return api_fail(std::string("clover apply launch failed: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"

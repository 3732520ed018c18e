// C ABI of the face exchange (qcuMrhsFaceBytes, qcuMrhsPackFaces,
// qcuMrhsDslashFaces): argument checks, then one launch per split dimension.
#include <hip/hip_runtime.h>

#include <string>

#include "api_internal.h"
#include "face_kernels.h"
#include "qcu_mrhs.h"

namespace qmrhs {
namespace {

// Common checks; on success *g holds the geometry.
bool face_args(const qcuMrhsGeom *gc, int prec, int nrhs, int parity, Geom *g, std::string *err)
{
  if (!geom_from_c(gc, g, err)) return false;
  if (!prec_bytes(prec)) return *err = "unsupported precision", false;
  if (nrhs < 1) return *err = "nrhs must be positive", false;
  if (parity != 0 && parity != 1) return *err = "parity must be 0 or 1", false;
  return true;
}

bool aligned(const void *p, int prec) { return uintptr_t(p) % uintptr_t(prec_bytes(prec)) == 0; }

// Both face buffers of every split dimension present and aligned (checked
// before anything is launched).
bool faces_ok(const Geom &g, int prec, const void *const *a, const void *const *b, std::string *err)
{
  for (int mu = 0; mu < 4; ++mu) {
    if (g.grid[mu] <= 1) continue;
    if (!a[mu] || !b[mu]) return *err = "null face buffer for split dimension " + std::to_string(mu), false;
    if (!aligned(a[mu], prec) || !aligned(b[mu], prec)) return *err = "misaligned face buffer", false;
  }
  return true;
}

}  // namespace
}  // namespace qmrhs

using namespace qmrhs;

extern "C" {

size_t qcuMrhsFaceBytes(const qcuMrhsGeom *gc, int nrhs, int prec, int mu)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err) || nrhs < 1 || !prec_bytes(prec) || mu < 0 || mu > 3) return 0;
  return size_t(face_sites(g, mu)) * 2 * g.N * nrhs * prec_bytes(prec);
}

int qcuMrhsPackFaces(void *const *to_bwd, void *const *to_fwd, const void *in, const void *gauge, int prec, int nrhs,
                     int parity, int dagger, const qcuMrhsGeom *gc, void *stream)
{
  Geom g;
  std::string err;
  if (!face_args(gc, prec, nrhs, parity, &g, &err)) return api_fail(err);
  if (!to_bwd || !to_fwd || !in || !gauge) return api_fail("null face array, spinor or gauge");
  if (!aligned(in, prec) || !aligned(gauge, prec)) return api_fail("misaligned spinor or gauge");
  if (!faces_ok(g, prec, to_bwd, to_fwd, &err)) return api_fail(err);
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int mu = 0; mu < 4; ++mu) {
    if (g.grid[mu] <= 1) continue;
    const hipError_t e = launch_pack_faces(to_bwd[mu], to_fwd[mu], in, gauge, prec, nrhs, parity, dagger, g, mu, st);
    if (e != hipSuccess) return api_fail(std::string("face pack launch failed: ") + hipGetErrorString(e));
  }
  return 0;
}

int qcuMrhsDslashFaces(void *out, const void *const *from_bwd, const void *const *from_fwd, const void *gauge,
                       int prec, int nrhs, int parity, int dagger, const qcuMrhsGeom *gc, void *stream)
{
  Geom g;
  std::string err;
  if (!face_args(gc, prec, nrhs, parity, &g, &err)) return api_fail(err);
  if (!out || !from_bwd || !from_fwd || !gauge) return api_fail("null face array, spinor or gauge");
  if (!aligned(out, prec) || !aligned(gauge, prec)) return api_fail("misaligned spinor or gauge");
  if (!faces_ok(g, prec, from_bwd, from_fwd, &err)) return api_fail(err);
  hipStream_t st = static_cast<hipStream_t>(stream);
  // One launch per dimension, in stream order: a site on an edge takes terms
  // from two dimensions, never from two lanes of one launch.
  for (int mu = 0; mu < 4; ++mu) {
    if (g.grid[mu] <= 1) continue;
    const hipError_t e =
        launch_dslash_faces(out, from_bwd[mu], from_fwd[mu], gauge, prec, nrhs, parity, dagger, g, mu, st);
    if (e != hipSuccess) return api_fail(std::string("face launch failed: ") + hipGetErrorString(e));
  }
  return 0;
}

}  // extern "C"

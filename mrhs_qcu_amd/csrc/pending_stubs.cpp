// Entry points whose implementation lands with the solver and batching units.
#include "api_internal.h"
#include "qcu.h"
#include "qcu_mrhs.h"

extern "C" {
int qcuMrhsCgSolve(void *, const void *, const void *, int, const qcuMrhsGeom *, qcuMrhsCgParam *, double *, void *)
{
  return qmrhs::api_fail("solver not built");
}
int qcuMrhsBatchCreate(const qcuMrhsProblem *, int, void **, void *) { return qmrhs::api_fail("batching not built"); }
int qcuMrhsBatchRun(void *, void *) { return qmrhs::api_fail("batching not built"); }
int qcuMrhsBatchDestroy(void *) { return qmrhs::api_fail("batching not built"); }
void cg_inverter(void *, void *, void *, QcuParam *, double, double) {}
}

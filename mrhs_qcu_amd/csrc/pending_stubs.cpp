// Entry points whose implementation lands with the solver and batching units.
#include "api_internal.h"
#include "qcu.h"
#include "qcu_mrhs.h"

extern "C" {
int qcuMrhsBatchCreate(const qcuMrhsProblem *, int, void **, void *) { return qmrhs::api_fail("batching not built"); }
int qcuMrhsBatchRun(void *, void *) { return qmrhs::api_fail("batching not built"); }
int qcuMrhsBatchDestroy(void *) { return qmrhs::api_fail("batching not built"); }
}

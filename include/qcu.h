/* Public C ABI of the MRHS Wilson dslash library.
 *
 * The declarations are identical to the reference's include/qcu.h:1-34 so that a
 * caller built against the reference links against this library unchanged.
 * All void* field arguments are HIP device pointers; QcuParam / QcuGrid_t are
 * host structs copied at initGridSize.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct QcuParam_s {
  int lattice_size[4];
} QcuParam;

typedef struct QcuGrid_t {
  int grid_size[4];
} QcuGrid_t;

typedef QcuGrid_t QcuGrid;

/* Creates (or re-creates) the process-wide context for an Lx*Ly*Lz*Lt local
 * lattice with nColors colours (reference src/qcu.cu:127-135). */
void initGridSize(QcuGrid_t *grid, QcuParam *p_param, int nColors);
/* Queues one right-hand side; the input is copied at push time, the output
 * pointer is written by the next startDslash (reference src/qcu.cu:88-101). */
void pushBackFermions(void *fermionOut, void *fermionIn);
/* Applies the hopping term to every queued right-hand side in one launch and
 * writes the queued outputs (reference src/qcu.cu:58-86). */
void startDslash(int parity, int daggerFlag);
/* Copies the AoS gauge field [4][2][V/2][N][N] complex (reference src/qcu.cu:102-109). */
void loadQcuGauge(void *gauge);
/* Builds the Dirac operator (reference src/qcu.cu:42-56, 152-160). */
void getDslash(int dslashType, double mass, int nColors, int nInputs, int floatPrecision);

/* Deprecated in the reference (src/qcu.cu:170-171): no-ops. */
void dslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity);
void fullDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int dagger_flag);
/* Stub in the reference (src/qcu.cu:173). Implemented here: solves the full
 * Wilson system (1 - kappa*D) x = b by even-odd preconditioned CG. */
void cg_inverter(void *x_vector, void *b_vector, void *gauge, QcuParam *param, double p_max_prec, double p_kappa);

#ifdef __cplusplus
}
#endif

/* Extended C ABI of the MI355X MRHS Wilson dslash library.
 *
 * The reference exposes only the queue-style entry points of qcu.h; these
 * entry points expose the same operator on explicit, persistent device
 * buffers so that a caller (an FFI binding, a solver, a benchmark) can keep
 * fields resident in HBM in the native layout and avoid any per-call copy.
 *
 * Conventions
 *  - Every field pointer is a HIP device pointer; `stream` is a hipStream_t
 *    (NULL = the default stream). Calls are asynchronous on `stream` unless
 *    stated otherwise.
 *  - Return value 0 = success; nonzero = error, message via qcuMrhsLastError().
 *  - Precision codes: QCU_MRHS_HALF (IEEE binary16 storage, FP32 arithmetic),
 *    QCU_MRHS_SINGLE, QCU_MRHS_DOUBLE, QCU_MRHS_BF16 (bfloat16 storage, FP32
 *    arithmetic), QCU_MRHS_FP8 (OCP FP8 E4M3 spinors, 2 bytes per complex,
 *    saturating at +-448, with binary16 links: the gauge of an FP8 operator is
 *    packed and sized as QCU_MRHS_HALF; FP32 arithmetic; nColors = 3). They
 *    extend QCU_PRECISION of qcu_enum.h.
 *  - Layouts (complex numbers interleaved re, im; S = Lx*Ly*Lz*Lt/2 sites of
 *    one parity, checkerboard index idx = ((t*Lz+z)*Ly+y)*(Lx/2)+x/2):
 *      user spinor   [S][4][N]            (one buffer per RHS; the qcu.h layout)
 *      user gauge    [4][2][S][N][N]      (the qcu.h layout)
 *      packed spinor [S][4][N][nrhs]      (native; RHS innermost)
 *      native gauge  [2][S][8][N][N]      (per site: U_mu(x) for mu = 0..3, then
 *                                          U_mu(x - mu)^dagger for mu = 0..3)
 *  - The operator is the bare hopping term of the reference
 *    (wilson_dslash_4d.cuh:59-293):
 *      out(x) = sum_mu U_mu(x)(1 - s g_mu) in(x+mu) + U_mu(x-mu)^+ (1 + s g_mu) in(x-mu)
 *    for x of the output parity, s = +1 (dagger 0) or -1 (dagger 1).
 */
#pragma once

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QCU_MRHS_HALF = 0, QCU_MRHS_SINGLE = 1, QCU_MRHS_DOUBLE = 2, QCU_MRHS_BF16 = 4, QCU_MRHS_FP8 = 5 };

typedef struct qcuMrhsGeom_s {
  int L[4];    /* local lattice Lx, Ly, Lz, Lt (all even) */
  int grid[4]; /* process grid; a dimension with grid > 1 skips the hops that leave the local volume */
  int nColors; /* N: 2..8 */
} qcuMrhsGeom;

typedef struct qcuMrhsTile_s {
  int rhs_per_lane; /* 1, 2 or 4 right-hand sides per lane */
  int block;        /* lanes per workgroup: 64, 128, 192 or 256 */
  int link_lds;     /* 1: stage each workgroup's links in LDS */
  int t_block;      /* sites traversed in blocks of t_block time slices ... */
  int z_block;      /* ... and z_block z slices; (1, 0) = natural order, 0 = Lz */
  int cache_hints;  /* bit 0: non-temporal output stores; bit 1: non-temporal link loads;
                       bit 2: keep the hardware workgroup order (no XCD-contiguous remap);
                       bit 3 (y-streamed traversals only): XCD-cyclic order, the 8 XCDs work side by side
                       on adjacent halves of y-streamed blocks (ignored when the lattice does
                       not split into whole chunks of workgroups);
                       bit 4 (user AoS layout): each lane reads and writes its own RHS's
                       records instead of the site's lanes reading them together */
  int occupancy;    /* cap on resident workgroups per CU (3..8), 0 = no cap */
  int traversal;    /* bit 0 = 0: each (t_block, z_block) block in (t, z, y, x) order;
                       bit 0 = 1: the block streamed along y, (y, t, z, x) order inside it;
                       bit 1: hypercube mapping, one wavefront per 2x2x2x2 block of output
                       sites (neighbours inside the block shared across lanes); needs
                       P = nrhs / rhs_per_lane in {1, 2, 4, ..., 32}, block a multiple
                       of 8 P, link_lds = 1, nColors <= 4,
                       and t_block / z_block then count blocks (divisors of Lt/2, Lz/2);
                       bit 2 (bit 0 clear): column order, the lattice cut into columns of
                       t_block x z_block in (y, z) that XCD x walks slab by slab along t
                       (columns 8 g + x), each column a whole number of workgroups;
                       traversal 4: site-major kernels, columns in sites (divisors of Ly,
                       Lz); traversal 6: the hypercube mapping, columns in blocks (divisors
                       of Ly/2, Lz/2), run on calls without a hypercube kernel as the same
                       columns in sites, or where those do not fill whole workgroups as
                       y-streamed site-major blocks of one t slice and z_block z slices;
                       bit 3 (with bit 2 only): the columns also split x in two halves;
                       higher bits are not defined (a tile that sets them is rejected) */
} qcuMrhsTile;

const char *qcuMrhsLastError(void);
int qcuMrhsVersion(void);

size_t qcuMrhsSpinorBytes(const qcuMrhsGeom *g, int nrhs, int prec);
size_t qcuMrhsGaugeBytes(const qcuMrhsGeom *g, int prec);

/* ---- storage layer ---------------------------------------------------- */
/* user gauge (FP64 or FP32) -> native gauge in dst_prec. */
int qcuMrhsPackGauge(void *dst, int dst_prec, const void *src_user, int src_prec, const qcuMrhsGeom *g, void *stream);
/* nrhs user spinors (host array of device pointers) -> packed spinor. */
int qcuMrhsPackSpinors(void *dst_packed, int dst_prec, const void *const *src_user, int src_prec, int nrhs,
                       const qcuMrhsGeom *g, void *stream);
/* packed spinor -> nrhs user spinors. */
int qcuMrhsUnpackSpinors(void *const *dst_user, int dst_prec, const void *src_packed, int src_prec, int nrhs,
                         const qcuMrhsGeom *g, void *stream);
/* n complex numbers converted element-wise between precisions. */
int qcuMrhsConvert(void *dst, int dst_prec, const void *src, int src_prec, size_t n, void *stream);
/* Reference shift semantics for any N and element width (bytes per complex:
 * 4, 8 or 16): direction 0 = user -> reference SoA, 1 = SoA -> user. */
int qcuMrhsShiftSpinor(void *dst, const void *src, int direction, const qcuMrhsGeom *g, int bytes_per_complex,
                       void *stream);
int qcuMrhsShiftGauge(void *dst, const void *src, int direction, const qcuMrhsGeom *g, int bytes_per_complex,
                      void *stream);

/* ---- operator ----------------------------------------------------------- */
/* out = D in on packed fields; gauge in the native layout; all three in prec.
 * tile = NULL: the tuned (or default) tile for this shape. */
int qcuMrhsDslash(void *out, const void *in, const void *gauge, int prec, int nrhs, int parity, int dagger,
                  const qcuMrhsGeom *g, const qcuMrhsTile *tile, void *stream);
/* out = a * x + b * D in (x packed, same precision and parity as out). */
int qcuMrhsDslashXpay(void *out, const void *in, const void *x, double a, double b, const void *gauge, int prec,
                      int nrhs, int parity, int dagger, const qcuMrhsGeom *g, const qcuMrhsTile *tile, void *stream);
/* General form used by solvers: out = a * x + b * D in with separate storage
 * precisions for in (and x, and the gauge) and for out, and optional per-RHS
 * scales (device float arrays of nrhs, NULL = 1): the true input is
 * stored_in * sc_in[r], the true x is stored_x * sc_x[r], and out is stored
 * as true_out / sc_out[r]. norm2_out (device double array of nrhs, may be
 * NULL) receives |true out|^2 per RHS. x may be NULL (then out = b * D in). */
int qcuMrhsDslashEx(void *out, int out_prec, const void *in, const void *x, int in_prec, double a, double b,
                    const float *sc_in, const float *sc_x, const float *sc_out, double *norm2_out, const void *gauge,
                    int nrhs, int parity, int dagger, const qcuMrhsGeom *g, void *stream);
/* No counterpart in the reference. qcuMrhsDslashEx without any device
 * allocation, for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation): the same
 * arguments and kernels, with the per-RHS norm partials in a fixed 16 MiB
 * device array the library holds on each GPU until the process ends. Calls
 * whose norms need more (workgroups + 256) * nrhs * 8 bytes than that fail
 * with a message. Asynchronous on `stream`. Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (other streams, other host
 * threads) race; calls in one stream are ordered. */
int qcuMrhsDslashExNoAlloc(void *out, int out_prec, const void *in, const void *x, int in_prec, double a, double b,
                           const float *sc_in, const float *sc_x, const float *sc_out, double *norm2_out,
                           const void *gauge, int nrhs, int parity, int dagger, const qcuMrhsGeom *g, void *stream);
/* Hopping term on nrhs user spinors (host arrays of device pointers), native
 * gauge in prec; the qcu.h startDslash path without the queue. */
int qcuMrhsDslashUser(void *const *out_user, const void *const *in_user, const void *gauge, int prec, int nrhs,
                      int parity, int dagger, const qcuMrhsGeom *g, void *stream);
/* Hopping term through qcu::instantiateDslash: spinors in the reference SoA
 * layout (host arrays of device pointers), gauge in the reference SoA layout;
 * prec is QCU_MRHS_HALF, _SINGLE or _DOUBLE. Synchronous. */
int qcuMrhsDslashReferenceLayout(void *const *out_soa, void *const *in_soa, void *gauge_soa, int prec, int nrhs,
                                 int parity, int dagger, const qcuMrhsGeom *g);

/* ---- two-row SU(3) links (12 reals per link) -------------------------------
 * Gauge compression of the native layout for nColors = 3: each link is kept
 * as its first two rows, [2][S][8][2][3] in prec, and the kernel rebuilds
 * row 2 = conj(row 0 x row 1), which holds for SU(3). Link bytes drop by a
 * third. No counterpart in the reference, which loads full links
 * (include/kernels/storageLoadStore/loadAndStore.cuh:6-15; its
 * loadGauge2 at :17-27 is unused).
 * qcuMrhsPackGauge12 takes the user gauge [4][2][S][3][3] (FP64 or FP32),
 * checks every link (after dividing the time links that cross the t boundary
 * by t_sign: -1 when antiperiodic time is folded into the links, else +1) is
 * SU(3) to 1e-5 and fails otherwise; it synchronises the stream.
 * qcuMrhsDslash12 is qcuMrhsDslash on that layout, with the same t_sign; the
 * tile (NULL: the one tuned or chosen for the native layout) follows
 * qcuMrhsDslash. */
size_t qcuMrhsGauge12Bytes(const qcuMrhsGeom *g, int prec);
int qcuMrhsPackGauge12(void *dst, int dst_prec, const void *src_user, int src_prec, const qcuMrhsGeom *g, int t_sign,
                       void *stream);
int qcuMrhsDslash12(void *out, const void *in, const void *gauge12, int prec, int nrhs, int parity, int dagger,
                    int t_sign, const qcuMrhsGeom *g, const qcuMrhsTile *tile, void *stream);

/* ---- eight-real SU(3) links -------------------------------------------------
 * Each link as U01, U02, U10 and (arg U00, arg U20): [2][S][8][4] complex
 * slots in prec, 8 of 18 reals. The moduli of U00 and U20 follow from unit
 * rows and columns, the other four entries from unitarity and det = 1; the
 * rebuild divides by |U01|^2 + |U02|^2 and runs once per site while the
 * links are staged in LDS (the tile must have link_lds = 1). No counterpart
 * in the reference, which loads full links.
 * qcuMrhsPackGauge8 checks every link as qcuMrhsPackGauge12 does and also
 * refuses links with |U01|^2 + |U02|^2 < 1e-6 (a unit gauge, for example:
 * storage errors would grow by up to 1000x); it synchronises the stream.
 * qcuMrhsDslash8 is qcuMrhsDslash on that layout, with the same t_sign; tile
 * NULL takes the native layout's tile with link_lds = 1. */
size_t qcuMrhsGauge8Bytes(const qcuMrhsGeom *g, int prec);
int qcuMrhsPackGauge8(void *dst, int dst_prec, const void *src_user, int src_prec, const qcuMrhsGeom *g, int t_sign,
                      void *stream);
int qcuMrhsDslash8(void *out, const void *in, const void *gauge8, int prec, int nrhs, int parity, int dagger,
                   int t_sign, const qcuMrhsGeom *g, const qcuMrhsTile *tile, void *stream);

/* ---- clover term -----------------------------------------------------------
 * The operator DSLASH_CLOVER names (include/qcu_enum.h:12 of the reference,
 * which reserves 72 complex numbers per site for it,
 * include/basic_data/qcu_point.cuh:98-101, and implements nothing more):
 *   A(x) = 1 + coeff * sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
 *   sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu] (DeGrand-Rossi, the basis of the
 *   hopping term), F_{mu nu} = (Q - Q^+)/(8i) with Q the sum of the four
 *   plaquettes of the (mu, nu) plane around x (no trace subtraction).
 * A is two Hermitian (2N x 2N) blocks per site (spins {0,1} and {2,3});
 * field layout [2 parity][S][2][2N][2N] complex in prec, 8 N^2 complex per
 * site. For the Wilson-clover operator M = A - kappa D, coeff = -kappa c_sw / 2
 * in the common normalisation.
 * qcuMrhsCloverCompute builds A (and, with clover_inv non-NULL, A^-1 blockwise)
 * from the user gauge [4][2][S][N][N] (FP64 or FP32), arithmetic in FP64;
 * synchronous. qcuMrhsCloverApply: out = C in on the sites of `parity`, C the
 * given field (A or A^-1), packed spinors in prec. nColors 2..5. */
size_t qcuMrhsCloverBytes(const qcuMrhsGeom *g, int prec);
int qcuMrhsCloverCompute(void *clover, void *clover_inv, int prec, const void *gauge_user, int gauge_prec,
                         double coeff, const qcuMrhsGeom *g, void *stream);
int qcuMrhsCloverApply(void *out, const void *in, const void *clover, int prec, int nrhs, int parity,
                       const qcuMrhsGeom *g, void *stream);

/* ---- face exchange (one dimension split over processes) ------------------
 * No counterpart in the reference, whose multi-process layer is declared but
 * not implemented (include/qcuDesc/qcu_desc.h:53-97, boundary streams at
 * include/qcu_kernel_param.cuh:11). With g->grid[mu] > 1, qcuMrhsDslash skips
 * the terms that cross the local boundary of mu (the reference's semantics,
 * wilson_dslash_4d.cuh:60,87,...); these two calls supply them:
 *   qcuMrhsPackFaces  (on `in`, parity 1 - parity): to_bwd[mu] = projected
 *     half spinors of the x_mu = 0 face, for the process behind in mu;
 *     to_fwd[mu] = U^+ times the projected half spinors of the x_mu = L_mu - 1
 *     face, for the process ahead;
 *   qcuMrhsDslashFaces (on `out`, parity `parity`): adds the boundary terms,
 *     from_fwd[mu] = the to_bwd[mu] of the process ahead, from_bwd[mu] = the
 *     to_fwd[mu] of the process behind.
 * Arrays of 4 device pointers; entries of unsplit dimensions are ignored.
 * Face buffers hold qcuMrhsFaceBytes(g, nrhs, prec, mu) bytes, [F][2][N][nrhs]
 * in prec (the precision of the spinors and the native gauge). Moving the
 * buffers between processes (RCCL send/recv, peer copies) is the caller's. */
size_t qcuMrhsFaceBytes(const qcuMrhsGeom *g, int nrhs, int prec, int mu);
int qcuMrhsPackFaces(void *const *to_bwd, void *const *to_fwd, const void *in, const void *gauge, int prec, int nrhs,
                     int parity, int dagger, const qcuMrhsGeom *g, void *stream);
int qcuMrhsDslashFaces(void *out, const void *const *from_bwd, const void *const *from_fwd, const void *gauge,
                       int prec, int nrhs, int parity, int dagger, const qcuMrhsGeom *g, void *stream);

/* ---- tuning ------------------------------------------------------------- */
/* Times every instantiated tile on the given fields (out is overwritten),
 * records the fastest for (N, nrhs, prec, lattice) and returns it. With
 * QCU_MRHS_TUNE_CACHE=<file> in the environment, results persist in that text
 * file (keyed by GPU architecture) and are reused by later processes. With
 * QCU_MRHS_TUNE_LOG=1 every candidate tile and its time go to stderr.
 * Shapes listed in the pinned tile table (tiles_<arch>.txt beside the
 * library, or the file named by QCU_MRHS_TILES; same line format) count as
 * tuned: the dispatcher and the solver use those tiles without tuning. */
int qcuMrhsTune(void *out, const void *in, const void *gauge, int prec, int nrhs, const qcuMrhsGeom *g,
                qcuMrhsTile *best, double *best_us, void *stream);
/* Tile the dispatcher uses for a shape (tuned if known, else the default). */
int qcuMrhsGetTile(const qcuMrhsGeom *g, int prec, int nrhs, qcuMrhsTile *tile);
/* Sets the tile the dispatcher, qcuMrhsDslashEx and the solver use for a
 * shape in this process (also appended to the QCU_MRHS_TUNE_CACHE file when
 * that is set). The tile is checked when it is next launched.
 * GetTile / SetTile with prec | QCU_MRHS_TILE_USER_LAYOUT name the tile of
 * the user-layout path (qcu.h startDslash, qcuMrhsDslashUser); only its
 * hypercube tiles (traversal bit 1, block a multiple of 8 nrhs) change that
 * path, which otherwise keeps its built-in traversal. */
#define QCU_MRHS_TILE_USER_LAYOUT 8
int qcuMrhsSetTile(const qcuMrhsGeom *g, int prec, int nrhs, const qcuMrhsTile *tile);
/* Number of instantiated dslash kernels (for diagnostics). */
int qcuMrhsKernelCount(void);

/* ---- measurement -------------------------------------------------------- */
/* Streaming bandwidth of the GPU, the ceiling the dslash's roofline fraction
 * is read against: mode 0 reads `bytes` of src, 1 writes `bytes` of dst,
 * 2 copies src to dst, 3 reads and rewrites dst in place, 4 copies the 8-byte
 * elements of src to dst at a 128-byte stride (one FP32 right-hand side into
 * its slot of a 16-RHS packed field; dst must hold 16 x bytes); 16-byte
 * accesses over every CU, one warm-up pass, then `reps` timed passes. *gbs
 * receives the useful bytes moved (read + written) per second in GB/s.
 * Synchronous. */
int qcuMrhsStreamBandwidth(int mode, void *dst, const void *src, size_t bytes, int reps, double *gbs, void *stream);

/* Link x half-spinor block products of the SU(N) hop, Y (N x 2k) = U (N x N)
 * H (N x 2k) complex, timed on the vector ALUs (the production kernels'
 * form) and on the matrix cores (v_mfma_f32_16x16x4_f32 on the real
 * embedding [[Ur, -Ui], [Ui, Ur]]); operands reloaded per block from
 * cache-resident tables (resident = 0: products plus operand loads) or held in
 * registers (resident = 1: products alone). No counterpart in the reference:
 * it backs the decision on an MFMA path for N >= 4. tflops[0] / tflops[1] receive the
 * useful FP32 rate (8 flops per complex multiply-add) of the VALU / MFMA form,
 * max_rel_err[0] / [1] the error of each form against an FP64 host product
 * over 1024 blocks. 2 <= N <= 8, 1 <= nrhs <= 32. Synchronous. */
int qcuMrhsLinkBlockBench(int N, int nrhs, int reps, int resident, double *tflops, double *max_rel_err,
                          void *stream);

/* ---- solver ------------------------------------------------------------- */
typedef struct qcuMrhsCgParam_s {
  double kappa;          /* hopping parameter; M^ = 1 - kappa^2 D_eo D_oe */
  double tol;            /* target ||M^+b - M^+M^ x|| / ||M^+b|| per RHS */
  int max_iter;          /* iteration cap */
  int reliable_interval; /* mixed precision: iterations between FP64 residual updates */
  int sloppy_prec;       /* QCU_MRHS_DOUBLE: FP64 CG; HALF / BF16 / SINGLE / FP8: mixed precision */
  double reliable_delta; /* mixed precision: true residual recomputed once the iterated residual of a
                            RHS fell by this factor since its last update (0 = precision default:
                            HALF 3e-3, SINGLE 1e-3, BF16 1e-1, FP8 3e-1) */
  /* outputs */
  int iterations;
  int reliable_updates;
  double seconds;
} qcuMrhsCgParam;

/* Solves M^+ M^ x = M^+ b on the even checkerboard for nrhs right-hand sides.
 * x, b: packed FP64 even-parity spinors; gauge: native FP64 gauge.
 * true_res (host, nrhs doubles, may be NULL) receives the final true relative
 * residual of each RHS recomputed in FP64. Synchronous. */
int qcuMrhsCgSolve(void *x, const void *b, const void *gauge_f64, int nrhs, const qcuMrhsGeom *g, qcuMrhsCgParam *p,
                   double *true_res, void *stream);
/* The solver keeps its work vectors between calls (no allocation when the
 * shape repeats) and tunes its operator tiles on first use; this releases
 * the work vectors. */
/* The same solver for the Wilson-clover operator: on the even checkerboard
 * M^ = A_ee - kappa^2 D_eo A_oo^-1 D_oe (A, A^-1 from qcuMrhsCloverCompute in
 * FP64; the hopping normalisation of M = A - kappa D). Solves M^+M^ x = M^+ b
 * with the same parameters; with a 16-bit sloppy precision A and A^-1 are
 * also kept in it. nColors 2..5. */
int qcuMrhsCgSolveClover(void *x, const void *b, const void *gauge_f64, const void *clover_f64,
                         const void *clover_inv_f64, int nrhs, const qcuMrhsGeom *g, qcuMrhsCgParam *p,
                         double *true_res, void *stream);
int qcuMrhsSolverRelease(void);

/* ---- batching ------------------------------------------------------------ */
typedef struct qcuMrhsProblem_s {
  qcuMrhsGeom geom;
  int nrhs, parity, dagger;
  void *out;         /* packed FP32 */
  const void *in;    /* packed FP32 */
  const void *gauge; /* native FP32 */
} qcuMrhsProblem;

/* Builds one replayable launch (a HIP graph) applying the dslash to every
 * problem, with a tuned tile per problem. With every nColors in 2..5 the
 * graph is a single work-list kernel covering all problems (per-N tile
 * family: 2 RHS per lane with links staged in LDS for N <= 4, 1 RHS per lane
 * for N = 5; 128 lanes per workgroup; each problem's tuned traversal);
 * otherwise one kernel per problem. */
int qcuMrhsBatchCreate(const qcuMrhsProblem *probs, int nprob, void **handle, void *stream);
int qcuMrhsBatchRun(void *handle, void *stream);
/* The tile the batch runs problem i with (the caller's problem order). */
int qcuMrhsBatchGetTile(void *handle, int i, qcuMrhsTile *tile);
int qcuMrhsBatchDestroy(void *handle);

#ifdef __cplusplus
}
#endif

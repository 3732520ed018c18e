// Vector kernels and device scalar state of the multi-RHS CG (host entry
// points; all asynchronous on `st`).
#pragma once
#include <hip/hip_runtime.h>

namespace qmrhs {

constexpr int kBlasBlock = 256;
constexpr int kBlasMaxRhs = 256;

// Per-RHS scalars, all device arrays of length k.
struct CgState {
  double *rr, *rr_new, *pap, *alpha, *beta, *pnorm2, *pnorm, *b2, *rdotp, *gamma;
  int *done, *sloppy_conv;
  float *sc_p, *sc_t, *sc_q, *sc_t2;  // sloppy storage scales (true = stored * sc)
  float *sc_p_old;                    // scale of the stored p before cg_beta replaced sc_p
  float scale_div = 1024.0f;          // sc_p = (bound of |p|) / scale_div; t, q, t2 16x, 4x, 64x that
};

// Workgroups of every vector kernel for n complex numbers and k RHS (the
// partial-sum buffers hold blas_num_blocks(n, k) * k doubles).
long long blas_num_blocks(long long n, int k);

// x += alpha p (x = NULL: skipped) ; p = gamma r + beta p (gamma = NULL: 1) ;
// ps = p / sc_p ; partial |p|^2. p = NULL: the direction is kept only in ps,
// the old one read as ps * sc_old (sc_old = NULL: 1). The old direction is
// read from p_src / ps_src, the new one written to p / ps (out of place, or
// the same buffer).
hipError_t blas_update_p(int prec_k, int prec_s, void *p, const void *p_src, void *ps, const void *ps_src,
                         const void *r, const double *beta, const double *gamma, void *x, const double *alpha,
                         const float *sc_p, const float *sc_old, double *partial, long long n, int k, hipStream_t st);
// Search directions of CG steps whose x update is still to be applied: the
// stored direction of step j (precision prec_p of blas_flush), its per-RHS
// step length alpha_j and storage scale sc_j (true p = stored p * sc_j; null: 1).
constexpr int kMaxDirSteps = 8;
struct DirSteps {
  int count = 0;
  const void *p[kMaxDirSteps] = {};
  const double *alpha[kMaxDirSteps] = {};
  const float *sc[kMaxDirSteps] = {};
};
// xh (FP64) += sum over the steps of alpha_j * sc_j * p_j, in step order.
hipError_t blas_flush(int prec_p, void *xh, const DirSteps &steps, long long n, int k, hipStream_t st);
hipError_t blas_norm2(int prec, const void *v, double *partial, long long n, int k, hipStream_t st);
// r = b - Ax (b FP64, Ax FP64 or, with FP32 r, FP32; difference in FP64, r
// stored in prec_k; Ax may be r itself), partials |r|^2 and Re <r, p> with
// the true p = stored p (prec_p) * sc_p (NULL: 1).
hipError_t blas_residual(int prec_k, void *r, const void *b, const void *Ax, int prec_ax, const void *p, int prec_p,
                         const float *sc_p, double *partial_rr, double *partial_rp, long long n, int k,
                         hipStream_t st);
hipError_t blas_zero(int prec, void *v, long long n, int k, hipStream_t st);
// out[r] = sum over nb workgroups of partial[b][r]; stage: device scratch of
// kReduceStages * k doubles.
constexpr int kReduceStages = 256;
hipError_t blas_reduce(const double *partial, long long nb, int k, double *out, double *stage, hipStream_t st);
hipError_t cg_alpha(const CgState &s, int k, hipStream_t st);
hipError_t cg_beta(const CgState &s, int k, double stop2, hipStream_t st);
hipError_t cg_orth(const CgState &s, int k, hipStream_t st);

}  // namespace qmrhs

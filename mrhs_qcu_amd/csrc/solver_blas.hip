// Vector kernels of the multi-RHS CG on packed fields [S][4][N][k].
//
// Every kernel is grid-stride with a thread count that is a multiple of k, so
// each lane keeps one right-hand side for its whole walk and accumulates its
// per-RHS sums in a register; a workgroup folds them into LDS and writes one
// partial per RHS: [nblocks][k], summed by blas_reduce (two passes). No global atomics,
// so every reduction is deterministic.
#include <hip/hip_runtime.h>

#include "qcu_device.h"
#include "solver_blas.h"

namespace qmrhs {

namespace {

template <class C> struct Acc;
template <> struct Acc<c64> { using T = double; };
template <> struct Acc<c32> { using T = float; };
template <> struct Acc<c16> { using T = float; };
template <> struct Acc<cb16> { using T = float; };
template <> struct Acc<c8> { using T = float; };

__device__ __forceinline__ void block_partial(double v, int rhs, int k, double *partial)
{
  __shared__ double red[kBlasMaxRhs];
  for (int i = threadIdx.x; i < k; i += blockDim.x) red[i] = 0.0;
  __syncthreads();
  atomicAdd(&red[rhs], v);
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) partial[size_t(blockIdx.x) * k + i] = red[i];
}

// block_partial for V consecutive right-hand sides rhs0 .. rhs0 + V - 1.
template <int V>
__device__ __forceinline__ void block_partial_v(const double (&v)[V], int rhs0, int k, double *partial)
{
  __shared__ double red[kBlasMaxRhs];
  for (int i = threadIdx.x; i < k; i += blockDim.x) red[i] = 0.0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < V; ++j) atomicAdd(&red[rhs0 + j], v[j]);
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) partial[size_t(blockIdx.x) * k + i] = red[i];
}

// V consecutive complex numbers of a packed field, one vector access.
template <class C, int V>
__device__ __forceinline__ CVec<C, V> ldv(const C *p, long long e)
{
  return *reinterpret_cast<const CVec<C, V> *>(p + e);
}
template <class C, int V>
__device__ __forceinline__ void stv(C *p, long long e, const CVec<C, V> &v)
{
  *reinterpret_cast<CVec<C, V> *>(p + e) = v;
}

// x += alpha p (x null: skipped) ; p = gamma r + beta p (gamma null: 1) ;
// p_s = p / sc_p (sloppy copy, may be null) ; partial |p|^2. With p null the
// search direction lives only in p_s: the old p is p_s * sc_old (null: 1).
// The old direction is read from p_src / ps_src and the new one written to p
// / ps (out of place, or the same buffer: each element is read before it is
// written by the same lane). A lane takes V consecutive complex numbers per
// step (V right-hand sides of one site row: k % V == 0), so a 16-bit field
// moves in 16-byte accesses.
template <class CK, class CS, int V>
__global__ void __launch_bounds__(kBlasBlock) update_p_kernel(CK *p, const CK *p_src, CS *ps, const CS *ps_src,
                                                              const CK *__restrict__ r,
                                                              const double *__restrict__ beta,
                                                              const double *__restrict__ gamma,
                                                              CK *__restrict__ x, const double *__restrict__ alpha,
                                                              const float *__restrict__ sc_p,
                                                              const float *__restrict__ sc_old, double *partial,
                                                              long long n, int k)
{
  using A = typename Acc<CK>::T;
  const long long T = (long long)gridDim.x * blockDim.x;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs0 = int((g * V) % k);  // constant along the walk: T is a multiple of k
  A b[V], c[V], al[V], so[V];
  float inv[V];
  double acc[V];
#pragma unroll
  for (int j = 0; j < V; ++j) {
    const int rhs = rhs0 + j;
    b[j] = A(beta[rhs]);
    c[j] = gamma ? A(gamma[rhs]) : A(1);
    al[j] = x ? A(alpha[rhs]) : A(0);
    inv[j] = sc_p ? 1.0f / sc_p[rhs] : 1.0f;
    so[j] = sc_old ? A(sc_old[rhs]) : A(1);
    acc[j] = 0.0;
  }
  const long long ng = n / V;
  for (long long gi = g; gi < ng; gi += T) {
    const long long e = gi * V;
    cpx<A> pp[V];
    if (p_src) {
      const CVec<CK, V> v = ldv<CK, V>(p_src, e);
#pragma unroll
      for (int j = 0; j < V; ++j) pp[j] = ld_cvt<A>(v.v[j]);
    } else {
      const CVec<CS, V> v = ldv<CS, V>(ps_src, e);
#pragma unroll
      for (int j = 0; j < V; ++j) {
        pp[j] = ld_cvt<A>(v.v[j]);
        pp[j].re *= so[j];
        pp[j].im *= so[j];
      }
    }
    const CVec<CK, V> rv = ldv<CK, V>(r, e);
    if (x) {
      CVec<CK, V> xv = ldv<CK, V>(x, e);
#pragma unroll
      for (int j = 0; j < V; ++j) {
        cpx<A> xx = ld_cvt<A>(xv.v[j]);
        xx.re += al[j] * pp[j].re;
        xx.im += al[j] * pp[j].im;
        xv.v[j] = StCvt<CK>::go(xx);
      }
      stv<CK, V>(x, e, xv);
    }
    CVec<CK, V> po;
    CVec<CS, V> pso;
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const cpx<A> rr = ld_cvt<A>(rv.v[j]);
      pp[j].re = c[j] * rr.re + b[j] * pp[j].re;
      pp[j].im = c[j] * rr.im + b[j] * pp[j].im;
      po.v[j] = StCvt<CK>::go(pp[j]);
      pso.v[j] = StCvt<CS>::go(cpx<float>{float(pp[j].re) * inv[j], float(pp[j].im) * inv[j]});
      acc[j] += double(pp[j].re) * pp[j].re + double(pp[j].im) * pp[j].im;
    }
    if (p) stv<CK, V>(p, e, po);
    if (ps) stv<CS, V>(ps, e, pso);
  }
  block_partial_v<V>(acc, rhs0, k, partial);
}

// xh += sum_j alpha_j sc_j p_j over the stored search directions of the steps
// not yet applied (sc_j null: 1), in FP64, in step order; V consecutive
// complex numbers per lane and step, as update_p.
template <class CP, int V>
__global__ void __launch_bounds__(kBlasBlock) flush_kernel(c64 *__restrict__ xh, const DirSteps steps, long long n,
                                                           int k)
{
  const long long T = (long long)gridDim.x * blockDim.x;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs0 = int((g * V) % k);
  double coef[kMaxDirSteps][V];
#pragma unroll
  for (int s = 0; s < kMaxDirSteps; ++s)
#pragma unroll
    for (int j = 0; j < V; ++j)
      coef[s][j] = s < steps.count
                       ? steps.alpha[s][rhs0 + j] * (steps.sc[s] ? double(steps.sc[s][rhs0 + j]) : 1.0)
                       : 0.0;
  const long long ng = n / V;
  for (long long gi = g; gi < ng; gi += T) {
    const long long e = gi * V;
    CVec<c64, V> h = ldv<c64, V>(xh, e);
#pragma unroll
    for (int s = 0; s < kMaxDirSteps; ++s) {
      if (s < steps.count) {
        const CVec<CP, V> pv = ldv<CP, V>(static_cast<const CP *>(steps.p[s]), e);
#pragma unroll
        for (int j = 0; j < V; ++j) {
          const cpx<double> pp = ld_cvt<double>(pv.v[j]);
          h.v[j].x += coef[s][j] * pp.re;
          h.v[j].y += coef[s][j] * pp.im;
        }
      }
    }
    stv<c64, V>(xh, e, h);
  }
}

// partial |v|^2
template <class C>
__global__ void __launch_bounds__(kBlasBlock) norm2_kernel(const C *__restrict__ v, double *partial, long long n,
                                                           int k)
{
  using A = typename Acc<C>::T;
  const long long T = (long long)gridDim.x * blockDim.x;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = int(g % k);
  double acc = 0.0;
  for (long long e = g; e < n; e += T) {
    cpx<A> z = ld_cvt<A>(v[e]);
    acc += double(z.re) * z.re + double(z.im) * z.im;
  }
  block_partial(acc, rhs, k, partial);
}

// r = b - Ax computed in FP64 (Ax stored as CA) and stored in the Krylov
// precision CK; partials |r|^2 (of the FP64 values) and Re <r, p> (of the
// stored r). Ax may be r itself (each element is read before it is written).
template <class CK, class CA, class CP>
__global__ void __launch_bounds__(kBlasBlock) residual_kernel(CK *r, const c64 *__restrict__ b,
                                                              const CA *Ax, const CP *__restrict__ p,
                                                              const float *__restrict__ sc_p, double *partial_rr,
                                                              double *partial_rp, long long n, int k)
{
  using A = typename Acc<CK>::T;
  const long long T = (long long)gridDim.x * blockDim.x;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = int(g % k);
  const double sp = sc_p ? double(sc_p[rhs]) : 1.0;
  double acc = 0.0, dot = 0.0;
  for (long long e = g; e < n; e += T) {
    const c64 bb = b[e];
    const cpx<double> aa = ld_cvt<double>(Ax[e]);
    const cpx<double> rr{bb.x - aa.re, bb.y - aa.im};
    const CK rs = StCvt<CK>::go(rr);
    r[e] = rs;
    acc += rr.re * rr.re + rr.im * rr.im;
    const cpx<A> rv = ld_cvt<A>(rs);
    const cpx<A> pv = ld_cvt<A>(p[e]);
    dot += double(rv.re) * pv.re + double(rv.im) * pv.im;
  }
  dot *= sp;
  block_partial(acc, rhs, k, partial_rr);
  __syncthreads();
  block_partial(dot, rhs, k, partial_rp);
}

// Fill with zeros (a kernel, so the solver loop stays capturable in a graph).
template <class C>
__global__ void __launch_bounds__(kBlasBlock) zero_kernel(C *__restrict__ v, long long n)
{
  const long long T = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += T)
    v[e] = StCvt<C>::go(cpx<float>{0.f, 0.f});
}

// out[r] = sum_b partial[b][r] in two passes. Pass 1: G workgroups walk the
// flat [nb][k] array with coalesced loads; the thread count per workgroup is
// a multiple of k, so every thread keeps one RHS; each workgroup folds its
// threads per RHS in a fixed order into stage[G][k]. Pass 2: one workgroup
// per RHS sums the G stage values. Deterministic, no atomics.
__global__ void __launch_bounds__(256) reduce_pass1(const double *__restrict__ partial, long long n, int k,
                                                    double *__restrict__ stage)
{
  __shared__ double red[256];
  const long long T = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += T) acc += partial[e];
  red[threadIdx.x] = acc;
  __syncthreads();
  if ((int)threadIdx.x < k) {
    double sum = 0.0;
    for (int t = threadIdx.x; t < (int)blockDim.x; t += k) sum += red[t];
    stage[size_t(blockIdx.x) * k + threadIdx.x] = sum;
  }
}

// One workgroup per RHS: the G stage values are loaded in parallel and
// folded by a fixed-order tree.
__global__ void __launch_bounds__(256) reduce_pass2(const double *__restrict__ stage, int G, int k,
                                                    double *__restrict__ out)
{
  __shared__ double red[256];
  const int r = blockIdx.x;
  red[threadIdx.x] = (int)threadIdx.x < G ? stage[size_t(threadIdx.x) * k + r] : 0.0;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[r] = red[0];
}

// alpha = rr / pAp for live RHS; pnorm = sqrt(sum |p|^2).
__global__ void alpha_kernel(CgState s, int k)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const double pap = s.pap[r];
  s.alpha[r] = (s.done[r] || !(pap > 0.0)) ? 0.0 : s.rr[r] / pap;
  s.pnorm[r] = sqrt(s.pnorm2[r]);
}

// beta = rr_new / rr; rr = rr_new; next sloppy scales for p and the
// intermediate fields (bounds that keep binary16 away from overflow).
__global__ void beta_kernel(CgState s, int k, double stop2)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const double rn = s.rr_new[r];
  const bool live = !s.done[r];
  const double beta = (live && s.rr[r] > 0.0) ? rn / s.rr[r] : 0.0;
  s.beta[r] = beta;
  if (live) s.rr[r] = rn;
  s.sloppy_conv[r] = (rn <= stop2 * s.b2[r]) ? 1 : 0;
  const double bound = sqrt(rn) + fabs(beta) * s.pnorm[r];
  float sp = float(bound / double(s.scale_div));
  if (!(sp > 1e-30f)) sp = 1e-30f;
  s.sc_p_old[r] = s.sc_p[r];
  s.sc_p[r] = sp;
  s.sc_t[r] = 16.0f * sp;
  s.sc_q[r] = 4.0f * sp;
  s.sc_t2[r] = 64.0f * sp;
}

// rr = |b^|^2, beta = 0, done only for a zero right-hand side, scales for p = r.
__global__ void init_state_kernel(CgState s, int k)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const double b2 = s.b2[r];
  s.rr[r] = b2;
  s.rr_new[r] = b2;
  s.beta[r] = 0.0;
  s.alpha[r] = 0.0;
  s.pnorm[r] = 0.0;
  s.pnorm2[r] = 0.0;
  s.done[r] = b2 > 0.0 ? 0 : 1;
  s.sloppy_conv[r] = s.done[r];
  float sp = float(sqrt(b2) / double(s.scale_div));
  if (!(sp > 1e-30f)) sp = 1e-30f;
  s.sc_p[r] = sp;
  s.sc_p_old[r] = sp;
  s.sc_t[r] = 16.0f * sp;
  s.sc_q[r] = 4.0f * sp;
  s.sc_t2[r] = 64.0f * sp;
}

// After a reliable update: the true residual norm takes the place of the
// iterated one (the following beta uses it), and the right-hand sides that
// meet the tolerance are marked done.
__global__ void check_true_kernel(CgState s, const double *rr_true, int k, double stop2)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const double rt = rr_true[r];
  if (!s.done[r]) s.rr_new[r] = rt;
  if (rt <= stop2 * s.b2[r]) s.done[r] = 1;
}

// After a reliable update and cg_beta: the old search direction is made
// orthogonal to the new (true) residual before it enters p = r + beta p, so
// p = gamma r + beta p with gamma = 1 - beta <r, p> / |r|^2; the sloppy
// storage scales follow the new bound |gamma| |r| + |beta| |p|.
__global__ void orth_kernel(CgState s, int k)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const double rr = s.rr[r];
  const double gamma = (!s.done[r] && rr > 0.0) ? 1.0 - s.beta[r] * s.rdotp[r] / rr : 1.0;
  s.gamma[r] = gamma;
  const double bound = fabs(gamma) * sqrt(rr) + fabs(s.beta[r]) * s.pnorm[r];
  float sp = float(bound / double(s.scale_div));
  if (!(sp > 1e-30f)) sp = 1e-30f;
  s.sc_p[r] = sp;
  s.sc_t[r] = 16.0f * sp;
  s.sc_q[r] = 4.0f * sp;
  s.sc_t2[r] = 64.0f * sp;
}

// Workgroup count: a multiple of q (the least q with q * kBlasBlock % k == 0,
// so the thread count is a multiple of k), about 8 workgroups per CU.
unsigned grid_for(long long n, int k)
{
  long long q = 1;
  while ((long long)kBlasBlock * q % k) ++q;
  long long nb = (2048 + q - 1) / q * q;
  const long long need = (n + kBlasBlock - 1) / kBlasBlock;
  if (nb > need) nb = (need + q - 1) / q * q;
  if (nb < q) nb = q;
  return unsigned(nb);
}

}  // namespace

long long blas_num_blocks(long long n, int k) { return grid_for(n, k); }

#define QM_GRID(n, k) dim3(grid_for(n, k)), dim3(kBlasBlock), 0, st

hipError_t blas_update_p(int prec_k, int prec_s, void *p, const void *p_src, void *ps, const void *ps_src,
                         const void *r, const double *beta, const double *gamma, void *x, const double *alpha,
                         const float *sc_p, const float *sc_old, double *partial, long long n, int k, hipStream_t st)
{
#define QM_UP_V(CK, CS, V)                                                                                          \
  hipLaunchKernelGGL((update_p_kernel<CK, CS, V>), QM_GRID(n, k), static_cast<CK *>(p),                           \
                     static_cast<const CK *>(p_src), static_cast<CS *>(ps), static_cast<const CS *>(ps_src),          \
                     static_cast<const CK *>(r), beta, gamma, static_cast<CK *>(x), alpha, sc_p, sc_old, partial, n, k)
#define QM_UP(CK, CS)          \
  if (k % 4 == 0) {            \
    QM_UP_V(CK, CS, 4);        \
  } else {                     \
    QM_UP_V(CK, CS, 1);        \
  }
  if (prec_k == PREC_DOUBLE) {
    QM_UP(c64, c64);
  } else {
    switch (prec_s) {
      case PREC_HALF: QM_UP(c32, c16); break;
      case PREC_BF16: QM_UP(c32, cb16); break;
      case PREC_FP8: QM_UP(c32, c8); break;
      default: QM_UP(c32, c32);
    }
  }
#undef QM_UP
#undef QM_UP_V
  return hipGetLastError();
}

hipError_t blas_norm2(int prec, const void *v, double *partial, long long n, int k, hipStream_t st)
{
  switch (prec) {
    case PREC_DOUBLE:
      hipLaunchKernelGGL(norm2_kernel<c64>, QM_GRID(n, k), static_cast<const c64 *>(v), partial, n, k);
      break;
    case PREC_SINGLE:
      hipLaunchKernelGGL(norm2_kernel<c32>, QM_GRID(n, k), static_cast<const c32 *>(v), partial, n, k);
      break;
    case PREC_HALF:
      hipLaunchKernelGGL(norm2_kernel<c16>, QM_GRID(n, k), static_cast<const c16 *>(v), partial, n, k);
      break;
    case PREC_FP8:
      hipLaunchKernelGGL(norm2_kernel<c8>, QM_GRID(n, k), static_cast<const c8 *>(v), partial, n, k);
      break;
    default:
      hipLaunchKernelGGL(norm2_kernel<cb16>, QM_GRID(n, k), static_cast<const cb16 *>(v), partial, n, k);
  }
  return hipGetLastError();
}

hipError_t blas_residual(int prec_k, void *r, const void *b, const void *Ax, int prec_ax, const void *p, int prec_p,
                         const float *sc_p, double *partial_rr, double *partial_rp, long long n, int k,
                         hipStream_t st)
{
#define QM_RES(CK, CA, CP)                                                                                       \
  hipLaunchKernelGGL((residual_kernel<CK, CA, CP>), QM_GRID(n, k), static_cast<CK *>(r),                          \
                     static_cast<const c64 *>(b), static_cast<const CA *>(Ax), static_cast<const CP *>(p), sc_p, \
                     partial_rr, partial_rp, n, k)
#define QM_RES_P(CK, CA)                         \
  switch (prec_p) {                              \
    case PREC_DOUBLE: QM_RES(CK, CA, c64); break; \
    case PREC_SINGLE: QM_RES(CK, CA, c32); break; \
    case PREC_HALF: QM_RES(CK, CA, c16); break;   \
    case PREC_FP8: QM_RES(CK, CA, c8); break;     \
    default: QM_RES(CK, CA, cb16);                \
  }
  if (prec_k == PREC_DOUBLE) {
    QM_RES_P(c64, c64)
  } else if (prec_ax == PREC_SINGLE) {
    QM_RES_P(c32, c32)
  } else {
    QM_RES_P(c32, c64)
  }
#undef QM_RES_P
#undef QM_RES
  return hipGetLastError();
}

hipError_t blas_flush(int prec_p, void *xh, const DirSteps &steps, long long n, int k, hipStream_t st)
{
  if (steps.count < 1) return hipSuccess;
  if (steps.count > kMaxDirSteps) return hipErrorInvalidValue;
  c64 *x = static_cast<c64 *>(xh);
#define QM_FL(CP)                                                                          \
  if (k % 4 == 0) hipLaunchKernelGGL((flush_kernel<CP, 4>), QM_GRID(n, k), x, steps, n, k); \
  else hipLaunchKernelGGL((flush_kernel<CP, 1>), QM_GRID(n, k), x, steps, n, k)
  switch (prec_p) {
    case PREC_DOUBLE: QM_FL(c64); break;
    case PREC_SINGLE: QM_FL(c32); break;
    case PREC_HALF: QM_FL(c16); break;
    case PREC_FP8: QM_FL(c8); break;
    default: QM_FL(cb16);
  }
#undef QM_FL
  return hipGetLastError();
}

hipError_t blas_zero(int prec, void *v, long long n, int k, hipStream_t st)
{
  switch (prec) {
    case PREC_DOUBLE: hipLaunchKernelGGL(zero_kernel<c64>, QM_GRID(n, k), static_cast<c64 *>(v), n); break;
    case PREC_SINGLE: hipLaunchKernelGGL(zero_kernel<c32>, QM_GRID(n, k), static_cast<c32 *>(v), n); break;
    case PREC_HALF: hipLaunchKernelGGL(zero_kernel<c16>, QM_GRID(n, k), static_cast<c16 *>(v), n); break;
    case PREC_FP8: hipLaunchKernelGGL(zero_kernel<c8>, QM_GRID(n, k), static_cast<c8 *>(v), n); break;
    default: hipLaunchKernelGGL(zero_kernel<cb16>, QM_GRID(n, k), static_cast<cb16 *>(v), n);
  }
  return hipGetLastError();
}

hipError_t blas_reduce(const double *partial, long long nb, int k, double *out, double *stage, hipStream_t st)
{
  if (k < 1 || k > kBlasMaxRhs) return hipErrorInvalidValue;
  const int block = (256 / k) * k;
  const long long n = nb * k;
  long long G = (n + block * 16LL - 1) / (block * 16LL);
  if (G > kReduceStages) G = kReduceStages;
  if (G < 1) G = 1;
  hipLaunchKernelGGL(reduce_pass1, dim3(unsigned(G)), dim3(block), 0, st, partial, n, k, stage);
  hipLaunchKernelGGL(reduce_pass2, dim3(unsigned(k)), dim3(kReduceStages), 0, st, stage, int(G), k, out);
  return hipGetLastError();
}

hipError_t cg_init_state(const CgState &s, int k, hipStream_t st)
{
  hipLaunchKernelGGL(init_state_kernel, dim3((k + 63) / 64), dim3(64), 0, st, s, k);
  return hipGetLastError();
}

hipError_t cg_check_true(const CgState &s, const double *rr_true, int k, double stop2, hipStream_t st)
{
  hipLaunchKernelGGL(check_true_kernel, dim3((k + 63) / 64), dim3(64), 0, st, s, rr_true, k, stop2);
  return hipGetLastError();
}

hipError_t cg_alpha(const CgState &s, int k, hipStream_t st)
{
  hipLaunchKernelGGL(alpha_kernel, dim3((k + 63) / 64), dim3(64), 0, st, s, k);
  return hipGetLastError();
}

hipError_t cg_beta(const CgState &s, int k, double stop2, hipStream_t st)
{
  hipLaunchKernelGGL(beta_kernel, dim3((k + 63) / 64), dim3(64), 0, st, s, k, stop2);
  return hipGetLastError();
}

hipError_t cg_orth(const CgState &s, int k, hipStream_t st)
{
  hipLaunchKernelGGL(orth_kernel, dim3((k + 63) / 64), dim3(64), 0, st, s, k);
  return hipGetLastError();
}

}  // namespace qmrhs

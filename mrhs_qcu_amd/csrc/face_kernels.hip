// Face exchange of a domain-decomposed hopping term (one dimension mu split
// over processes, grid_mu > 1).
//
// The interior launch (qcuMrhsDslash with grid_mu > 1) skips every term that
// crosses the local boundary of mu, as the reference kernel does
// (include/kernels/wilson_dslash_4d.cuh:60,87,118,147,177,206,237,266). The
// two kernels here supply those terms from half spinors that the neighbouring
// process packs, so interior + faces equals the operator on the whole lattice:
//
//   pack   (sender, input parity q = 1 - p):
//     to_bwd: sites at x_mu = 0,        h = (1 - s g_mu) psi           (2N x k)
//             -> the process behind, for its forward hop
//     to_fwd: sites at x_mu = L_mu - 1, h = U_mu(x)^+ (1 + s g_mu) psi  (2N x k)
//             -> the process ahead, for its backward hop (it does not hold
//                this link, the sender does)
//   faces  (receiver, output parity p):
//     sites at x_mu = L_mu - 1: out += reconstruct(U_mu(x) from_fwd)
//     sites at x_mu = 0:        out += reconstruct(from_bwd)
//
// Face buffers are [F][2][N][k] in the spinor precision, F = sites of one
// parity on the face. Face site numbering (shared by sender and receiver,
// whose sites on a common face have the same coordinates other than x_mu):
//   mu = y, z, t: the checkerboard index order with x_mu removed;
//   mu = x:       row r = (t*Lz + z)*Ly + y of the face site, f = r / 2 (of
//                 each pair of rows exactly one holds a face site of a given
//                 parity at x = 0, the other at x = Lx - 1).
// Local extents are even, so local and global parities agree.
#include <hip/hip_runtime.h>

#include "dslash_kernel.h"
#include "face_kernels.h"

namespace qmrhs {

namespace {

struct FaceGeo {
  int subLx, Ly, Lz, Lt, S;
  int mu, F, k;
  int parity;  // output parity p of the operator the faces belong to
  float sign;
};

// Checkerboard index of face site f of dimension mu at x_mu = 0 (end = 0) or
// x_mu = L_mu - 1 (end = 1), for sites of parity par.
__device__ __forceinline__ int face_site(const FaceGeo &g, int f, int end, int par)
{
  switch (g.mu) {
    case 3: return (end ? g.Lt - 1 : 0) * (g.Lz * g.Ly * g.subLx) + f;
    case 2: {
      const int inner = g.Ly * g.subLx;
      const int t = f / inner;
      return ((t * g.Lz + (end ? g.Lz - 1 : 0)) * g.Ly) * g.subLx + (f - t * inner);
    }
    case 1: {
      const int tz = f / g.subLx;
      return (tz * g.Ly + (end ? g.Ly - 1 : 0)) * g.subLx + (f - tz * g.subLx);
    }
    default: {
      // x = 2 xh + (par ^ eo), eo = (y + z + t) & 1: x = 0 needs eo = par,
      // x = Lx - 1 needs eo = 1 - par.
      int r = 2 * f;
      const int y = r % g.Ly, zt = r / g.Ly;
      const int z = zt % g.Lz, t = zt / g.Lz;
      const int eo = (y + z + t) & 1;
      const int want = end ? 1 - par : par;
      r += (eo != want) ? 1 : 0;
      return r * g.subLx + (end ? g.subLx - 1 : 0);
    }
  }
}

template <int N, class A>
__device__ __forceinline__ void project(cpx<A> (&h)[2][N], const cpx<A> (&psi)[4][N], int D, A s)
{
  // Direction D is uniform over the launch; the switch selects the hop-table
  // entry as compile-time phases.
  switch (D) {
#define QM_PROJ(DD)                                                   \
  case DD: {                                                          \
    constexpr HopEntry e = hop_entry(DD);                             \
    _Pragma("unroll") for (int c = 0; c < N; ++c)                     \
    {                                                                 \
      h[0][c] = add_phase<e.k0>(psi[0][c], psi[e.a0][c], s);          \
      h[1][c] = add_phase<e.k1>(psi[1][c], psi[e.a1][c], s);          \
    }                                                                 \
    break;                                                            \
  }
    QM_PROJ(0)
    QM_PROJ(1)
    QM_PROJ(2)
    QM_PROJ(3)
    QM_PROJ(4)
    QM_PROJ(5)
    QM_PROJ(6)
    QM_PROJ(7)
#undef QM_PROJ
  }
}

// out[4][N] (RHS stride k) += the 4-spinor rebuilt from the half spinor g of
// direction D. D is a template argument so that every spin row index is a
// constant (a per-lane row index would put the spinor in scratch memory).
template <int D, int N, class C, class A>
__device__ __forceinline__ void add_reconstructed(C *po, const cpx<A> (&g)[2][N], A s, int k)
{
  constexpr HopEntry e = hop_entry(D);
  cpx<A> add[4][N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    add[0][i] = g[0][i];
    add[1][i] = g[1][i];
    add[e.a0][i] = add_phase<ConjPhase<e.k0>::v>(cpx<A>{A(0), A(0)}, g[0][i], s);
    add[e.a1][i] = add_phase<ConjPhase<e.k1>::v>(cpx<A>{A(0), A(0)}, g[1][i], s);
  }
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      cpx<A> v = ld_cvt<A>(po[size_t(sp * N + c) * k]);
      v.re += add[sp][c].re;
      v.im += add[sp][c].im;
      po[size_t(sp * N + c) * k] = StCvt<C>::go(v);
    }
}

// y = V x (dag = false) or V^+ x (dag = true) for the two colour vectors.
template <int N, class C, class A>
__device__ __forceinline__ void link_mul(cpx<A> (&y)[2][N], const C *V, const cpx<A> (&x)[2][N], bool dag)
{
#pragma unroll
  for (int i = 0; i < N; ++i) {
    cpx<A> a0{A(0), A(0)}, a1{A(0), A(0)};
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cpx<A> v = ld_cvt<A>(dag ? V[j * N + i] : V[i * N + j]);
      if (dag) v.im = -v.im;
      a0.re = fmaA(v.re, x[0][j].re, fmaA(-v.im, x[0][j].im, a0.re));
      a0.im = fmaA(v.re, x[0][j].im, fmaA(v.im, x[0][j].re, a0.im));
      a1.re = fmaA(v.re, x[1][j].re, fmaA(-v.im, x[1][j].im, a1.re));
      a1.im = fmaA(v.re, x[1][j].im, fmaA(v.im, x[1][j].re, a1.im));
    }
    y[0][i] = a0;
    y[1][i] = a1;
  }
}

// One lane per (face site, RHS); blockIdx.y = 0: to_bwd, 1: to_fwd.
template <int N, class C, class A>
__global__ void __launch_bounds__(256) pack_faces_kernel(C *__restrict__ to_bwd, C *__restrict__ to_fwd,
                                                         const C *__restrict__ in, const C *__restrict__ gauge,
                                                         FaceGeo g)
{
  const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= (long long)g.F * g.k) return;
  const int f = int(lane / g.k), r = int(lane - (long long)f * g.k);
  const int end = blockIdx.y;
  const int q = 1 - g.parity;
  const int site = face_site(g, f, end, q);
  const A s = A(g.sign);
  cpx<A> psi[4][N];
  const C *p = in + size_t(site) * 4 * N * g.k + r;
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < N; ++c) psi[sp][c] = ld_cvt<A>(p[size_t(sp * N + c) * g.k]);
  cpx<A> h[2][N], o[2][N];
  C *dst;
  if (!end) {
    project<N, A>(h, psi, 2 * g.mu, s);
    for (int j = 0; j < 2; ++j)
      for (int c = 0; c < N; ++c) o[j][c] = h[j][c];
    dst = to_bwd;
  } else {
    project<N, A>(h, psi, 2 * g.mu + 1, s);
    const C *V = gauge + (size_t(q) * g.S + site) * 8 * N * N + size_t(native_link_slot(2 * g.mu)) * N * N;
    link_mul<N, C, A>(o, V, h, true);
    dst = to_fwd;
  }
  C *d = dst + size_t(f) * 2 * N * g.k + r;
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int c = 0; c < N; ++c) d[size_t(j * N + c) * g.k] = StCvt<C>::go(o[j][c]);
}

// One lane per (face site, RHS); blockIdx.y = 1: sites at x_mu = L_mu - 1
// (forward hop from from_fwd), 0: sites at x_mu = 0 (backward hop from
// from_bwd). The two site sets are disjoint, so no two lanes of a launch
// update the same output.
template <int N, class C, class A>
__global__ void __launch_bounds__(256) dslash_faces_kernel(C *__restrict__ out, const C *__restrict__ from_bwd,
                                                           const C *__restrict__ from_fwd,
                                                           const C *__restrict__ gauge, FaceGeo g)
{
  const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= (long long)g.F * g.k) return;
  const int f = int(lane / g.k), r = int(lane - (long long)f * g.k);
  const int end = blockIdx.y;
  const int site = face_site(g, f, end, g.parity);
  const A s = A(g.sign);
  const C *src = (end ? from_fwd : from_bwd) + size_t(f) * 2 * N * g.k + r;
  cpx<A> h[2][N], o[2][N];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int c = 0; c < N; ++c) h[j][c] = ld_cvt<A>(src[size_t(j * N + c) * g.k]);
  int D;
  if (end) {
    const C *V =
        gauge + (size_t(g.parity) * g.S + site) * 8 * N * N + size_t(native_link_slot(2 * g.mu)) * N * N;
    link_mul<N, C, A>(o, V, h, false);
    D = 2 * g.mu;
  } else {
    for (int j = 0; j < 2; ++j)
      for (int c = 0; c < N; ++c) o[j][c] = h[j][c];
    D = 2 * g.mu + 1;
  }
  C *po = out + size_t(site) * 4 * N * g.k + r;
  // D is uniform over the launch (2 mu or 2 mu + 1 by workgroup row).
  switch (D) {
#define QM_ADD(DD) \
  case DD: add_reconstructed<DD, N, C, A>(po, o, s, g.k); break;
    QM_ADD(0)
    QM_ADD(1)
    QM_ADD(2)
    QM_ADD(3)
    QM_ADD(4)
    QM_ADD(5)
    QM_ADD(6)
    QM_ADD(7)
#undef QM_ADD
  }
}

FaceGeo face_geo(const Geom &geom, int mu, int k, int parity, int dagger)
{
  FaceGeo g;
  g.subLx = geom.L[0] / 2;
  g.Ly = geom.L[1];
  g.Lz = geom.L[2];
  g.Lt = geom.L[3];
  g.S = g.subLx * g.Ly * g.Lz * g.Lt;
  g.mu = mu;
  g.F = int(face_sites(geom, mu));
  g.k = k;
  g.parity = parity;
  g.sign = dagger ? -1.0f : 1.0f;
  return g;
}

template <int N, class C, class A>
hipError_t launch_typed(bool pack, void *o0, void *o1, const void *i0, const void *i1, const void *gauge,
                        const FaceGeo &g, hipStream_t st)
{
  const long long lanes = (long long)g.F * g.k;
  const dim3 grid(unsigned((lanes + 255) / 256), 2), block(256);
  if (pack)
    hipLaunchKernelGGL((pack_faces_kernel<N, C, A>), grid, block, 0, st, static_cast<C *>(o0), static_cast<C *>(o1),
                       static_cast<const C *>(i0), static_cast<const C *>(gauge), g);
  else
    hipLaunchKernelGGL((dslash_faces_kernel<N, C, A>), grid, block, 0, st, static_cast<C *>(o0),
                       static_cast<const C *>(i0), static_cast<const C *>(i1), static_cast<const C *>(gauge), g);
  return hipGetLastError();
}

template <class C, class A>
hipError_t launch_n(int N, bool pack, void *o0, void *o1, const void *i0, const void *i1, const void *gauge,
                    const FaceGeo &g, hipStream_t st)
{
  switch (N) {
    case 2: return launch_typed<2, C, A>(pack, o0, o1, i0, i1, gauge, g, st);
    case 3: return launch_typed<3, C, A>(pack, o0, o1, i0, i1, gauge, g, st);
    case 4: return launch_typed<4, C, A>(pack, o0, o1, i0, i1, gauge, g, st);
    case 5: return launch_typed<5, C, A>(pack, o0, o1, i0, i1, gauge, g, st);
    case 6: return launch_typed<6, C, A>(pack, o0, o1, i0, i1, gauge, g, st);
    case 7: return launch_typed<7, C, A>(pack, o0, o1, i0, i1, gauge, g, st);
    default: return launch_typed<8, C, A>(pack, o0, o1, i0, i1, gauge, g, st);
  }
}

hipError_t launch_any(int prec, int N, bool pack, void *o0, void *o1, const void *i0, const void *i1,
                      const void *gauge, const FaceGeo &g, hipStream_t st)
{
  switch (prec) {
    case PREC_DOUBLE: return launch_n<c64, double>(N, pack, o0, o1, i0, i1, gauge, g, st);
    case PREC_SINGLE: return launch_n<c32, float>(N, pack, o0, o1, i0, i1, gauge, g, st);
    case PREC_HALF: return launch_n<c16, float>(N, pack, o0, o1, i0, i1, gauge, g, st);
    default: return launch_n<cb16, float>(N, pack, o0, o1, i0, i1, gauge, g, st);
  }
}

}  // namespace

long long face_sites(const Geom &g, int mu)
{
  const long long S = (long long)(g.L[0] / 2) * g.L[1] * g.L[2] * g.L[3];
  return mu == 0 ? S / (g.L[0] / 2) / 2 : S / g.L[mu];
}

hipError_t launch_pack_faces(void *to_bwd, void *to_fwd, const void *in, const void *gauge, int prec, int k,
                             int parity, int dagger, const Geom &geom, int mu, hipStream_t st)
{
  const FaceGeo g = face_geo(geom, mu, k, parity, dagger);
  return launch_any(prec, geom.N, true, to_bwd, to_fwd, in, nullptr, gauge, g, st);
}

hipError_t launch_dslash_faces(void *out, const void *from_bwd, const void *from_fwd, const void *gauge, int prec,
                               int k, int parity, int dagger, const Geom &geom, int mu, hipStream_t st)
{
  const FaceGeo g = face_geo(geom, mu, k, parity, dagger);
  return launch_any(prec, geom.N, false, out, nullptr, from_bwd, from_fwd, gauge, g, st);
}

}  // namespace qmrhs

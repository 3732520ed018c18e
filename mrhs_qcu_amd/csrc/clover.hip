// Clover term: field construction from the gauge links, 6x6-block inversion
// and the site-local application to packed multi-RHS spinors.
//
//   A(x) = 1 + coeff * sum_{mu<nu} sigma_{mu nu} (x) F_{mu nu}(x)
//   sigma_{mu nu} = (i/2) [gamma_mu, gamma_nu]      (DeGrand-Rossi basis)
//   F_{mu nu}(x)  = (Q_{mu nu}(x) - Q_{mu nu}(x)^+) / (8i)
//   Q_{mu nu}(x)  = the four plaquettes in the (mu, nu) plane that start and
//                   end at x ("clover leaves"), each a product of 4 links.
//
// gamma_5 = diag(1, 1, -1, -1) in this basis, so every sigma_{mu nu} is block
// diagonal in the spin pairs {0, 1} and {2, 3}, and A(x) is two Hermitian
// (2N x 2N) blocks. Field layout: [2 parity][S][2 block][2N][2N] complex,
// row/column index (spin within the block) * N + colour; 8 N^2 complex per
// site (72 at N = 3, the per-site size of the reference's reserved clover
// field, include/basic_data/qcu_point.cuh:98-101).
#include <hip/hip_runtime.h>

#include "clover.h"
#include "qcu_device.h"

namespace qmrhs {

namespace {

struct CloverGeo {
  int Lx, Ly, Lz, Lt, S;
  double coeff;
  SigmaBlocks sig;
};

// Checkerboard parity and index of full coordinates (periodic wrap).
__device__ __forceinline__ void site_of(const CloverGeo &g, int x, int y, int z, int t, int &par, int &idx)
{
  x = (x + g.Lx) % g.Lx;
  y = (y + g.Ly) % g.Ly;
  z = (z + g.Lz) % g.Lz;
  t = (t + g.Lt) % g.Lt;
  par = (x + y + z + t) & 1;
  idx = ((t * g.Lz + z) * g.Ly + y) * (g.Lx / 2) + x / 2;
}

template <int N>
struct Mat {
  double re[N][N], im[N][N];
};

template <int N, class CSrc>
__device__ __forceinline__ void load_link(Mat<N> &m, const CSrc *gauge, const CloverGeo &g, int mu, const int (&c)[4],
                                          bool dag)
{
  int par, idx;
  site_of(g, c[0], c[1], c[2], c[3], par, idx);
  const CSrc *p = gauge + ((size_t(mu) * 2 + par) * g.S + idx) * N * N;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const CSrc v = dag ? p[j * N + i] : p[i * N + j];
      m.re[i][j] = double(v.x);
      m.im[i][j] = dag ? -double(v.y) : double(v.y);
    }
}

template <int N>
__device__ __forceinline__ void mul(Mat<N> &o, const Mat<N> &a, const Mat<N> &b)
{
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double r = 0.0, m = 0.0;
      for (int l = 0; l < N; ++l) {
        r += a.re[i][l] * b.re[l][j] - a.im[i][l] * b.im[l][j];
        m += a.re[i][l] * b.im[l][j] + a.im[i][l] * b.re[l][j];
      }
      o.re[i][j] = r;
      o.im[i][j] = m;
    }
}

// One leaf: L1 L2 L3 L4, each link given by direction, offset from x, dagger.
struct LinkRef {
  int mu, d[4];
  bool dag;
};

template <int N, class CSrc>
__device__ __forceinline__ void add_leaf(Mat<N> &q, const CSrc *gauge, const CloverGeo &g, const int (&x)[4],
                                         const LinkRef (&l)[4])
{
  Mat<N> acc, u, t;
  for (int k = 0; k < 4; ++k) {
    int c[4];
    for (int d = 0; d < 4; ++d) c[d] = x[d] + l[k].d[d];
    load_link<N, CSrc>(u, gauge, g, l[k].mu, c, l[k].dag);
    if (k == 0) {
      acc = u;
    } else {
      mul<N>(t, acc, u);
      acc = t;
    }
  }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      q.re[i][j] += acc.re[i][j];
      q.im[i][j] += acc.im[i][j];
    }
}

// One lane per (parity, site): A(x) in FP64.
template <int N, class CSrc>
__global__ void __launch_bounds__(64) clover_compute_kernel(c64 *__restrict__ out, const CSrc *__restrict__ gauge,
                                                            CloverGeo g)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2LL * g.S) return;
  const int par = int(i / g.S), idx = int(i - (long long)par * g.S);
  const int subLx = g.Lx / 2;
  const int xh = idx % subLx, r1 = idx / subLx;
  const int y = r1 % g.Ly, r2 = r1 / g.Ly;
  const int z = r2 % g.Lz, t = r2 / g.Lz;
  const int xx = 2 * xh + ((par + y + z + t) & 1);
  const int x[4] = {xx, y, z, t};
  constexpr int B = 2 * N;
  // A blocks accumulated in place: start from the identity.
  c64 *a = out + size_t(i) * 2 * B * B;
  for (int e = 0; e < 2 * B * B; ++e) a[e] = c64{0.0, 0.0};
  for (int b = 0; b < 2; ++b)
    for (int e = 0; e < B; ++e) a[b * B * B + e * B + e].x = 1.0;
  int plane = 0;
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++plane) {
      int em[4] = {0, 0, 0, 0}, en[4] = {0, 0, 0, 0};
      em[mu] = 1;
      en[nu] = 1;
      auto off = [](const int (&p)[4], int s1, const int (&q)[4], int s2, int (&o)[4]) {
        for (int d = 0; d < 4; ++d) o[d] = s1 * p[d] + s2 * q[d];
      };
      Mat<N> q;
      for (int ii = 0; ii < N; ++ii)
        for (int jj = 0; jj < N; ++jj) q.re[ii][jj] = q.im[ii][jj] = 0.0;
      LinkRef l[4];
      // U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
      l[0] = {mu, {0, 0, 0, 0}, false};
      l[1] = {nu, {0, 0, 0, 0}, false};
      off(em, 1, en, 0, l[1].d);
      l[2] = {mu, {0, 0, 0, 0}, true};
      off(em, 0, en, 1, l[2].d);
      l[3] = {nu, {0, 0, 0, 0}, true};
      add_leaf<N, CSrc>(q, gauge, g, x, l);
      // U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu)
      l[0] = {nu, {0, 0, 0, 0}, false};
      l[1] = {mu, {0, 0, 0, 0}, true};
      off(em, -1, en, 1, l[1].d);
      l[2] = {nu, {0, 0, 0, 0}, true};
      off(em, -1, en, 0, l[2].d);
      l[3] = {mu, {0, 0, 0, 0}, false};
      off(em, -1, en, 0, l[3].d);
      add_leaf<N, CSrc>(q, gauge, g, x, l);
      // U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu)
      l[0] = {mu, {0, 0, 0, 0}, true};
      off(em, -1, en, 0, l[0].d);
      l[1] = {nu, {0, 0, 0, 0}, true};
      off(em, -1, en, -1, l[1].d);
      l[2] = {mu, {0, 0, 0, 0}, false};
      off(em, -1, en, -1, l[2].d);
      l[3] = {nu, {0, 0, 0, 0}, false};
      off(em, 0, en, -1, l[3].d);
      add_leaf<N, CSrc>(q, gauge, g, x, l);
      // U_nu(x-nu)^+ U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^+
      l[0] = {nu, {0, 0, 0, 0}, true};
      off(em, 0, en, -1, l[0].d);
      l[1] = {mu, {0, 0, 0, 0}, false};
      off(em, 0, en, -1, l[1].d);
      l[2] = {nu, {0, 0, 0, 0}, false};
      off(em, 1, en, -1, l[2].d);
      l[3] = {mu, {0, 0, 0, 0}, true};
      add_leaf<N, CSrc>(q, gauge, g, x, l);
      // F = (Q - Q^+) / (8i): F_ij = -i/8 (Q_ij - conj(Q_ji))
      for (int b = 0; b < 2; ++b)
        for (int s1 = 0; s1 < 2; ++s1)
          for (int s2 = 0; s2 < 2; ++s2) {
            const double sr = g.sig.re[plane][b][s1][s2], si = g.sig.im[plane][b][s1][s2];
            if (sr == 0.0 && si == 0.0) continue;
            for (int c1 = 0; c1 < N; ++c1)
              for (int c2 = 0; c2 < N; ++c2) {
                const double dr = q.re[c1][c2] - q.re[c2][c1];
                const double di = q.im[c1][c2] + q.im[c2][c1];
                // F = -i/8 (dr + i di) = (di - i dr) / 8
                const double fr = di / 8.0, fi = -dr / 8.0;
                // coeff * sigma * F
                const double vr = g.coeff * (sr * fr - si * fi), vi = g.coeff * (sr * fi + si * fr);
                c64 &e = a[b * B * B + (s1 * N + c1) * B + (s2 * N + c2)];
                e.x += vr;
                e.y += vi;
              }
          }
    }
}

// One lane per (parity, site, block): Gauss-Jordan inverse of a (2N x 2N)
// block with partial pivoting, FP64, in the lane's scratch.
template <int N>
__global__ void __launch_bounds__(64) clover_invert_kernel(c64 *__restrict__ inv, const c64 *__restrict__ a,
                                                           long long nblocks)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblocks) return;
  constexpr int B = 2 * N;
  double mr[B][2 * B], mi[B][2 * B];
  const c64 *src = a + size_t(i) * B * B;
  for (int r = 0; r < B; ++r)
    for (int c = 0; c < B; ++c) {
      mr[r][c] = src[r * B + c].x;
      mi[r][c] = src[r * B + c].y;
      mr[r][B + c] = (r == c) ? 1.0 : 0.0;
      mi[r][B + c] = 0.0;
    }
  for (int col = 0; col < B; ++col) {
    int piv = col;
    double best = mr[col][col] * mr[col][col] + mi[col][col] * mi[col][col];
    for (int r = col + 1; r < B; ++r) {
      const double v = mr[r][col] * mr[r][col] + mi[r][col] * mi[r][col];
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (piv != col)
      for (int c = 0; c < 2 * B; ++c) {
        const double tr = mr[col][c], ti = mi[col][c];
        mr[col][c] = mr[piv][c];
        mi[col][c] = mi[piv][c];
        mr[piv][c] = tr;
        mi[piv][c] = ti;
      }
    // row col /= pivot
    const double pr = mr[col][col], pi = mi[col][col];
    const double den = pr * pr + pi * pi;
    const double ir = pr / den, ii = -pi / den;
    for (int c = 0; c < 2 * B; ++c) {
      const double vr = mr[col][c], vi = mi[col][c];
      mr[col][c] = vr * ir - vi * ii;
      mi[col][c] = vr * ii + vi * ir;
    }
    for (int r = 0; r < B; ++r) {
      if (r == col) continue;
      const double fr = mr[r][col], fi = mi[r][col];
      if (fr == 0.0 && fi == 0.0) continue;
      for (int c = 0; c < 2 * B; ++c) {
        mr[r][c] -= fr * mr[col][c] - fi * mi[col][c];
        mi[r][c] -= fr * mi[col][c] + fi * mr[col][c];
      }
    }
  }
  c64 *dst = inv + size_t(i) * B * B;
  for (int r = 0; r < B; ++r)
    for (int c = 0; c < B; ++c) dst[r * B + c] = c64{mr[r][B + c], mi[r][B + c]};
}

// out(x) = A(x) in(x) on sites of one parity, packed [S][4][N][k]. One lane
// per (site, RHS); the spinor loads are coalesced over the RHS. With STAGE
// the clover blocks of the workgroup's sites are first copied to LDS with
// 16-byte loads (each site's 8N^2 entries read from HBM once per workgroup),
// and the k lanes of a site read them there (same address: broadcast);
// otherwise every lane reads them from global memory.
template <int N, class C, class A, bool STAGE>
__global__ void __launch_bounds__(256) clover_apply_kernel(C *__restrict__ out, const C *__restrict__ in,
                                                           const C *__restrict__ clov, int parity, int S, int k)
{
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int B = 2 * N;
  constexpr int SITE = 2 * B * B;  // complex numbers per site
  const long long lane0 = (long long)blockIdx.x * blockDim.x;
  const long long lane = lane0 + threadIdx.x;
  const long long nl = (long long)S * k;
  const C *cl;
  const int site = int(lane / k), r = int(lane - (long long)site * k);
  if constexpr (STAGE) {
    static_assert((SITE * sizeof(C)) % 16 == 0, "site block must be whole 16-byte chunks");
    constexpr int CH = int(SITE * sizeof(C)) / 16;
    const int first = int(lane0 / k);
    long long lastl = lane0 + blockDim.x - 1;
    if (lastl >= nl) lastl = nl - 1;
    const int nsite = int(lastl / k) - first + 1;
    const uint4 *src = reinterpret_cast<const uint4 *>(clov + (size_t(parity) * S + first) * SITE);
    uint4 *dst = reinterpret_cast<uint4 *>(smem);
    for (int i = threadIdx.x; i < nsite * CH; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    cl = reinterpret_cast<const C *>(smem) + size_t(site - first) * SITE;
  } else {
    cl = clov + (size_t(parity) * S + site) * SITE;
  }
  if (lane >= nl) return;
  const C *pin = in + size_t(site) * 4 * N * k + r;
  cpx<A> v[4 * N];
#pragma unroll
  for (int e = 0; e < 4 * N; ++e) v[e] = ld_cvt<A>(pin[size_t(e) * k]);
  C *po = out + size_t(site) * 4 * N * k + r;
#pragma unroll
  for (int b = 0; b < 2; ++b) {
#pragma unroll
    for (int row = 0; row < B; ++row) {
      cpx<A> acc{A(0), A(0)};
#pragma unroll
      for (int c = 0; c < B; ++c) {
        const cpx<A> m = ld_cvt<A>(cl[b * B * B + row * B + c]);
        const cpx<A> x = v[b * B + c];
        acc.re += m.re * x.re - m.im * x.im;
        acc.im += m.re * x.im + m.im * x.re;
      }
      po[size_t(b * B + row) * k] = StCvt<C>::go(acc);
    }
  }
}

template <int N>
hipError_t compute_n(c64 *a, const void *gauge, int gauge_prec, const CloverGeo &g, hipStream_t st)
{
  const long long n = 2LL * g.S;
  const dim3 grid(unsigned((n + 63) / 64)), block(64);
  if (gauge_prec == PREC_DOUBLE)
    hipLaunchKernelGGL((clover_compute_kernel<N, c64>), grid, block, 0, st, a, static_cast<const c64 *>(gauge), g);
  else
    hipLaunchKernelGGL((clover_compute_kernel<N, c32>), grid, block, 0, st, a, static_cast<const c32 *>(gauge), g);
  return hipGetLastError();
}

template <int N>
hipError_t invert_n(c64 *inv, const c64 *a, long long nblocks, hipStream_t st)
{
  hipLaunchKernelGGL((clover_invert_kernel<N>), dim3(unsigned((nblocks + 63) / 64)), dim3(64), 0, st, inv, a,
                     nblocks);
  return hipGetLastError();
}

template <int N, class C, class A>
hipError_t apply_t(void *out, const void *in, const void *clov, int parity, int S, int k, hipStream_t st)
{
  const long long lanes = (long long)S * k;
  const dim3 grid(unsigned((lanes + 255) / 256)), block(256);
  // Stage in LDS when a workgroup covers few sites (k >= 8: at most 34) and
  // the field allows 16-byte loads.
  if (k >= 8 && uintptr_t(clov) % 16 == 0) {
    const size_t shm = size_t(255 / k + 2) * 8 * N * N * sizeof(C);
    hipLaunchKernelGGL((clover_apply_kernel<N, C, A, true>), grid, block, shm, st, static_cast<C *>(out),
                       static_cast<const C *>(in), static_cast<const C *>(clov), parity, S, k);
  } else {
    hipLaunchKernelGGL((clover_apply_kernel<N, C, A, false>), grid, block, 0, st, static_cast<C *>(out),
                       static_cast<const C *>(in), static_cast<const C *>(clov), parity, S, k);
  }
  return hipGetLastError();
}

template <int N>
hipError_t apply_n(int prec, void *out, const void *in, const void *clov, int parity, int S, int k, hipStream_t st)
{
  switch (prec) {
    case PREC_DOUBLE: return apply_t<N, c64, double>(out, in, clov, parity, S, k, st);
    case PREC_SINGLE: return apply_t<N, c32, float>(out, in, clov, parity, S, k, st);
    case PREC_HALF: return apply_t<N, c16, float>(out, in, clov, parity, S, k, st);
    default: return apply_t<N, cb16, float>(out, in, clov, parity, S, k, st);
  }
}

}  // namespace

hipError_t clover_compute(c64 *a, const void *gauge, int gauge_prec, const int *L, int N, double coeff,
                          const SigmaBlocks &sig, hipStream_t st)
{
  CloverGeo g;
  g.Lx = L[0];
  g.Ly = L[1];
  g.Lz = L[2];
  g.Lt = L[3];
  g.S = L[0] / 2 * L[1] * L[2] * L[3];
  g.coeff = coeff;
  g.sig = sig;
  switch (N) {
    case 2: return compute_n<2>(a, gauge, gauge_prec, g, st);
    case 3: return compute_n<3>(a, gauge, gauge_prec, g, st);
    case 4: return compute_n<4>(a, gauge, gauge_prec, g, st);
    default: return compute_n<5>(a, gauge, gauge_prec, g, st);
  }
}

hipError_t clover_invert(c64 *inv, const c64 *a, long long nblocks, int N, hipStream_t st)
{
  switch (N) {
    case 2: return invert_n<2>(inv, a, nblocks, st);
    case 3: return invert_n<3>(inv, a, nblocks, st);
    case 4: return invert_n<4>(inv, a, nblocks, st);
    default: return invert_n<5>(inv, a, nblocks, st);
  }
}

hipError_t clover_apply(int prec, void *out, const void *in, const void *clov, int parity, int S, int k, int N,
                        hipStream_t st)
{
  switch (N) {
    case 2: return apply_n<2>(prec, out, in, clov, parity, S, k, st);
    case 3: return apply_n<3>(prec, out, in, clov, parity, S, k, st);
    case 4: return apply_n<4>(prec, out, in, clov, parity, S, k, st);
    default: return apply_n<5>(prec, out, in, clov, parity, S, k, st);
  }
}

}  // namespace qmrhs

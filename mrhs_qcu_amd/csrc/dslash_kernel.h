// Multi-right-hand-side even-odd Wilson hopping term for SU(N) on gfx950.
//
// Operator (reference include/kernels/wilson_dslash_4d.cuh:59-293): for an
// output site x of parity p and s = +1 (daggerFlag 0) / -1 (daggerFlag 1),
//   out(x) = sum_mu [ U_mu(x) (1 - s g_mu) in(x+mu) + U_mu(x-mu)^+ (1 + s g_mu) in(x-mu) ]
// in the DeGrand-Rossi basis, periodic in all four directions, with the terms
// that cross the local boundary skipped in every dimension whose process grid
// is larger than one (reference wilson_dslash_4d.cuh:60,87,118,147,177,206,237,266).
//
// Each hop is done on half spinors: the 4N input components are projected to
// two colour vectors h0 = in_0 + k0 s in_a0 and h1 = in_1 + k1 s in_a1, both are
// multiplied by the link V (U_mu(x) forward, U_mu(x-mu)^+ backward), and the
// result is reconstructed as out_0 += V h0, out_1 += V h1,
// out_a0 += conj(k0) s V h0, out_a1 += conj(k1) s V h1 (table kHop below).
//
// Work decomposition: one lane owns one output site and R consecutive
// right-hand sides, so the link of a site is fetched once per R RHS (and once
// per workgroup when it is staged in LDS), and in the native layout the 8
// neighbour spinor blocks are read as contiguous RHS-innermost vectors.
#pragma once

#include <type_traits>

#include "qcu_device.h"

namespace qmrhs {

enum SpinorLayout : int {
  SL_PACKED = 0,    // native: [S][4][N][k] complex, RHS innermost
  SL_AOS_PTRS = 1,  // one pointer per RHS, each the user layout [S][4][N]
  SL_SOA_PTRS = 2,  // one pointer per RHS, each the reference SoA layout [4*N][S]
  SL_AOS_COOP = 3,  // the user AoS layout of SL_AOS_PTRS read and written by site groups
                    // (coop_issue / coop_finish / coop_store): a kernel variant, not a format
};
enum GaugeLayout : int {
  GL_NATIVE = 0,   // [2][S][8][N][N]: per site 4 forward links then 4 daggered backward links
  GL_REF_SOA = 1,  // reference SoA [4][2][N*N][S]; backward links daggered on the fly
  GL_NATIVE12 = 2, // SU(3) only: [2][S][8][2][3], the native links as their first two rows;
                   // row 2 = conj(row 0 x row 1) is rebuilt in registers
  GL_NATIVE8 = 3,  // SU(3) only: [2][S][8][4], eight reals per link: U01, U02, U10 and
                   // (arg U00, arg U20); rebuilt while the links are staged in LDS
};

// Stored complex numbers per link of a gauge layout.
__host__ __device__ constexpr int link_cplx(int GL, int N)
{
  return GL == GL_NATIVE12 ? 2 * N : (GL == GL_NATIVE8 ? 4 : N * N);
}
// Layouts that store each site's 8 links as one block ([2][S][8][...]).
__host__ __device__ constexpr bool site_blocked(int GL)
{
  return GL == GL_NATIVE || GL == GL_NATIVE12 || GL == GL_NATIVE8;
}

constexpr int kMaxPtrs = 32;  // RHS per launch for the pointer layouts
struct PtrPack {
  void *p[kMaxPtrs];
};

struct DslashArgs {
  void *out;
  const void *in;
  const void *gauge;
  const void *x;          // xpay operand (same layout and type as in), or null
  const float *sc_in;     // per-RHS: true input = stored input * sc_in[r]      (null = 1)
  const float *sc_x;      // per-RHS: true x     = stored x     * sc_x[r]       (null = 1)
  const float *sc_out;    // per-RHS: stored out = true out / sc_out[r]         (null = 1)
  double xa, xb;          // true out = xa * x + xb * D(in)   (x null: xb * D(in))
  double *partial;        // per-workgroup per-RHS sum of |true out|^2, [nblocks][k] (nullable)
  // Fused CG update (packed layout, FP32 or FP64 output; rx_r null: off):
  // with v = the true out, r -= alpha v in place (a packed field of the
  // output type), out is not written and partial receives |r|^2 instead.
  void *rx_r;
  const double *rx_alpha;
  float sign;             // +1 (daggerFlag 0) or -1 (daggerFlag 1)
  int k;                  // RHS count (RHS stride of the packed layout)
  int P;                  // lanes per site = k / R
  int parity;             // output parity
  int subLx, Ly, Lz, Lt;
  int S;                  // sites per parity
  int skip[4];            // dimension has grid > 1
  int tblock, zblock;     // traversal order (t_out, z_out, t_in, z_in, y, x/2) in blocks of
                          // tblock time and zblock z slices (natural order: 1, Lz)
  int nt_store;           // output written with non-temporal stores
  int nt_link;            // links read with non-temporal loads
  int xcd_map;            // remap workgroups so each XCD owns a contiguous range
  uint32_t nlanes;        // S * P
  uint32_t nblocks;
  FastDiv div_P, div_subLx, div_row, div_tb, div_zb, div_nzb, div_S, div_Ly;
  int ystream;  // traversal (t_out, z_out, y, t_in, z_in, x/2) instead of (t_out, z_out, t_in, z_in, y, x/2)
  int t_flip;   // GL_NATIVE12: the time links that cross the t boundary carry a factor -1 (antiperiodic
                // time folded into the links), which two stored rows cannot express
  int aos_vec;  // pointer layouts: every buffer 16-byte aligned (16-byte loads and stores)
  int plain_out;  // packed output is D(in) as it is: no xpay operand, scales, norms or fused update
  uint32_t efac_off;  // > 0 (LDS kernels): byte offset in dynamic LDS of the staged per-RHS epilogue
                      // factors [4][k] in the accumulation type (xb sc_in, xa sc_x, 1 / sc_out, rx_alpha)
  uint32_t xcd_gran;  // > 0: XCD-cyclic workgroup order in chunks of xcd_gran workgroups (xcd_cyclic)
  uint32_t coop_off;  // SL_AOS_COOP: byte offset in dynamic LDS of the group exchange area (16 B
                      // per piece, blockDim.x * pieces-per-record pieces)
  // Hypercube kernels (dslash_hcube.h): the launch walks a lattice of
  // 2x2x2x2 blocks (subLx x Ly/2 x Lz/2 x Lt/2) in the traversal order above,
  // with tblock / zblock counted in blocks; 2^hc_lp lanes per site.
  uint32_t hc_lp;
  int bLy, bLz;
  FastDiv hdiv_Ly, hdiv_row, hdiv_nzb;
  // Column order (traversal bit 2): the lattice (site-major kernels, units
  // of sites, nT = Lt) or the block lattice (hypercube kernels, units of
  // 2x2x2x2 blocks, nT = Lt / 2) is cut into columns of col_cy x col_cz in
  // (y, z), whole in x and t; the logical site (block) number is
  // ((column * nT + T) * slab + b) with b = (z_in * col_cy + y_in) * subLx +
  // xh inside the column's slab at time T. Every column is a whole number
  // (col_grp) of workgroups; XCD x walks column 8 g + x slab by slab from
  // T = 0 to nT - 1, for g = 0, 1, ... (col_remap); the launch has
  // 8 ceil(col_ncol / 8) col_grp workgroups, the ones past the last column
  // idle.
  // With traversal bit 3 the columns also split x in two (col_sx = subLx / 2
  // checkerboard x indices each, col_nx = 2; otherwise col_sx = subLx,
  // col_nx = 1), the x part of the column index running fastest.
  int col_order;
  uint32_t col_cy, col_cz, col_ncol, col_ncy, col_nT, col_slab, col_grp, col_sx, col_nx;
  FastDiv cdiv_slab, cdiv_nT, cdiv_cy, cdiv_ncy, cdiv_grp, cdiv_sx, cdiv_nx;
  uint32_t ngrid;  // workgroups launched: nblocks, or the padded column-order grid
  // Pointer layouts only. Kept last: the packed kernels read only the fields
  // above, which then share a few scalar-cache lines of the argument block
  // (at 24^3x48, 16 RHS, FP32 the packed kernel took 4 us longer with these
  // 0.8 KB of pointer arrays placed before its fields).
  int xptr;               // x_ptrs holds a per-RHS xpay operand
  PtrPack out_ptrs, in_ptrs, x_ptrs;
};

// Hop table: direction D = 2*mu + (backward); projection partner spins a0/a1
// and phases k0/k1 (0: +1, 1: -1, 2: +i, 3: -i). Read off the reference's
// projector lines wilson_dslash_4d.cuh:74-81 (X+), 102-112 (X-), 134-141 (Y+),
// 162-171 (Y-), 193-200 (Z+), 222-231 (Z-), 253-260 (T+), 281-290 (T-).
struct HopEntry {
  int a0, k0, a1, k1;
};
__host__ __device__ constexpr HopEntry hop_entry(int D)
{
  switch (D) {
    case 0: return {3, 3, 2, 3};  // X+ : h0 = in0 - i s in3, h1 = in1 - i s in2
    case 1: return {3, 2, 2, 2};  // X- : h0 = in0 + i s in3, h1 = in1 + i s in2
    case 2: return {3, 0, 2, 1};  // Y+ : h0 = in0 +   s in3, h1 = in1 -   s in2
    case 3: return {3, 1, 2, 0};  // Y- : h0 = in0 -   s in3, h1 = in1 +   s in2
    case 4: return {2, 3, 3, 2};  // Z+ : h0 = in0 - i s in2, h1 = in1 + i s in3
    case 5: return {2, 2, 3, 3};  // Z- : h0 = in0 + i s in2, h1 = in1 - i s in3
    case 6: return {2, 1, 3, 1};  // T+ : h0 = in0 -   s in2, h1 = in1 -   s in3
    default: return {2, 0, 3, 0};  // T- : h0 = in0 +   s in2, h1 = in1 +   s in3
  }
}

// Slot of direction D inside a site's 8-link block of the native gauge layout.
__host__ __device__ constexpr int native_link_slot(int D) { return (D & 1) ? 4 + (D >> 1) : (D >> 1); }

template <class A>
__device__ __forceinline__ A fmaA(A a, A b, A c)
{
  return __builtin_fmaf(a, b, c);
}
template <>
__device__ __forceinline__ double fmaA<double>(double a, double b, double c)
{
  return __builtin_fma(a, b, c);
}

// base + (phase * s) * z, phase code 0: +1, 1: -1, 2: +i, 3: -i; one FMA per
// real component (the sign s = +-1 is a run-time operand of the FMA).
template <int KP, class A>
__device__ __forceinline__ cpx<A> add_phase(cpx<A> base, cpx<A> z, A s)
{
  if constexpr (KP == 0) return cpx<A>{fmaA(s, z.re, base.re), fmaA(s, z.im, base.im)};
  else if constexpr (KP == 1) return cpx<A>{fmaA(-s, z.re, base.re), fmaA(-s, z.im, base.im)};
  else if constexpr (KP == 2) return cpx<A>{fmaA(-s, z.im, base.re), fmaA(s, z.re, base.im)};
  else return cpx<A>{fmaA(s, z.im, base.re), fmaA(-s, z.re, base.im)};
}
template <int KP> struct ConjPhase { static constexpr int v = KP < 2 ? KP : (KP == 2 ? 3 : 2); };

// acc += the hop in direction D for one RHS, given the link V and the full
// neighbour spinor psi[4][N]: 2N projections, two N x N complex mat-vecs as
// FMA chains (one multiply + 2N-1 FMAs per real output) and 4N
// reconstruction updates.
template <int N, int D, class A>
__device__ __forceinline__ void hop_one(cpx<A> (&acc)[4][N], const cpx<A> (&V)[N][N], const cpx<A> (&psi)[4][N], A s)
{
  constexpr HopEntry e = hop_entry(D);
  cpx<A> h0[N], h1[N];
#pragma unroll
  for (int c = 0; c < N; ++c) {
    h0[c] = add_phase<e.k0>(psi[0][c], psi[e.a0][c], s);
    h1[c] = add_phase<e.k1>(psi[1][c], psi[e.a1][c], s);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    cpx<A> g0, g1;
    g0.re = V[i][0].re * h0[0].re;
    g0.im = V[i][0].re * h0[0].im;
    g1.re = V[i][0].re * h1[0].re;
    g1.im = V[i][0].re * h1[0].im;
    g0.re = fmaA(-V[i][0].im, h0[0].im, g0.re);
    g0.im = fmaA(V[i][0].im, h0[0].re, g0.im);
    g1.re = fmaA(-V[i][0].im, h1[0].im, g1.re);
    g1.im = fmaA(V[i][0].im, h1[0].re, g1.im);
#pragma unroll
    for (int j = 1; j < N; ++j) {
      const cpx<A> v = V[i][j];
      g0.re = fmaA(v.re, h0[j].re, g0.re);
      g0.im = fmaA(v.re, h0[j].im, g0.im);
      g1.re = fmaA(v.re, h1[j].re, g1.re);
      g1.im = fmaA(v.re, h1[j].im, g1.im);
      g0.re = fmaA(-v.im, h0[j].im, g0.re);
      g0.im = fmaA(v.im, h0[j].re, g0.im);
      g1.re = fmaA(-v.im, h1[j].im, g1.re);
      g1.im = fmaA(v.im, h1[j].re, g1.im);
    }
    acc[0][i].re += g0.re;
    acc[0][i].im += g0.im;
    acc[1][i].re += g1.re;
    acc[1][i].im += g1.im;
    acc[e.a0][i] = add_phase<ConjPhase<e.k0>::v>(acc[e.a0][i], g0, s);
    acc[e.a1][i] = add_phase<ConjPhase<e.k1>::v>(acc[e.a1][i], g1, s);
  }
}

// The hop for all R right-hand sides of a lane with the link read from LDS
// one row at a time inside the mat-vec (each row used by every RHS), for
// large N, where a whole N x N link in registers does not fit beside the
// accumulators and two neighbour spinors.
template <int N, int D, int R, class A, class CG>
__device__ __forceinline__ void hop_rows(cpx<A> (&acc)[R][4][N], const CG *vp, const cpx<A> (&psi)[R][4][N], A s)
{
  constexpr HopEntry e = hop_entry(D);
  cpx<A> h0[R][N], h1[R][N];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      h0[r][c] = add_phase<e.k0>(psi[r][0][c], psi[r][e.a0][c], s);
      h1[r][c] = add_phase<e.k1>(psi[r][1][c], psi[r][e.a1][c], s);
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    cpx<A> v[N];
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = ld_cvt<A>(vp[i * N + j]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cpx<A> g0{A(0), A(0)}, g1{A(0), A(0)};
#pragma unroll
      for (int j = 0; j < N; ++j) {
        g0.re = fmaA(v[j].re, h0[r][j].re, g0.re);
        g0.im = fmaA(v[j].re, h0[r][j].im, g0.im);
        g1.re = fmaA(v[j].re, h1[r][j].re, g1.re);
        g1.im = fmaA(v[j].re, h1[r][j].im, g1.im);
        g0.re = fmaA(-v[j].im, h0[r][j].im, g0.re);
        g0.im = fmaA(v[j].im, h0[r][j].re, g0.im);
        g1.re = fmaA(-v[j].im, h1[r][j].im, g1.re);
        g1.im = fmaA(v[j].im, h1[r][j].re, g1.im);
      }
      acc[r][0][i].re += g0.re;
      acc[r][0][i].im += g0.im;
      acc[r][1][i].re += g1.re;
      acc[r][1][i].im += g1.im;
      acc[r][e.a0][i] = add_phase<ConjPhase<e.k0>::v>(acc[r][e.a0][i], g0, s);
      acc[r][e.a1][i] = add_phase<ConjPhase<e.k1>::v>(acc[r][e.a1][i], g1, s);
    }
  }
}

// Large-N threshold of the row-wise link path (native links, from LDS when
// staged, else from global memory; no link is prefetched into registers).
constexpr int kRowLinkN = 5;

// Site coordinates and neighbours of one output site.
struct SiteGeo {
  int idx;       // checkerboard index ((t*Lz+z)*Ly+y)*subLx+xh
  int nbr[8];    // neighbour index per direction D
  int live[8];   // 0 if the term is skipped at a local boundary
  int xh;        // checkerboard x index
  // X- neighbour sharing: the X- neighbour of a site is the X+ neighbour of
  // the site before it in its x row (same y, z, t). When that site's lanes sit
  // in the same wavefront (lane - xstride), the X- spinor is taken from them
  // with a cross-lane read instead of a second load; xown marks the lanes
  // that load it themselves (first site of a row or of a wavefront).
  int xown, xsrc;
  // GL_NATIVE12 with t_flip: negate the rebuilt T+ link (t = Lt - 1) / T- link (t = 0).
  int tneg[2];
};

// 32-bit cross-lane read of lane `src` (ds_bpermute); doubles take two.
__device__ __forceinline__ float lane_read(float v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ double lane_read(double v, int src) { return __shfl(v, src, 64); }

// Column order: the logical workgroup of hardware workgroup hw (XCD hw % 8
// walks column 8 g + hw % 8, g = (hw / 8) / col_grp), or 0xffffffff for a
// workgroup past the last column.
__device__ __forceinline__ uint32_t col_remap(uint32_t hw, const DslashArgs &a)
{
  const uint32_t i = hw >> 3;
  const uint32_t g = fdiv(i, a.cdiv_grp);
  const uint32_t col = 8u * g + (hw & 7u);
  return col < a.col_ncol ? col * a.col_grp + (i - g * a.col_grp) : 0xffffffffu;
}

// Logical site number -> coordinates for the traversal orders
//   (t_out, z_out, t_in, z_in, y, x/2)   blocks of t_block x z_block slices, or
//   (t_out, z_out, y, t_in, z_in, x/2)   the same blocks streamed along y, so
// that an XCD's L2 holds the three y planes of a block (with their t and z
// halo) that the neighbour loads of one plane touch.
__device__ __forceinline__ void decode_site(uint32_t sl, const DslashArgs &a, uint32_t &xh, uint32_t &y, uint32_t &z,
                                            uint32_t &t)
{
  if (a.col_order) {
    const uint32_t q = fdiv(sl, a.cdiv_slab);  // column * Lt + t
    const uint32_t b = sl - q * a.col_slab;
    const uint32_t colx = fdiv(q, a.cdiv_nT);
    t = q - colx * a.col_nT;
    const uint32_t col = fdiv(colx, a.cdiv_nx);
    const uint32_t r = fdiv(b, a.cdiv_sx);
    xh = (colx - col * a.col_nx) * a.col_sx + (b - r * a.col_sx);
    const uint32_t zin = fdiv(r, a.cdiv_cy);
    const uint32_t yin = r - zin * a.col_cy;
    const uint32_t cz = fdiv(col, a.cdiv_ncy);
    y = (col - cz * a.col_ncy) * a.col_cy + yin;
    z = cz * a.col_cz + zin;
    return;
  }
  if (a.ystream) {
    const uint32_t q = fdiv(sl, a.div_subLx);
    xh = sl - q * a.subLx;
    const uint32_t q2 = fdiv(q, a.div_zb);
    const uint32_t zin = q - q2 * a.zblock;
    const uint32_t q3 = fdiv(q2, a.div_tb);
    const uint32_t tin = q2 - q3 * a.tblock;
    const uint32_t q4 = fdiv(q3, a.div_Ly);
    y = q3 - q4 * a.Ly;
    const uint32_t tout = fdiv(q4, a.div_nzb);
    const uint32_t zout = q4 - tout * (a.Lz / a.zblock);
    z = zout * a.zblock + zin;
    t = tout * a.tblock + tin;
    return;
  }
  const uint32_t row = uint32_t(a.subLx) * a.Ly;
  const uint32_t q = fdiv(sl, a.div_row);
  const uint32_t r = sl - q * row;
  y = fdiv(r, a.div_subLx);
  xh = r - y * a.subLx;
  const uint32_t q2 = fdiv(q, a.div_zb);
  const uint32_t zin = q - q2 * a.zblock;
  const uint32_t q3 = fdiv(q2, a.div_tb);
  const uint32_t tin = q2 - q3 * a.tblock;
  const uint32_t tout = fdiv(q3, a.div_nzb);
  const uint32_t zout = q3 - tout * (a.Lz / a.zblock);
  z = zout * a.zblock + zin;
  t = tout * a.tblock + tin;
}

__device__ __forceinline__ SiteGeo site_geometry(uint32_t sl, const DslashArgs &a)
{
  SiteGeo g;
  const uint32_t row = uint32_t(a.subLx) * a.Ly;
  uint32_t xh, y, z, t;
  decode_site(sl, a, xh, y, z, t);
  const int idx = int(((t * a.Lz + z) * a.Ly + y) * a.subLx + xh);
  g.idx = idx;
  const int eo = int((y + z + t) & 1u);
  const int p = a.parity;
  // x: the checkerboard x index moves only when the full x coordinate crosses a pair.
  const int xf = (eo != p) ? (int(xh) + 1 == a.subLx ? 0 : int(xh) + 1) : int(xh);
  const int xb = (eo == p) ? (xh == 0 ? a.subLx - 1 : int(xh) - 1) : int(xh);
  g.nbr[0] = idx - int(xh) + xf;
  g.nbr[1] = idx - int(xh) + xb;
  const int sy = a.subLx, sz = int(row), st = int(row) * a.Lz;
  g.nbr[2] = idx + (int(y) + 1 == a.Ly ? -(a.Ly - 1) * sy : sy);
  g.nbr[3] = idx + (y == 0 ? (a.Ly - 1) * sy : -sy);
  g.nbr[4] = idx + (int(z) + 1 == a.Lz ? -(a.Lz - 1) * sz : sz);
  g.nbr[5] = idx + (z == 0 ? (a.Lz - 1) * sz : -sz);
  g.nbr[6] = idx + (int(t) + 1 == a.Lt ? -(a.Lt - 1) * st : st);
  g.nbr[7] = idx + (t == 0 ? (a.Lt - 1) * st : -st);
  g.xh = int(xh);
  g.live[0] = !(a.skip[0] && int(xh) == a.subLx - 1 && p != eo);
  g.live[1] = !(a.skip[0] && xh == 0 && p == eo);
  g.live[2] = !(a.skip[1] && int(y) == a.Ly - 1);
  g.live[3] = !(a.skip[1] && y == 0);
  g.live[4] = !(a.skip[2] && int(z) == a.Lz - 1);
  g.live[5] = !(a.skip[2] && z == 0);
  g.live[6] = !(a.skip[3] && int(t) == a.Lt - 1);
  g.live[7] = !(a.skip[3] && t == 0);
  g.tneg[0] = a.t_flip && int(t) == a.Lt - 1;
  g.tneg[1] = a.t_flip && t == 0;
  return g;
}

// Load a neighbour spinor (all 4N components) for R right-hand sides.
template <int N, int R, int SL, class CS, class A>
__device__ __forceinline__ void load_spinor(cpx<A> (&psi)[R][4][N], const DslashArgs &a, const void *base, int site,
                                            int slot)
{
  if constexpr (SL == SL_PACKED) {
    const CVec<CS, R> *p = reinterpret_cast<const CVec<CS, R> *>(static_cast<const CS *>(base) +
                                                                  size_t(site) * 4 * N * a.k + size_t(slot) * R);
    const size_t stride = size_t(a.k) / R;  // in CVec units between consecutive (spin,colour)
#pragma unroll
    for (int sc = 0; sc < 4 * N; ++sc) {
      CVec<CS, R> v = p[sc * stride];
#pragma unroll
      for (int r = 0; r < R; ++r) psi[r][sc / N][sc % N] = ld_cvt<A>(v.v[r]);
    }
  } else if constexpr (SL == SL_AOS_PTRS) {
    static_assert(R == 1, "pointer layouts hold one RHS per lane");
    const CS *p = static_cast<const CS *>(a.in_ptrs.p[slot]) + size_t(site) * 4 * N;
    constexpr int C = 16 / int(sizeof(CS));  // complex numbers per 16-byte load
    if constexpr (C > 1) {
      if (a.aos_vec) {  // every buffer 16-byte aligned: the site's 4N components in 16-byte pieces
#pragma unroll
        for (int q = 0; q < 4 * N / C; ++q) {
          const CVec<CS, C> v = reinterpret_cast<const CVec<CS, C> *>(p)[q];
#pragma unroll
          for (int j = 0; j < C; ++j) psi[0][(q * C + j) / N][(q * C + j) % N] = ld_cvt<A>(v.v[j]);
        }
        return;
      }
    }
#pragma unroll
    for (int sc = 0; sc < 4 * N; ++sc) psi[0][sc / N][sc % N] = ld_cvt<A>(p[sc]);
  } else {
    static_assert(R == 1, "pointer layouts hold one RHS per lane");
    const CS *p = static_cast<const CS *>(a.in_ptrs.p[slot]) + site;
#pragma unroll
    for (int sc = 0; sc < 4 * N; ++sc) psi[0][sc / N][sc % N] = ld_cvt<A>(p[size_t(sc) * a.S]);
  }
}

// Load the link of direction D for the output site (native layout: from a
// site-contiguous block of 8 links, in global memory or staged in LDS).
template <int N, class CG, class A>
__device__ __forceinline__ void load_link_native(cpx<A> (&V)[N][N], const CG *blk8, int D)
{
  const CG *p = blk8 + D * N * N;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) V[i][j] = ld_cvt<A>(p[i * N + j]);
}

// Compressed SU(3) link: rows 0 and 1 stored, row 2 = conj(row 0 x row 1)
// (unitarity and det = 1); neg flips the sign of the whole link.
template <int N, class CG, class A>
__device__ __forceinline__ void load_link_native12(cpx<A> (&V)[N][N], const CG *blk8, int slot, bool neg)
{
  static_assert(N == 3, "two-row links are SU(3) only");
  const CG *p = blk8 + slot * 2 * N;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) V[i][j] = ld_cvt<A>(p[i * N + j]);
  // (a x b)_j = a_{j+1} b_{j+2} - a_{j+2} b_{j+1}; row 2 is its conjugate.
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const cpx<A> a1 = V[0][(j + 1) % 3], a2 = V[0][(j + 2) % 3], b1 = V[1][(j + 1) % 3], b2 = V[1][(j + 2) % 3];
    const A re = fmaA(a1.re, b2.re, fmaA(-a1.im, b2.im, fmaA(-a2.re, b1.re, a2.im * b1.im)));
    const A im = fmaA(a1.re, b2.im, fmaA(a1.im, b2.re, fmaA(-a2.re, b1.im, -(a2.im * b1.re))));
    V[2][j] = cpx<A>{re, -im};
  }
  if (neg) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) V[i][j] = cpx<A>{-V[i][j].re, -V[i][j].im};
  }
}

// Eight-real SU(3) link (slot `slot` of a site block): a = row 0, b = row 1,
// c = row 2 with a1, a2, b0 stored and the phases of a0 and c0. The moduli
// follow from unit rows and columns, |a0|^2 = 1 - n with n = |a1|^2 + |a2|^2
// and |c0|^2 = n - |b0|^2, and the rest from unitarity with det = 1:
//   b1 = -(c0* a2* + a0* a1 b0) / n    b2 = (c0* a1* - a0* a2 b0) / n
//   c1 = (b0* a2* - a0* a1 c0) / n     c2 = -(b0* a1* + a0* a2 c0) / n
// (n > 0 is checked when the field is packed); neg flips the whole link.
template <class A> __device__ __forceinline__ cpx<A> cmul(cpx<A> x, cpx<A> y)
{
  return cpx<A>{fmaA(x.re, y.re, -(x.im * y.im)), fmaA(x.re, y.im, x.im * y.re)};
}
template <class A> __device__ __forceinline__ cpx<A> cconj(cpx<A> x) { return cpx<A>{x.re, -x.im}; }
template <class A> __device__ __forceinline__ void sin_cos(A x, A *sn, A *cs);
template <> __device__ __forceinline__ void sin_cos<float>(float x, float *sn, float *cs) { sincosf(x, sn, cs); }
template <> __device__ __forceinline__ void sin_cos<double>(double x, double *sn, double *cs) { sincos(x, sn, cs); }

template <int N, class CG, class A>
__device__ __forceinline__ void load_link_native8(cpx<A> (&V)[N][N], const CG *blk8, int slot, bool neg)
{
  static_assert(N == 3, "eight-real links are SU(3) only");
  const CG *p = blk8 + slot * 4;
  const cpx<A> a1 = ld_cvt<A>(p[0]), a2 = ld_cvt<A>(p[1]), b0 = ld_cvt<A>(p[2]), th = ld_cvt<A>(p[3]);
  const A n = fmaA(a1.re, a1.re, fmaA(a1.im, a1.im, fmaA(a2.re, a2.re, a2.im * a2.im)));
  const A nb0 = fmaA(b0.re, b0.re, b0.im * b0.im);
  const A ra0 = sqrt(A(1) - n > A(0) ? A(1) - n : A(0));
  const A rc0 = sqrt(n - nb0 > A(0) ? n - nb0 : A(0));
  A sa, ca, sc, cc;
  sin_cos<A>(th.re, &sa, &ca);
  sin_cos<A>(th.im, &sc, &cc);
  const cpx<A> a0{ra0 * ca, ra0 * sa}, c0{rc0 * cc, rc0 * sc};
  const A inv = A(1) / n;
  const cpx<A> a0c = cconj(a0), b0c = cconj(b0), c0c = cconj(c0), a1c = cconj(a1), a2c = cconj(a2);
  const cpx<A> t1 = cmul(c0c, a2c), t2 = cmul(a0c, cmul(a1, b0)), t3 = cmul(c0c, a1c), t4 = cmul(a0c, cmul(a2, b0));
  const cpx<A> t5 = cmul(b0c, a2c), t6 = cmul(a0c, cmul(a1, c0)), t7 = cmul(b0c, a1c), t8 = cmul(a0c, cmul(a2, c0));
  V[0][0] = a0;
  V[0][1] = a1;
  V[0][2] = a2;
  V[1][0] = b0;
  V[1][1] = cpx<A>{-(t1.re + t2.re) * inv, -(t1.im + t2.im) * inv};
  V[1][2] = cpx<A>{(t3.re - t4.re) * inv, (t3.im - t4.im) * inv};
  V[2][0] = c0;
  V[2][1] = cpx<A>{(t5.re - t6.re) * inv, (t5.im - t6.im) * inv};
  V[2][2] = cpx<A>{-(t7.re + t8.re) * inv, -(t7.im + t8.im) * inv};
  if (neg) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) V[i][j] = cpx<A>{-V[i][j].re, -V[i][j].im};
  }
}

// Reference SoA gauge [4][2][N*N][S]: forward U_mu(x) at x's parity, backward
// U_mu(x-mu) at the neighbour (opposite parity), conjugate-transposed here.
template <int N, class CG, class A>
__device__ __forceinline__ void load_link_refsoa(cpx<A> (&V)[N][N], const DslashArgs &a, int D, int idx, int nbr)
{
  const int mu = D >> 1;
  const bool bwd = D & 1;
  const int par = bwd ? 1 - a.parity : a.parity;
  const int site = bwd ? nbr : idx;
  const CG *p = static_cast<const CG *>(a.gauge) + size_t(mu * 2 + par) * N * N * a.S + site;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cpx<A> u = ld_cvt<A>(p[size_t(i * N + j) * a.S]);
      if (bwd) V[j][i] = cpx<A>{u.re, -u.im};
      else V[i][j] = u;
    }
}

// Link of direction D for the output site (native layout: from the site's
// 8-link block, in global memory or staged in LDS; reference layout: gathered).
template <int D, int N, int GL, class CG, class A>
__device__ __forceinline__ void fetch_link(cpx<A> (&V)[N][N], const DslashArgs &a, const SiteGeo &g, const CG *blk8)
{
  if constexpr (GL == GL_NATIVE) load_link_native<N, CG, A>(V, blk8, native_link_slot(D));
  else if constexpr (GL == GL_NATIVE12)
    load_link_native12<N, CG, A>(V, blk8, native_link_slot(D), D >= 6 ? g.tneg[D - 6] != 0 : false);
  else load_link_refsoa<N, CG, A>(V, a, D, g.idx, g.nbr[D]);
}

// Software-pipelined sweep over the 8 directions: the neighbour spinor (and
// link) of direction D+1 are requested before the arithmetic of direction D,
// so each wave keeps one direction of loads in flight while it computes.
// Every neighbour index is in range (periodic wrap), so the prefetch is
// unconditional; a skipped boundary term only skips the arithmetic.
// LL: links are staged in LDS; they are then read right before use (LDS
// latency is short) instead of being prefetched into a second register set.
template <int D, int N, int R, int SL, int GL, class CS, class CG, class A, bool LL>
__device__ __forceinline__ void sweep(cpx<A> (&acc)[R][4][N], const DslashArgs &a, const SiteGeo &g, int slot,
                                      const CG *blk8, A s, cpx<A> (&psi_cur)[R][4][N], cpx<A> (&V_cur)[N][N],
                                      cpx<A> (&psi_nxt)[R][4][N], cpx<A> (&V_nxt)[N][N])
{
  constexpr bool ROWS = GL == GL_NATIVE && N >= kRowLinkN;
  if constexpr (D == 0) {
    // X- data: every lane reads its predecessor's X+ spinor, then the lanes
    // without a predecessor in the row overwrite it with their own load.
    if constexpr (!LL && !ROWS) fetch_link<1, N, GL, CG, A>(V_nxt, a, g, blk8);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int sp = 0; sp < 4; ++sp)
#pragma unroll
        for (int c = 0; c < N; ++c)
          psi_nxt[r][sp][c] = cpx<A>{lane_read(psi_cur[r][sp][c].re, g.xsrc), lane_read(psi_cur[r][sp][c].im, g.xsrc)};
    if (g.xown) load_spinor<N, R, SL, CS, A>(psi_nxt, a, a.in, g.nbr[1], slot);
  } else if constexpr (D + 1 < 8) {
    load_spinor<N, R, SL, CS, A>(psi_nxt, a, a.in, g.nbr[D + 1], slot);
    if constexpr (!LL && !ROWS) fetch_link<D + 1, N, GL, CG, A>(V_nxt, a, g, blk8);
  }
  if (g.live[D]) {
    if constexpr (ROWS) {
      hop_rows<N, D, R, A, CG>(acc, blk8 + native_link_slot(D) * N * N, psi_cur, s);
    } else {
      if constexpr (LL) fetch_link<D, N, GL, CG, A>(V_cur, a, g, blk8);
#pragma unroll
      for (int r = 0; r < R; ++r) hop_one<N, D, A>(acc[r], V_cur, psi_cur[r], s);
    }
  }
  if constexpr (D + 1 < 8) {
    if constexpr (LL)
      sweep<D + 1, N, R, SL, GL, CS, CG, A, LL>(acc, a, g, slot, blk8, s, psi_nxt, V_cur, psi_cur, V_nxt);
    else
      sweep<D + 1, N, R, SL, GL, CS, CG, A, LL>(acc, a, g, slot, blk8, s, psi_nxt, V_nxt, psi_cur, V_cur);
  }
}

// ---------------------------------------------------------------------------
// User AoS pointers read by site groups (SL_AOS_COOP). The P = k lanes of a
// site (one RHS each; P a power of two, so a group never straddles a
// wavefront) load the P records of a neighbour site together: lane `slot`
// loads the 16-byte pieces m = slot + P q, q < Q (Q pieces per record), i.e.
// piece m % Q of RHS m / Q's record. One load instruction of the wavefront
// then covers whole records of a few buffers at the wavefront's consecutive
// sites (~10 cache lines for 16 FP32 RHS) where a lane reading its own record
// touches ~48 (every lane a different buffer, 16 bytes of each 96-byte
// record). The pieces are exchanged through the group's LDS area of P Q
// pieces (each record in a slot of QS >= Q pieces) and every lane reads its own
// record. Every lane of a group takes part
// in each exchange (the group's branch conditions are uniform over it).
template <int N, class CS>
struct AosRec {
  static constexpr int Q = 4 * N * int(sizeof(CS)) / 16;  // 16-byte pieces per record
  static constexpr int C = 16 / int(sizeof(CS));          // complex numbers per piece
  // Pieces per record slot of the LDS exchange area: odd, so that lanes
  // reading their own records at a stride of QS pieces spread over the banks.
  static constexpr int QS = Q | 1;
  static_assert((4 * N * int(sizeof(CS))) % 16 == 0, "records in whole 16-byte pieces");
};
// The Q pieces a lane holds between issue and exchange.
template <int N, class CS>
using AosPieces = u32x4[AosRec<N, CS>::Q];

// The base of each piece this lane loads: piece m = slot + P q is piece
// m % Q of RHS m / Q's record (read from the argument block once, not once
// per direction).
template <int N, class CS>
struct AosBases {
  const u32x4 *p[AosRec<N, CS>::Q];
};
template <int N, class CS>
__device__ __forceinline__ AosBases<N, CS> coop_bases(const DslashArgs &a, int slot)
{
  constexpr int Q = AosRec<N, CS>::Q;
  AosBases<N, CS> b;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = slot + a.P * q;
    const int r = m / Q;
    b.p[q] = static_cast<const u32x4 *>(a.in_ptrs.p[r]) + (m - r * Q);
  }
  return b;
}

// Issue the group's loads of the neighbour record at `site` (no wait).
template <int N, class CS>
__device__ __forceinline__ void coop_issue(AosPieces<N, CS> &v, const AosBases<N, CS> &b, int site)
{
  constexpr int Q = AosRec<N, CS>::Q;
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = b.p[q][size_t(site) * Q];
}

// Exchange the issued pieces through the group's LDS area `grp` and convert
// this lane's record into psi.
template <int N, class CS, class A>
__device__ __forceinline__ void coop_finish(cpx<A> (&psi)[1][4][N], const AosPieces<N, CS> &v, u32x4 *grp,
                                            int slot, int P)
{
  constexpr int Q = AosRec<N, CS>::Q, C = AosRec<N, CS>::C, QS = AosRec<N, CS>::QS;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = slot + P * q;
    const int r = m / Q;
    grp[r * QS + (m - r * Q)] = v[q];
  }
  // The group is inside one wavefront: LDS accesses of a wavefront complete
  // in order, so only the compiler's order needs holding.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const CVec<CS, C> w = __builtin_bit_cast(CVec<CS, C>, grp[slot * QS + q]);
#pragma unroll
    for (int j = 0; j < C; ++j) psi[0][(q * C + j) / N][(q * C + j) % N] = ld_cvt<A>(w.v[j]);
  }
  __builtin_amdgcn_wave_barrier();  // the next exchange's stores stay after these loads
}

// Output of the group: each lane's record (times fin) into the LDS area,
// then lane `slot` stores pieces m = slot + P q to RHS m / Q at `site`.
template <int N, class CO, class A>
__device__ __forceinline__ void coop_store(const DslashArgs &a, const cpx<A> (&acc)[1][4][N], int site, int slot,
                                           u32x4 *grp, A fin, A &nrm)
{
  constexpr int Q = AosRec<N, CO>::Q, C = AosRec<N, CO>::C, QS = AosRec<N, CO>::QS;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    CVec<CO, C> o;
#pragma unroll
    for (int j = 0; j < C; ++j) {
      cpx<A> v = acc[0][(q * C + j) / N][(q * C + j) % N];
      v.re *= fin;
      v.im *= fin;
      nrm += v.re * v.re + v.im * v.im;
      o.v[j] = StCvt<CO>::go(v);
    }
    grp[slot * QS + q] = __builtin_bit_cast(u32x4, o);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = slot + a.P * q;
    const int r = m / Q;
    static_cast<u32x4 *>(a.out_ptrs.p[r])[size_t(site) * Q + (m - r * Q)] = grp[r * QS + (m - r * Q)];
  }
}

// The sweep of the group-read kernel: direction D + 1's loads are issued
// before direction D's arithmetic and exchanged after it. X- comes from the
// lane holding the same RHS of the site before in the row (its X+), loaded
// by the groups at the start of a row (g.xown, uniform over a group). Links
// from LDS.
template <int D, int N, class CS, class CG, class A>
__device__ __forceinline__ void sweep_coop(cpx<A> (&acc)[1][4][N], const DslashArgs &a, const SiteGeo &g, int slot,
                                           const CG *blk8, A s, cpx<A> (&psi_cur)[1][4][N],
                                           cpx<A> (&psi_nxt)[1][4][N], u32x4 *grp, const AosBases<N, CS> &bp)
{
  AosPieces<N, CS> v;
  if constexpr (D == 0) {
#pragma unroll
    for (int sp = 0; sp < 4; ++sp)
#pragma unroll
      for (int c = 0; c < N; ++c)
        psi_nxt[0][sp][c] = cpx<A>{lane_read(psi_cur[0][sp][c].re, g.xsrc), lane_read(psi_cur[0][sp][c].im, g.xsrc)};
    if (g.xown) coop_issue<N, CS>(v, bp, g.nbr[1]);
  } else if constexpr (D + 1 < 8) {
    coop_issue<N, CS>(v, bp, g.nbr[D + 1]);
  }
  if (g.live[D]) {
    cpx<A> V[N][N];
    fetch_link<D, N, GL_NATIVE, CG, A>(V, a, g, blk8);
    hop_one<N, D, A>(acc[0], V, psi_cur[0], s);
  }
  if constexpr (D + 1 < 8) {
    if constexpr (D == 0) {
      if (g.xown) coop_finish<N, CS, A>(psi_nxt, v, grp, slot, a.P);
    } else {
      coop_finish<N, CS, A>(psi_nxt, v, grp, slot, a.P);
    }
    sweep_coop<D + 1, N, CS, CG, A>(acc, a, g, slot, blk8, s, psi_nxt, psi_cur, grp, bp);
  }
}

// ---------------------------------------------------------------------------
// 16-bit storage paths (binary16 and bfloat16). Spinor and link components
// stay packed (re, im) pairs of 16-bit values and every complex
// multiply-accumulate of the mat-vec is two dot2 instructions
// (v_dot2_f32_f16 / v_dot2_f32_bf16: d = s0.lo*s1.lo + s0.hi*s1.hi + s2,
// products and sum in FP32):
//   Im(V h) = dot2(V, hs),  Re(V h) = dot2(Vc, hs),
// with hs = (Im h, Re h) and Vc = (-Im V, Re V). The projection produces hs
// directly and Vc is formed once per link, shared by the R right-hand sides
// of the lane. binary16 projects with one packed-half FMA per colour (the
// swap folded into its operand selection); bfloat16 has no packed arithmetic,
// so it projects in FP32 and rounds the swapped pair back with one
// v_cvt_pk_bf16_f32 (round to nearest even).
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

template <class C> struct Pair;

template <> struct Pair<c16> {
  using V = h2;
  __device__ __forceinline__ static V load(const c16 &c)
  {
    V r;
    r.x = c.x;
    r.y = c.y;
    return r;
  }
  __device__ __forceinline__ static V swap(V h) { return __builtin_shufflevector(h, h, 1, 0); }
  // swap(base + (phase * s) * z) for the phase codes of add_phase.
  template <int KP>
  __device__ __forceinline__ static V proj_swapped(V base, V z, float s)
  {
    const V ss = {_Float16(s), _Float16(s)};
    if constexpr (KP == 0) return __builtin_elementwise_fma(ss, swap(z), swap(base));
    else if constexpr (KP == 1) return __builtin_elementwise_fma(-ss, swap(z), swap(base));
    else if constexpr (KP == 2) {  // i s z = s (-Im z, Re z); swapped: (s Re z, -s Im z)
      const V sg = {ss.x, _Float16(-ss.y)};
      return __builtin_elementwise_fma(sg, z, swap(base));
    } else {  // -i s z = s (Im z, -Re z); swapped: (-s Re z, s Im z)
      const V sg = {_Float16(-ss.x), ss.y};
      return __builtin_elementwise_fma(sg, z, swap(base));
    }
  }
  __device__ __forceinline__ static V link_c(V v)  // (-Im V, Re V)
  {
    const V m = {_Float16(-1.0f), _Float16(1.0f)};
    return swap(v) * m;
  }
  __device__ __forceinline__ static float dot(V a, V b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }
  // a.x b.x + a.y b.y from zero. Written out so that it stays the
  // three-operand form with an inline-constant 0 accumulator: the compiler
  // otherwise turns dot(a, b, 0) into the two-operand v_dot2c after a
  // v_mov of the zero, one extra instruction per chain (384 per lane in the
  // 16-RHS FP16 hypercube kernel).
  __device__ __forceinline__ static float dot0(V a, V b)
  {
    float r;
    asm("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
  __device__ __forceinline__ static f2v to_f(V v) { return __builtin_convertvector(v, f2v); }
  __device__ __forceinline__ static V from_f(f2v v) { return __builtin_convertvector(v, V); }
  __device__ __forceinline__ static V neg(V v) { return __builtin_bit_cast(V, __builtin_bit_cast(uint32_t, v) ^ 0x80008000u); }
};

template <> struct Pair<cb16> {
  using V = bf2;
  __device__ __forceinline__ static V load(const cb16 &c) { return __builtin_bit_cast(V, c); }
  template <int KP>
  __device__ __forceinline__ static V proj_swapped(V base, V z, float s)
  {
    const f2v b = __builtin_convertvector(base, f2v), w = __builtin_convertvector(z, f2v);
    float hr, hi;
    if constexpr (KP == 0) {
      hr = fmaf(s, w.x, b.x);
      hi = fmaf(s, w.y, b.y);
    } else if constexpr (KP == 1) {
      hr = fmaf(-s, w.x, b.x);
      hi = fmaf(-s, w.y, b.y);
    } else if constexpr (KP == 2) {
      hr = fmaf(-s, w.y, b.x);
      hi = fmaf(s, w.x, b.y);
    } else {
      hr = fmaf(s, w.y, b.x);
      hi = fmaf(-s, w.x, b.y);
    }
    const f2v hs = {hi, hr};
    return __builtin_convertvector(hs, V);
  }
  __device__ __forceinline__ static V link_c(V v)  // (-Im V, Re V): swap the halves, flip the low sign bit
  {
    const uint32_t u = __builtin_bit_cast(uint32_t, v);
    return __builtin_bit_cast(V, ((u >> 16) | (u << 16)) ^ 0x8000u);
  }
  __device__ __forceinline__ static float dot(V a, V b, float c)
  {
    return __builtin_amdgcn_fdot2_f32_bf16(a, b, c, false);
  }
  // a.x b.x + a.y b.y from zero (see Pair<c16>::dot0).
  __device__ __forceinline__ static float dot0(V a, V b)
  {
    float r;
    asm("v_dot2_f32_bf16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
  __device__ __forceinline__ static f2v to_f(V v) { return __builtin_convertvector(v, f2v); }
  __device__ __forceinline__ static V from_f(f2v v) { return __builtin_convertvector(v, V); }
  __device__ __forceinline__ static V neg(V v) { return __builtin_bit_cast(V, __builtin_bit_cast(uint32_t, v) ^ 0x80008000u); }
};

// acc += hop in direction D for one RHS (16-bit storage).
template <class C, int N, int D>
__device__ __forceinline__ void hop_one_p(cpx<float> (&acc)[4][N], const typename Pair<C>::V (&V)[N][N],
                                          const typename Pair<C>::V (&Vc)[N][N],
                                          const typename Pair<C>::V (&psi)[4][N], float s)
{
  using P = Pair<C>;
  constexpr HopEntry e = hop_entry(D);
  typename P::V h0s[N], h1s[N];
#pragma unroll
  for (int c = 0; c < N; ++c) {
    h0s[c] = P::template proj_swapped<e.k0>(psi[0][c], psi[e.a0][c], s);
    h1s[c] = P::template proj_swapped<e.k1>(psi[1][c], psi[e.a1][c], s);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    cpx<float> g0, g1;
    g0.re = P::dot0(Vc[i][0], h0s[0]);
    g0.im = P::dot0(V[i][0], h0s[0]);
    g1.re = P::dot0(Vc[i][0], h1s[0]);
    g1.im = P::dot0(V[i][0], h1s[0]);
#pragma unroll
    for (int j = 1; j < N; ++j) {
      g0.re = P::dot(Vc[i][j], h0s[j], g0.re);
      g0.im = P::dot(V[i][j], h0s[j], g0.im);
      g1.re = P::dot(Vc[i][j], h1s[j], g1.re);
      g1.im = P::dot(V[i][j], h1s[j], g1.im);
    }
    acc[0][i].re += g0.re;
    acc[0][i].im += g0.im;
    acc[1][i].re += g1.re;
    acc[1][i].im += g1.im;
    acc[e.a0][i] = add_phase<ConjPhase<e.k0>::v>(acc[e.a0][i], g0, s);
    acc[e.a1][i] = add_phase<ConjPhase<e.k1>::v>(acc[e.a1][i], g1, s);
  }
}

template <class C, int N, int R>
__device__ __forceinline__ void load_spinor_p(typename Pair<C>::V (&psi)[R][4][N], const DslashArgs &a, int site,
                                              int slot)
{
  const CVec<C, R> *p = reinterpret_cast<const CVec<C, R> *>(static_cast<const C *>(a.in) +
                                                              size_t(site) * 4 * N * a.k + size_t(slot) * R);
  const size_t stride = size_t(a.k) / R;
#pragma unroll
  for (int sc = 0; sc < 4 * N; ++sc) {
    CVec<C, R> v = p[sc * stride];
#pragma unroll
    for (int r = 0; r < R; ++r) psi[r][sc / N][sc % N] = Pair<C>::load(v.v[r]);
  }
}

template <class C, int N>
__device__ __forceinline__ void load_link_p(typename Pair<C>::V (&V)[N][N], const C *blk8, int slot)
{
  const C *p = blk8 + slot * N * N;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) V[i][j] = Pair<C>::load(p[i * N + j]);
}

// Link of direction D as 16-bit pairs: native layout as stored; two-row
// SU(3) layout with row 2 = conj(row 0 x row 1) formed in FP32 and rounded
// back to the storage type (and the whole link negated where t_flip asks).
template <class C, int N, int GL>
__device__ __forceinline__ void load_link_pg(typename Pair<C>::V (&V)[N][N], const C *blk8, int D, const SiteGeo &g)
{
  if constexpr (GL == GL_NATIVE) {
    load_link_p<C, N>(V, blk8, native_link_slot(D));
  } else {
    static_assert(N == 3, "two-row links are SU(3) only");
    using P = Pair<C>;
    const C *p = blk8 + native_link_slot(D) * 2 * N;
    f2v f[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        V[i][j] = P::load(p[i * N + j]);
        f[i][j] = P::to_f(V[i][j]);
      }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const f2v a1 = f[0][(j + 1) % 3], a2 = f[0][(j + 2) % 3], b1 = f[1][(j + 1) % 3], b2 = f[1][(j + 2) % 3];
      const float re = fmaf(a1.x, b2.x, fmaf(-a1.y, b2.y, fmaf(-a2.x, b1.x, a2.y * b1.y)));
      const float im = fmaf(a1.x, b2.y, fmaf(a1.y, b2.x, fmaf(-a2.x, b1.y, -(a2.y * b1.x))));
      const f2v r2 = {re, -im};
      V[2][j] = P::from_f(r2);
    }
    if (D >= 6 && g.tneg[D >= 6 ? D - 6 : 0]) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) V[i][j] = P::neg(V[i][j]);
    }
  }
}

// 16-bit hop for all R right-hand sides of a lane with the link read one row
// at a time where it is used (large N; see hop_rows).
template <class C, int N, int D, int R>
__device__ __forceinline__ void hop_rows_p(cpx<float> (&acc)[R][4][N], const C *vp,
                                           const typename Pair<C>::V (&psi)[R][4][N], float s)
{
  using P = Pair<C>;
  constexpr HopEntry e = hop_entry(D);
  typename P::V h0s[R][N], h1s[R][N];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      h0s[r][c] = P::template proj_swapped<e.k0>(psi[r][0][c], psi[r][e.a0][c], s);
      h1s[r][c] = P::template proj_swapped<e.k1>(psi[r][1][c], psi[r][e.a1][c], s);
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    typename P::V v[N], vc[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      v[j] = P::load(vp[i * N + j]);
      vc[j] = P::link_c(v[j]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cpx<float> g0{0.f, 0.f}, g1{0.f, 0.f};
#pragma unroll
      for (int j = 0; j < N; ++j) {
        g0.re = P::dot(vc[j], h0s[r][j], g0.re);
        g0.im = P::dot(v[j], h0s[r][j], g0.im);
        g1.re = P::dot(vc[j], h1s[r][j], g1.re);
        g1.im = P::dot(v[j], h1s[r][j], g1.im);
      }
      acc[r][0][i].re += g0.re;
      acc[r][0][i].im += g0.im;
      acc[r][1][i].re += g1.re;
      acc[r][1][i].im += g1.im;
      acc[r][e.a0][i] = add_phase<ConjPhase<e.k0>::v>(acc[r][e.a0][i], g0, s);
      acc[r][e.a1][i] = add_phase<ConjPhase<e.k1>::v>(acc[r][e.a1][i], g1, s);
    }
  }
}

template <class C, int D, int N, int R, bool LL, int GL>
__device__ __forceinline__ void sweep_p(cpx<float> (&acc)[R][4][N], const DslashArgs &a, const SiteGeo &g, int slot,
                                        const C *blk8, float s, typename Pair<C>::V (&psi_cur)[R][4][N],
                                        typename Pair<C>::V (&V_cur)[N][N], typename Pair<C>::V (&psi_nxt)[R][4][N],
                                        typename Pair<C>::V (&V_nxt)[N][N])
{
  constexpr bool ROWS = GL == GL_NATIVE && N >= kRowLinkN;
  if constexpr (D == 0) {
    if constexpr (!LL && !ROWS) load_link_pg<C, N, GL>(V_nxt, blk8, 1, g);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int sp = 0; sp < 4; ++sp)
#pragma unroll
        for (int c = 0; c < N; ++c)
          psi_nxt[r][sp][c] = __builtin_bit_cast(
              typename Pair<C>::V, lane_read(__builtin_bit_cast(float, psi_cur[r][sp][c]), g.xsrc));
    if (g.xown) load_spinor_p<C, N, R>(psi_nxt, a, g.nbr[1], slot);
  } else if constexpr (D + 1 < 8) {
    load_spinor_p<C, N, R>(psi_nxt, a, g.nbr[D + 1], slot);
    if constexpr (!LL && !ROWS) load_link_pg<C, N, GL>(V_nxt, blk8, D + 1, g);
  }
  if (g.live[D]) {
    if constexpr (ROWS) {
      hop_rows_p<C, N, D, R>(acc, blk8 + native_link_slot(D) * N * N, psi_cur, s);
    } else {
      if constexpr (LL) load_link_pg<C, N, GL>(V_cur, blk8, D, g);
      typename Pair<C>::V Vc[N][N];
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) Vc[i][j] = Pair<C>::link_c(V_cur[i][j]);
#pragma unroll
      for (int r = 0; r < R; ++r) hop_one_p<C, N, D>(acc[r], V_cur, Vc, psi_cur[r], s);
    }
  }

  if constexpr (D + 1 < 8) {
    if constexpr (LL)
      sweep_p<C, D + 1, N, R, LL, GL>(acc, a, g, slot, blk8, s, psi_nxt, V_cur, psi_cur, V_nxt);
    else
      sweep_p<C, D + 1, N, R, LL, GL>(acc, a, g, slot, blk8, s, psi_nxt, V_nxt, psi_cur, V_cur);
  }
}

// Checkerboard index of a logical site number.
__device__ __forceinline__ uint32_t logical_to_phys(uint32_t sl, const DslashArgs &a)
{
  uint32_t xh, y, z, t;
  decode_site(sl, a, xh, y, z, t);
  return ((t * a.Lz + z) * a.Ly + y) * a.subLx + xh;
}

// Store / load of a small vector with non-temporal hints (16, 8 or 4 byte pieces).
template <class V>
__device__ __forceinline__ void store_nt(V *p, const V &v)
{
  static_assert(sizeof(V) % 4 == 0, "store size");
  if constexpr (sizeof(V) % 16 == 0) {
    const u32x4 *src = reinterpret_cast<const u32x4 *>(&v);
    u32x4 *dst = reinterpret_cast<u32x4 *>(p);
#pragma unroll
    for (int i = 0; i < int(sizeof(V) / 16); ++i) __builtin_nontemporal_store(src[i], dst + i);
  } else if constexpr (sizeof(V) % 8 == 0) {
    const u32x2 *src = reinterpret_cast<const u32x2 *>(&v);
    u32x2 *dst = reinterpret_cast<u32x2 *>(p);
#pragma unroll
    for (int i = 0; i < int(sizeof(V) / 8); ++i) __builtin_nontemporal_store(src[i], dst + i);
  } else {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&v);
    uint32_t *dst = reinterpret_cast<uint32_t *>(p);
#pragma unroll
    for (int i = 0; i < int(sizeof(V) / 4); ++i) __builtin_nontemporal_store(src[i], dst + i);
  }
}

// Epilogue of the packed layout for one lane (R right-hand sides of output
// site idx): true out = xa * x + xb * sc_in * acc, stored = true / sc_out,
// optional per-RHS |true out|^2 into nrm; or the fused CG update
// r -= alpha * true out with |r|^2 into nrm. ef: the per-RHS factors staged
// in LDS (null: read from global memory).
template <int N, int R, class CS, class CO, class A>
__device__ __forceinline__ void epilogue_packed(const DslashArgs &a, const cpx<A> (&acc)[R][4][N], int idx,
                                                uint32_t slot, const A *ef, A (&nrm)[R])
{
  A fin[R], fx[R], fo[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rhs = int(slot) * R + r;
    if (ef) {
      fin[r] = ef[rhs];
      fx[r] = ef[a.k + rhs];
      fo[r] = ef[2 * a.k + rhs];
    } else {
      fin[r] = A(a.xb) * (a.sc_in ? A(a.sc_in[rhs]) : A(1));
      fx[r] = A(a.xa) * (a.sc_x ? A(a.sc_x[rhs]) : A(1));
      fo[r] = a.sc_out ? A(1) / A(a.sc_out[rhs]) : A(1);
    }
  }
  if constexpr ((std::is_same<CO, c32>::value && std::is_same<A, float>::value) ||
                (std::is_same<CO, c64>::value && std::is_same<A, double>::value)) {
    if (a.rx_r) {
      const size_t base = size_t(idx) * 4 * N * a.k + size_t(slot) * R;
      const size_t stride = size_t(a.k) / R;
      CVec<CO, R> *pr = reinterpret_cast<CVec<CO, R> *>(static_cast<CO *>(a.rx_r) + base);
      const CVec<CS, R> *px =
          a.x ? reinterpret_cast<const CVec<CS, R> *>(static_cast<const CS *>(a.x) + base) : nullptr;
      A al[R];
#pragma unroll
      for (int r = 0; r < R; ++r) al[r] = ef ? ef[3 * a.k + int(slot) * R + r] : A(a.rx_alpha[int(slot) * R + r]);
#pragma unroll
      for (int sc = 0; sc < 4 * N; ++sc) {
        CVec<CS, R> xv;
        if (px) xv = px[sc * stride];
        CVec<CO, R> rv = pr[sc * stride];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          cpx<A> v = acc[r][sc / N][sc % N];
          v.re *= fin[r];
          v.im *= fin[r];
          if (px) {
            const cpx<A> xx = ld_cvt<A>(xv.v[r]);
            v.re += fx[r] * xx.re;
            v.im += fx[r] * xx.im;
          }
          rv.v[r].x -= al[r] * v.re;
          rv.v[r].y -= al[r] * v.im;
          nrm[r] += rv.v[r].x * rv.v[r].x + rv.v[r].y * rv.v[r].y;
        }
        pr[sc * stride] = rv;
      }
      return;
    }
  }
  const size_t base = size_t(idx) * 4 * N * a.k + size_t(slot) * R;
  const size_t stride = size_t(a.k) / R;
  CVec<CO, R> *po = reinterpret_cast<CVec<CO, R> *>(static_cast<CO *>(a.out) + base);
  const CVec<CS, R> *px = a.x ? reinterpret_cast<const CVec<CS, R> *>(static_cast<const CS *>(a.x) + base) : nullptr;
  if (a.plain_out) {
    // The epilogue runs after the wave's last load has returned, so its
    // instructions are not hidden behind memory: store acc directly.
#pragma unroll
    for (int sc = 0; sc < 4 * N; ++sc) {
      CVec<CO, R> o;
#pragma unroll
      for (int r = 0; r < R; ++r) o.v[r] = StCvt<CO>::go(acc[r][sc / N][sc % N]);
      if (a.nt_store) store_nt(po + sc * stride, o);
      else po[sc * stride] = o;
    }
  } else if (!px) {
    // Scaled output without an xpay operand: |fin acc|^2 = fin^2 |acc|^2
    // and stored = acc (fin / sc_out), two multiplies fewer per value.
    A gs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) gs[r] = fin[r] * fo[r];
#pragma unroll
    for (int sc = 0; sc < 4 * N; ++sc) {
      CVec<CO, R> o;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const cpx<A> v = acc[r][sc / N][sc % N];
        nrm[r] = fmaA(v.re, v.re, fmaA(v.im, v.im, nrm[r]));
        o.v[r] = StCvt<CO>::go(cpx<A>{v.re * gs[r], v.im * gs[r]});
      }
      if (a.nt_store) store_nt(po + sc * stride, o);
      else po[sc * stride] = o;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) nrm[r] *= fin[r] * fin[r];
  } else {
#pragma unroll
    for (int sc = 0; sc < 4 * N; ++sc) {
      CVec<CO, R> o;
      CVec<CS, R> xv;
      if (px) xv = px[sc * stride];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cpx<A> v = acc[r][sc / N][sc % N];
        v.re *= fin[r];
        v.im *= fin[r];
        if (px) {
          const cpx<A> xx = ld_cvt<A>(xv.v[r]);
          v.re += fx[r] * xx.re;
          v.im += fx[r] * xx.im;
        }
        nrm[r] += v.re * v.re + v.im * v.im;
        v.re *= fo[r];
        v.im *= fo[r];
        o.v[r] = StCvt<CO>::go(v);
      }
      if (a.nt_store) store_nt(po + sc * stride, o);
      else po[sc * stride] = o;
    }
  }
}

// Epilogue of the pointer layouts (one RHS per lane, RHS slot `slot`):
// out = xb sc_in acc (+ xa sc_x x with a per-RHS xpay operand), written to
// that RHS's buffer at site idx; |out|^2 into nrm.
template <int N, int SL, class CS, class CO, class A>
__device__ __forceinline__ void epilogue_ptrs(const DslashArgs &a, const cpx<A> (&acc)[1][4][N], int idx,
                                              uint32_t slot, A (&nrm)[1])
{
  const A fin[1] = {A(a.xb) * (a.sc_in ? A(a.sc_in[slot]) : A(1))};
  const A fx[1] = {A(a.xa) * (a.sc_x ? A(a.sc_x[slot]) : A(1))};
  constexpr int C = 16 / int(sizeof(CO));
  bool vec_done = false;
  if constexpr (SL == SL_AOS_PTRS && C > 1) {
    if (a.aos_vec && !a.xptr) {
      CVec<CO, C> *pv = reinterpret_cast<CVec<CO, C> *>(static_cast<CO *>(a.out_ptrs.p[slot]) +
                                                        size_t(idx) * 4 * N);
#pragma unroll
      for (int q = 0; q < 4 * N / C; ++q) {
        CVec<CO, C> o;
#pragma unroll
        for (int j = 0; j < C; ++j) {
          cpx<A> v = acc[0][(q * C + j) / N][(q * C + j) % N];
          v.re *= fin[0];
          v.im *= fin[0];
          nrm[0] += v.re * v.re + v.im * v.im;
          o.v[j] = StCvt<CO>::go(v);
        }
        pv[q] = o;
      }
      vec_done = true;
    }
  }
  if (!vec_done) {
    CO *po;
    const CS *px = nullptr;  // xpay operand of this RHS (same layout as the input)
    size_t es;
    if constexpr (SL == SL_AOS_PTRS) {
      po = static_cast<CO *>(a.out_ptrs.p[slot]) + size_t(idx) * 4 * N;
      if (a.xptr) px = static_cast<const CS *>(a.x_ptrs.p[slot]) + size_t(idx) * 4 * N;
      es = 1;
    } else {
      po = static_cast<CO *>(a.out_ptrs.p[slot]) + idx;
      if (a.xptr) px = static_cast<const CS *>(a.x_ptrs.p[slot]) + idx;
      es = size_t(a.S);
    }
#pragma unroll
    for (int sc = 0; sc < 4 * N; ++sc) {
      cpx<A> v = acc[0][sc / N][sc % N];
      v.re *= fin[0];
      v.im *= fin[0];
      if (px) {
        const cpx<A> xx = ld_cvt<A>(px[sc * es]);
        v.re = fmaA(fx[0], xx.re, v.re);
        v.im = fmaA(fx[0], xx.im, v.im);
      }
      nrm[0] += v.re * v.re + v.im * v.im;
      po[sc * es] = StCvt<CO>::go(v);
    }
  }
}

// Per-RHS partial sums of |out|^2 of one workgroup (logical index blk) into
// a.partial[blk][k] (no atomics on global memory: a second pass reduces the
// [nblocks][k] partials). Every lane of the workgroup calls it.
template <int R, class A>
__device__ __forceinline__ void workgroup_partials(const DslashArgs &a, uint32_t blk, bool active, uint32_t slot,
                                                   const A (&nrm)[R])
{
  if (!a.partial) return;
  __shared__ double red[256];
  __syncthreads();
  for (int i = threadIdx.x; i < a.k; i += blockDim.x) red[i] = 0.0;
  __syncthreads();
  if (active) {
#pragma unroll
    for (int r = 0; r < R; ++r) atomicAdd(&red[int(slot) * R + r], double(nrm[r]));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < a.k; i += blockDim.x) a.partial[size_t(blk) * a.k + i] = red[i];
}

// One kernel body for every layout / precision / tile combination.
//   CS: stored input (and xpay operand) complex, CO: stored output complex,
//   CG: stored link complex, A: accumulation real type.
//   SITE_MINOR: lanes enumerate sites fastest (reference SoA layout).
// hw_blk: the workgroup's index among the launch's a.nblocks workgroups in
// hardware dispatch order (blockIdx.x for a launch of its own; the offset
// index inside a batched work-list launch, whose ranges start at multiples of
// the XCD count so that hw_blk % 8 is still the XCD).
template <int N, int R, int SL, int GL, bool LDS, bool SITE_MINOR, class CS, class CG, class CO, class A>
__device__ __forceinline__ void dslash_body(const DslashArgs &a, uint32_t hw_blk)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t blk = a.col_order ? col_remap(hw_blk, a)
                                   : (a.xcd_gran ? xcd_cyclic(hw_blk, a.xcd_gran)
                                                 : (a.xcd_map ? xcd_remap(hw_blk, a.nblocks) : hw_blk));
  // Idle column-order workgroups (past the last column) leave before any
  // barrier; the whole workgroup takes this branch.
  if (blk >= a.nblocks) return;
  const uint32_t lane0 = blk * blockDim.x;
  const uint32_t gid = lane0 + threadIdx.x;
  const bool active = gid < a.nlanes;

  uint32_t sl, slot;
  if constexpr (SITE_MINOR) {
    slot = fdiv(gid, a.div_S);
    sl = gid - slot * a.S;
  } else {
    sl = fdiv(gid, a.div_P);
    slot = gid - sl * a.P;
  }

  const CG *blk8 = nullptr;
  // Two-row links staged in LDS are rebuilt to full links while staging, once
  // per site and link instead of once per lane that reads them: the sweep
  // then sees the native layout (GLS).
  static_assert(GL != GL_NATIVE8 || LDS, "eight-real links are rebuilt only while staging in LDS");
  constexpr int GLS = (LDS && (GL == GL_NATIVE12 || GL == GL_NATIVE8)) ? int(GL_NATIVE) : GL;
  constexpr int LC = link_cplx(GLS, N);  // complex numbers per link as the sweep reads them
  if constexpr (site_blocked(GL) && LDS) {
    static_assert(!SITE_MINOR, "LDS link staging needs site-major lanes");
    // Stage the 8 links of every site of this workgroup once; all k/R lanes of
    // a site then read them from LDS (broadcast) instead of from L1/L2.
    const uint32_t first = fdiv(lane0, a.div_P);
    uint32_t lastl = lane0 + blockDim.x - 1;
    if (lastl >= a.nlanes) lastl = a.nlanes - 1;
    const uint32_t nsite = fdiv(lastl, a.div_P) - first + 1;
    if constexpr (GL == GL_NATIVE) {
      constexpr int chunk16 = (8 * LC * int(sizeof(CG))) / 16;
      u32x4 *dst = reinterpret_cast<u32x4 *>(smem);
      const u32x4 *src = static_cast<const u32x4 *>(a.gauge);
      for (uint32_t i = threadIdx.x; i < nsite * chunk16; i += blockDim.x) {
        const uint32_t sloc = i / chunk16, w = i - sloc * chunk16;
        const uint32_t phys = logical_to_phys(first + sloc, a);
        const u32x4 *ps = src + (size_t(a.parity) * a.S + phys) * chunk16 + w;
        dst[i] = a.nt_link ? __builtin_nontemporal_load(ps) : *ps;
      }
    } else {
      // One thread per (site, link): read the stored reals, rebuild the full
      // link (and the t-boundary sign), store it in the native slot.
      using W = typename std::conditional<std::is_same<CG, c64>::value, double, float>::type;
      CG *dst = reinterpret_cast<CG *>(smem);
      for (uint32_t i = threadIdx.x; i < nsite * 8; i += blockDim.x) {
        const uint32_t sloc = i >> 3, d = i & 7u;
        const uint32_t phys = logical_to_phys(first + sloc, a);
        const CG *src = static_cast<const CG *>(a.gauge) + (size_t(a.parity) * a.S + phys) * 8 * link_cplx(GL, N);
        bool neg = false;
        if (a.t_flip && (d == 3 || d == 7)) {
          const uint32_t t = phys / (uint32_t(a.subLx) * a.Ly * a.Lz);
          neg = d == 3 ? int(t) == a.Lt - 1 : t == 0;
        }
        cpx<W> V[N][N];
        if constexpr (GL == GL_NATIVE12) load_link_native12<N, CG, W>(V, src, int(d), neg);  // d: slot in the block
        else load_link_native8<N, CG, W>(V, src, int(d), neg);
        CG *q = dst + (size_t(sloc) * 8 + d) * LC;
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
          for (int c = 0; c < N; ++c) q[r * N + c] = StCvt<CG>::go(V[r][c]);
      }
    }
    if (a.efac_off) {
      // The per-RHS factors of the epilogue, loaded here beside the links so
      // that their latency is not paid after the last neighbour load.
      A *ef = reinterpret_cast<A *>(smem + a.efac_off);
      for (int i = threadIdx.x; i < a.k; i += blockDim.x) {
        ef[i] = A(a.xb) * (a.sc_in ? A(a.sc_in[i]) : A(1));
        ef[a.k + i] = A(a.xa) * (a.sc_x ? A(a.sc_x[i]) : A(1));
        ef[2 * a.k + i] = a.sc_out ? A(1) / A(a.sc_out[i]) : A(1);
        ef[3 * a.k + i] = a.rx_alpha ? A(a.rx_alpha[i]) : A(0);
      }
    }
    __syncthreads();
    if (active) blk8 = reinterpret_cast<const CG *>(smem) + size_t(sl - first) * 8 * LC;
  }
  A nrm[R];
#pragma unroll
  for (int r = 0; r < R; ++r) nrm[r] = A(0);
  if (active) {
    SiteGeo g = site_geometry(sl, a);
    {
      // Lane gid - xstride holds the same RHS slot of logical site sl - 1,
      // which is x index xh - 1 of the same row (x is the fastest coordinate
      // of every traversal order) unless xh starts a row: xh = 0, or in a
      // column order that splits x, the first x index of the column.
      const int xstride = SITE_MINOR ? 1 : int(a.P);
      const int lw = int(threadIdx.x & 63u);
      const uint32_t xu = uint32_t(g.xh);
      const bool row_start = a.col_order ? xu == fdiv(xu, a.cdiv_sx) * a.col_sx : xu == 0;
      g.xown = (lw < xstride || row_start) ? 1 : 0;
      g.xsrc = g.xown ? lw : lw - xstride;
    }
    if constexpr (site_blocked(GL) && !LDS)
      blk8 = static_cast<const CG *>(a.gauge) + (size_t(a.parity) * a.S + g.idx) * 8 * LC;

    const A s = A(a.sign);
    cpx<A> acc[R][4][N];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int sp = 0; sp < 4; ++sp)
#pragma unroll
        for (int c = 0; c < N; ++c) acc[r][sp][c] = cpx<A>{A(-0.0), A(-0.0)};  // x + (-0) == x: folds

    u32x4 *grp = nullptr;  // SL_AOS_COOP: this site group's LDS exchange area
    if constexpr (SL == SL_AOS_COOP) {
      static_assert(GL == GL_NATIVE && LDS && !SITE_MINOR && R == 1 && N < kRowLinkN,
                    "group-read AoS kernels: native links staged in LDS, site-major lanes, one RHS per lane");
      grp = reinterpret_cast<u32x4 *>(smem + a.coop_off) + size_t(threadIdx.x - slot) * AosRec<N, CS>::QS;
      cpx<A> psi0[1][4][N], psi1[1][4][N];
      const AosBases<N, CS> bp = coop_bases<N, CS>(a, int(slot));
      {
        AosPieces<N, CS> v;
        coop_issue<N, CS>(v, bp, g.nbr[0]);
        coop_finish<N, CS, A>(psi0, v, grp, int(slot), a.P);
      }
      sweep_coop<0, N, CS, CG, A>(acc, a, g, int(slot), blk8, s, psi0, psi1, grp, bp);
    } else if constexpr (SL == SL_PACKED && site_blocked(GL) && std::is_same<CS, CG>::value &&
                  (std::is_same<CS, c16>::value || std::is_same<CS, cb16>::value) && std::is_same<A, float>::value) {
      using PV = typename Pair<CS>::V;
      PV psi0[R][4][N], psi1[R][4][N], V0[N][N], V1[N][N];
      load_spinor_p<CS, N, R>(psi0, a, g.nbr[0], slot);
      if constexpr (!LDS && !(GL == GL_NATIVE && N >= kRowLinkN)) load_link_pg<CS, N, GL>(V0, blk8, 0, g);
      sweep_p<CS, 0, N, R, LDS, GLS>(acc, a, g, slot, blk8, float(s), psi0, V0, psi1, V1);
    } else {
      cpx<A> psi0[R][4][N], psi1[R][4][N], V0[N][N], V1[N][N];
      load_spinor<N, R, SL, CS, A>(psi0, a, a.in, g.nbr[0], slot);
      constexpr bool LL = LDS && site_blocked(GL);
      if constexpr (!LL && !(GL == GL_NATIVE && N >= kRowLinkN)) fetch_link<0, N, GL, CG, A>(V0, a, g, blk8);
      sweep<0, N, R, SL, GLS, CS, CG, A, LL>(acc, a, g, slot, blk8, s, psi0, V0, psi1, V1);
    }

    // Epilogue: true out = xa * x + xb * sc_in * acc; stored = true / sc_out.
    const A *ef = nullptr;
    if constexpr (site_blocked(GL) && LDS)
      if (a.efac_off) ef = reinterpret_cast<const A *>(smem + a.efac_off);
    if constexpr (SL == SL_PACKED) {
      epilogue_packed<N, R, CS, CO, A>(a, acc, g.idx, slot, ef, nrm);
    } else if constexpr (SL == SL_AOS_COOP) {
      const A fin = A(a.xb) * (a.sc_in ? A(a.sc_in[slot]) : A(1));
      coop_store<N, CO, A>(a, acc, g.idx, int(slot), grp, fin, nrm[0]);
    } else {
      epilogue_ptrs<N, SL, CS, CO, A>(a, acc, g.idx, slot, nrm);
    }
  }

  workgroup_partials<R, A>(a, blk, active, slot, nrm);
}

// The kernel. KF > 0: the RHS count is the compile-time constant KF (packed
// layout only), so every row offset of a site block is an immediate and the
// per-row 64-bit address arithmetic disappears (24^3x48, 16 RHS, FP32:
// 343 -> 334 us; 3,007 -> 2,778 VALU instructions in the kernel, 182 -> 178
// VGPRs). The arithmetic and its order are unchanged, so results are
// bit-identical to KF = 0.
template <int N, int R, int SL, int GL, bool LDS, bool SITE_MINOR, class CS, class CG, class CO, class A, int KF = 0>
__global__ void __launch_bounds__(256) dslash_kernel(const DslashArgs a)
{
  static_assert(KF == 0 || SL == SL_PACKED, "a compile-time RHS count is for the packed layout");
  if constexpr (KF > 0) {
    DslashArgs b = a;
    b.k = KF;
    dslash_body<N, R, SL, GL, LDS, SITE_MINOR, CS, CG, CO, A>(b, blockIdx.x);
  } else {
    dslash_body<N, R, SL, GL, LDS, SITE_MINOR, CS, CG, CO, A>(a, blockIdx.x);
  }
}

}  // namespace qmrhs

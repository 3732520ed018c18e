// Hypercube work mapping of the packed multi-RHS dslash.
//
// The output sites of one parity are taken in 2x2x2x2 blocks of the full
// lattice: in checkerboard coordinates a block is one x index xh and the
// corners y = 2Y + a, z = 2Z + b, t = 2T + c (a, b, c in {0, 1}), eight
// output sites. The eight opposite-parity sites of the same block are their
// "inner" neighbours: every output site has exactly one inner neighbour per
// dimension (x: the other site of its x pair, which has the same checkerboard
// index; y, z, t: the site of the corner with that bit flipped, again at the
// same checkerboard index) and one "outer" neighbour per dimension outside
// the block. With P = k / R lanes per site, a block takes 8 P consecutive
// lanes, corner-major inside each wavefront: lane = 8 P block' + P corner +
// slot for P <= 8; for P > 8 the slots are split over P / 8 wavefronts of 8
// corners x 8 slots each. The corner partner in y, z, t is then the lane at
// distance Pl, 2 Pl, 4 Pl (Pl = min(P, 8)), always in the same wavefront.
//
// Each lane loads the spinor of its own inner site once and computes from it
// the half spinor its y, z and t partners need; the three half spinors cross
// lanes (ds_bpermute), so a block reads 8 inner and 32 outer spinors, 5 per
// output site instead of 8 (each input site is read once as inner and four
// times as outer). The forward and backward hops of a dimension differ only
// in the sign of the projector and reconstruction phases and in the link
// slot, so a lane's hop direction (forward or backward, fixed by its corner
// and, for x, by the parity of its row) is a per-lane sign: the hop of
// direction 2 mu + 1 is that of 2 mu with s -> -s (hop_entry's phase pairs).
//
// Operator and link layout are those of dslash_kernel.h (reference
// include/kernels/wilson_dslash_4d.cuh:59-293); the terms are summed in
// another order, so results agree with the site-major kernels to rounding.
#pragma once

#include "dslash_kernel.h"

namespace qmrhs {

// Block coordinates (xh, Y, Z, T) of logical block number hb, in the
// traversal order of dslash_kernel.h applied to the block lattice
// subLx x bLy x bLz x (Lt / 2) (tblock / zblock counted in blocks).
__device__ __forceinline__ void hc_decode(uint32_t hb, const DslashArgs &a, uint32_t &xh, uint32_t &Y, uint32_t &Z,
                                          uint32_t &T)
{
  if (a.col_order) {
    const uint32_t q = fdiv(hb, a.cdiv_slab);  // column * nT + T
    const uint32_t b = hb - q * a.col_slab;
    const uint32_t colx = fdiv(q, a.cdiv_nT);
    T = q - colx * a.col_nT;
    const uint32_t col = fdiv(colx, a.cdiv_nx);
    const uint32_t r = fdiv(b, a.cdiv_sx);
    xh = (colx - col * a.col_nx) * a.col_sx + (b - r * a.col_sx);
    const uint32_t zin = fdiv(r, a.cdiv_cy);
    const uint32_t yin = r - zin * a.col_cy;
    const uint32_t cz = fdiv(col, a.cdiv_ncy);
    Y = (col - cz * a.col_ncy) * a.col_cy + yin;
    Z = cz * a.col_cz + zin;
    return;
  }
  const uint32_t nzb = a.hdiv_nzb.d;
  if (a.ystream) {
    const uint32_t q = fdiv(hb, a.div_subLx);
    xh = hb - q * a.subLx;
    const uint32_t q2 = fdiv(q, a.div_zb);
    const uint32_t zin = q - q2 * a.zblock;
    const uint32_t q3 = fdiv(q2, a.div_tb);
    const uint32_t tin = q2 - q3 * a.tblock;
    const uint32_t q4 = fdiv(q3, a.hdiv_Ly);
    Y = q3 - q4 * a.bLy;
    const uint32_t tout = fdiv(q4, a.hdiv_nzb);
    const uint32_t zout = q4 - tout * nzb;
    Z = zout * a.zblock + zin;
    T = tout * a.tblock + tin;
    return;
  }
  const uint32_t row = uint32_t(a.subLx) * a.bLy;
  const uint32_t q = fdiv(hb, a.hdiv_row);
  const uint32_t r = hb - q * row;
  Y = fdiv(r, a.div_subLx);
  xh = r - Y * a.subLx;
  const uint32_t q2 = fdiv(q, a.div_zb);
  const uint32_t zin = q - q2 * a.zblock;
  const uint32_t q3 = fdiv(q2, a.div_tb);
  const uint32_t tin = q2 - q3 * a.tblock;
  const uint32_t tout = fdiv(q3, a.hdiv_nzb);
  const uint32_t zout = q3 - tout * nzb;
  Z = zout * a.zblock + zin;
  T = tout * a.tblock + tin;
}

struct HcSite {
  int idx, xh, y, z, t;
};

// Output site number sw = 8 * block + corner (corner bits: y, z, t).
__device__ __forceinline__ HcSite hc_site(uint32_t sw, const DslashArgs &a)
{
  uint32_t xh, Y, Z, T;
  hc_decode(sw >> 3, a, xh, Y, Z, T);
  const uint32_t c = sw & 7u;
  HcSite s;
  s.xh = int(xh);
  s.y = int(2 * Y + (c & 1u));
  s.z = int(2 * Z + ((c >> 1) & 1u));
  s.t = int(2 * T + (c >> 2));
  s.idx = ((s.t * a.Lz + s.z) * a.Ly + s.y) * a.subLx + s.xh;
  return s;
}

// Half spinor of the hop in dimension MU with sign s (s = +-sign: forward /
// backward): h0 = in_0 + k0 s in_a0, h1 = in_1 + k1 s in_a1.
template <int N, int MU, class A>
__device__ __forceinline__ void hc_proj(cpx<A> (&h)[2][N], const cpx<A> (&psi)[4][N], A s)
{
  constexpr HopEntry e = hop_entry(2 * MU);
#pragma unroll
  for (int c = 0; c < N; ++c) {
    h[0][c] = add_phase<e.k0>(psi[0][c], psi[e.a0][c], s);
    h[1][c] = add_phase<e.k1>(psi[1][c], psi[e.a1][c], s);
  }
}

// acc += reconstruction of V h (the half spinor of hc_proj with the same s).
template <int N, int MU, class A>
__device__ __forceinline__ void hc_apply(cpx<A> (&acc)[4][N], const cpx<A> (&V)[N][N], const cpx<A> (&h)[2][N], A s)
{
  constexpr HopEntry e = hop_entry(2 * MU);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    cpx<A> g0, g1;
    g0.re = V[i][0].re * h[0][0].re;
    g0.im = V[i][0].re * h[0][0].im;
    g1.re = V[i][0].re * h[1][0].re;
    g1.im = V[i][0].re * h[1][0].im;
    g0.re = fmaA(-V[i][0].im, h[0][0].im, g0.re);
    g0.im = fmaA(V[i][0].im, h[0][0].re, g0.im);
    g1.re = fmaA(-V[i][0].im, h[1][0].im, g1.re);
    g1.im = fmaA(V[i][0].im, h[1][0].re, g1.im);
#pragma unroll
    for (int j = 1; j < N; ++j) {
      const cpx<A> v = V[i][j];
      g0.re = fmaA(v.re, h[0][j].re, g0.re);
      g0.im = fmaA(v.re, h[0][j].im, g0.im);
      g1.re = fmaA(v.re, h[1][j].re, g1.re);
      g1.im = fmaA(v.re, h[1][j].im, g1.im);
      g0.re = fmaA(-v.im, h[0][j].im, g0.re);
      g0.im = fmaA(v.im, h[0][j].re, g0.im);
      g1.re = fmaA(-v.im, h[1][j].im, g1.re);
      g1.im = fmaA(v.im, h[1][j].re, g1.im);
    }
    acc[0][i].re += g0.re;
    acc[0][i].im += g0.im;
    acc[1][i].re += g1.re;
    acc[1][i].im += g1.im;
    acc[e.a0][i] = add_phase<ConjPhase<e.k0>::v>(acc[e.a0][i], g0, s);
    acc[e.a1][i] = add_phase<ConjPhase<e.k1>::v>(acc[e.a1][i], g1, s);
  }
}

// Full hop in dimension MU for the R right-hand sides of a lane: link from
// slot lslot of the site's LDS block, neighbour spinor psi, sign s.
template <int N, int MU, int R, class CG, class A>
__device__ __forceinline__ void hc_hop(cpx<A> (&acc)[R][4][N], const cpx<A> (&psi)[R][4][N], const CG *blk8, int lslot,
                                       A s)
{
  cpx<A> V[N][N];
  load_link_native<N, CG, A>(V, blk8, lslot);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    cpx<A> h[2][N];
    hc_proj<N, MU>(h, psi[r], s);
    hc_apply<N, MU>(acc[r], V, h, s);
  }
}

// Inner hop in dimension MU (1..3) through the corner partner at lane
// distance m: each lane projects its own inner spinor for the partner (whose
// sign is -s), the half spinors cross lanes, and each lane applies its link.
template <int N, int MU, int R, class CG, class A>
__device__ __forceinline__ void hc_hop_xchg(cpx<A> (&acc)[R][4][N], const cpx<A> (&psi)[R][4][N], const CG *blk8,
                                            int lslot, A s, int m)
{
  cpx<A> V[N][N];
  load_link_native<N, CG, A>(V, blk8, lslot);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    cpx<A> h[2][N];
    hc_proj<N, MU>(h, psi[r], -s);
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int c = 0; c < N; ++c) {
        h[q][c].re = __shfl_xor(h[q][c].re, m, 64);
        h[q][c].im = __shfl_xor(h[q][c].im, m, 64);
      }
    hc_apply<N, MU>(acc[r], V, h, s);
  }
}

// 16-bit storage (binary16 / bfloat16 pairs, FP32 accumulation with dot2;
// the scheme of hop_one_p): the projection yields swapped half spinors hs,
// which are what crosses lanes (one 32-bit value per complex number).
template <class C, int N, int MU>
__device__ __forceinline__ void hc_proj_p(typename Pair<C>::V (&hs)[2][N], const typename Pair<C>::V (&psi)[4][N],
                                          float s)
{
  constexpr HopEntry e = hop_entry(2 * MU);
#pragma unroll
  for (int c = 0; c < N; ++c) {
    hs[0][c] = Pair<C>::template proj_swapped<e.k0>(psi[0][c], psi[e.a0][c], s);
    hs[1][c] = Pair<C>::template proj_swapped<e.k1>(psi[1][c], psi[e.a1][c], s);
  }
}

template <class C, int N, int MU>
__device__ __forceinline__ void hc_apply_p(cpx<float> (&acc)[4][N], const typename Pair<C>::V (&V)[N][N],
                                           const typename Pair<C>::V (&Vc)[N][N],
                                           const typename Pair<C>::V (&hs)[2][N], float s)
{
  using P = Pair<C>;
  constexpr HopEntry e = hop_entry(2 * MU);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    cpx<float> g0, g1;
    g0.re = P::dot0(Vc[i][0], hs[0][0]);
    g0.im = P::dot0(V[i][0], hs[0][0]);
    g1.re = P::dot0(Vc[i][0], hs[1][0]);
    g1.im = P::dot0(V[i][0], hs[1][0]);
#pragma unroll
    for (int j = 1; j < N; ++j) {
      g0.re = P::dot(Vc[i][j], hs[0][j], g0.re);
      g0.im = P::dot(V[i][j], hs[0][j], g0.im);
      g1.re = P::dot(Vc[i][j], hs[1][j], g1.re);
      g1.im = P::dot(V[i][j], hs[1][j], g1.im);
    }
    acc[0][i].re += g0.re;
    acc[0][i].im += g0.im;
    acc[1][i].re += g1.re;
    acc[1][i].im += g1.im;
    acc[e.a0][i] = add_phase<ConjPhase<e.k0>::v>(acc[e.a0][i], g0, s);
    acc[e.a1][i] = add_phase<ConjPhase<e.k1>::v>(acc[e.a1][i], g1, s);
  }
}

template <class C, int N>
__device__ __forceinline__ void hc_link_p(typename Pair<C>::V (&V)[N][N], typename Pair<C>::V (&Vc)[N][N],
                                          const C *blk8, int lslot)
{
  load_link_p<C, N>(V, blk8, lslot);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) Vc[i][j] = Pair<C>::link_c(V[i][j]);
}

template <class C, int N, int MU, int R>
__device__ __forceinline__ void hc_hop_p(cpx<float> (&acc)[R][4][N], const typename Pair<C>::V (&psi)[R][4][N],
                                         const C *blk8, int lslot, float s)
{
  typename Pair<C>::V V[N][N], Vc[N][N];
  hc_link_p<C, N>(V, Vc, blk8, lslot);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    typename Pair<C>::V hs[2][N];
    hc_proj_p<C, N, MU>(hs, psi[r], s);
    hc_apply_p<C, N, MU>(acc[r], V, Vc, hs, s);
  }
}

template <class C, int N, int MU, int R>
__device__ __forceinline__ void hc_hop_xchg_p(cpx<float> (&acc)[R][4][N], const typename Pair<C>::V (&psi)[R][4][N],
                                              const C *blk8, int lslot, float s, int m)
{
  using PV = typename Pair<C>::V;
  PV V[N][N], Vc[N][N];
  hc_link_p<C, N>(V, Vc, blk8, lslot);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    PV hs[2][N];
    hc_proj_p<C, N, MU>(hs, psi[r], -s);
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int c = 0; c < N; ++c)
        hs[q][c] = __builtin_bit_cast(PV, __shfl_xor(__builtin_bit_cast(float, hs[q][c]), m, 64));
    hc_apply_p<C, N, MU>(acc[r], V, Vc, hs, s);
  }
}

// Kernel body. Packed spinors (or the user AoS layout through per-RHS
// pointers, one RHS per lane), native links staged in LDS, N <= 4 (links in
// registers), P = k / R lanes per site a power of two <= 32 with 8 P dividing
// the workgroup. The launch has S * P lanes; S is a multiple of 8 (even Ly,
// Lz, Lt), so every block is whole and the corner partners of an active lane
// are active.
template <int N, int R, int SL, class CS, class CG, class CO, class A>
__device__ __forceinline__ void dslash_hc_body(const DslashArgs &a, uint32_t hw_blk)
{
  static_assert(N < kRowLinkN, "hypercube kernels hold whole links in registers");
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
  // Lanes of one block: 8 P = 2^gb. With P <= 8 (Pl = P) the block is 8 P
  // consecutive lanes, corner-major; with P > 8 (Pl = 8) its RHS slots are
  // split over P / 8 wavefronts, each holding all 8 corners for 8 slots
  // (lane = 64 hi + 8 corner + lo, slot = 8 hi + lo), so the corner partners
  // stay inside one wavefront.
  const uint32_t lp = a.hc_lp;
  const uint32_t lpl = lp < 3u ? lp : 3u;
  const uint32_t gb = 3u + lp;
  const uint32_t rr = gid & ((1u << gb) - 1u);
  const uint32_t hi = rr >> (3u + lpl);
  const uint32_t r2 = rr & ((1u << (3u + lpl)) - 1u);
  const uint32_t cn = r2 >> lpl;
  const uint32_t slot = (hi << lpl) | (r2 & ((1u << lpl) - 1u));
  const uint32_t sw = ((gid >> gb) << 3) | cn;  // output site number: 8 * block + corner
  const uint32_t Pl = 1u << lpl;
  constexpr int LC = N * N;
  {
    // The 8 links of every output site of this workgroup, once, into LDS
    // (workgroups hold whole blocks: the block size divides the workgroup).
    const uint32_t first = lane0 >> lp;
    const uint32_t end = (lane0 + blockDim.x < a.nlanes) ? lane0 + blockDim.x : a.nlanes;
    const uint32_t nsite = (end - lane0) >> lp;
    constexpr int chunk16 = (8 * LC * int(sizeof(CG))) / 16;
    static_assert((8 * LC * int(sizeof(CG))) % 16 == 0, "site link block in 16-byte pieces");
    u32x4 *dst = reinterpret_cast<u32x4 *>(smem);
    const u32x4 *src = static_cast<const u32x4 *>(a.gauge);
    for (uint32_t i = threadIdx.x; i < nsite * chunk16; i += blockDim.x) {
      const uint32_t sloc = i / chunk16, w = i - sloc * chunk16;
      const uint32_t phys = uint32_t(hc_site(first + sloc, a).idx);
      const u32x4 *ps = src + (size_t(a.parity) * a.S + phys) * chunk16 + w;
      dst[i] = a.nt_link ? __builtin_nontemporal_load(ps) : *ps;
    }
    if (a.efac_off) {
      A *ef = reinterpret_cast<A *>(smem + a.efac_off);
      for (int i = threadIdx.x; i < a.k; i += blockDim.x) {
        ef[i] = A(a.xb) * (a.sc_in ? A(a.sc_in[i]) : A(1));
        ef[a.k + i] = A(a.xa) * (a.sc_x ? A(a.sc_x[i]) : A(1));
        ef[2 * a.k + i] = a.sc_out ? A(1) / A(a.sc_out[i]) : A(1);
        ef[3 * a.k + i] = a.rx_alpha ? A(a.rx_alpha[i]) : A(0);
      }
    }
    __syncthreads();
  }
  A nrm[R];
#pragma unroll
  for (int r = 0; r < R; ++r) nrm[r] = A(0);
  if (active) {
    const HcSite st = hc_site(sw, a);
    const int idx = st.idx;
    // Full x = 2 xh + e: the inner x neighbour is ahead (X+) for e = 0.
    const int e = a.parity ^ ((st.y + st.z + st.t) & 1);
    const A s = A(a.sign);
    const A sx = e ? -s : s, sy = (cn & 1u) ? -s : s, sz = (cn & 2u) ? -s : s, stt = (cn & 4u) ? -s : s;
    const int dy = a.subLx, dz = a.subLx * a.Ly, dt = dz * a.Lz;
    // Outer neighbours (periodic wrap) and whether their term is live.
    const int nx = e ? (st.xh + 1 == a.subLx ? idx - st.xh : idx + 1) : (st.xh == 0 ? idx + a.subLx - 1 : idx - 1);
    const int ny = (cn & 1u) ? (st.y + 1 == a.Ly ? idx - (a.Ly - 1) * dy : idx + dy)
                             : (st.y == 0 ? idx + (a.Ly - 1) * dy : idx - dy);
    const int nz = (cn & 2u) ? (st.z + 1 == a.Lz ? idx - (a.Lz - 1) * dz : idx + dz)
                             : (st.z == 0 ? idx + (a.Lz - 1) * dz : idx - dz);
    const int nt = (cn & 4u) ? (st.t + 1 == a.Lt ? idx - (a.Lt - 1) * dt : idx + dt)
                             : (st.t == 0 ? idx + (a.Lt - 1) * dt : idx - dt);
    const bool lx = !(a.skip[0] && (e ? st.xh == a.subLx - 1 : st.xh == 0));
    const bool ly = !(a.skip[1] && ((cn & 1u) ? st.y == a.Ly - 1 : st.y == 0));
    const bool lz = !(a.skip[2] && ((cn & 2u) ? st.z == a.Lz - 1 : st.z == 0));
    const bool lt = !(a.skip[3] && ((cn & 4u) ? st.t == a.Lt - 1 : st.t == 0));
    const CG *blk8 = reinterpret_cast<const CG *>(smem) + size_t(sw - (lane0 >> lp)) * 8 * LC;
    const A *ef = a.efac_off ? reinterpret_cast<const A *>(smem + a.efac_off) : nullptr;

    cpx<A> acc[R][4][N];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int sp = 0; sp < 4; ++sp)
#pragma unroll
        for (int c = 0; c < N; ++c) acc[r][sp][c] = cpx<A>{A(-0.0), A(-0.0)};

    constexpr bool kPair = SL == SL_PACKED && std::is_same<CS, CG>::value &&
                           (std::is_same<CS, c16>::value || std::is_same<CS, cb16>::value) &&
                           std::is_same<A, float>::value;
    // Inner hops: x from the own spinor, y / z / t through the partners;
    // then the outer hops (the opposite direction of the inner one), one load
    // ahead. Link slots: forward mu, backward 4 + mu.
    if constexpr (kPair) {
      using PV = typename Pair<CS>::V;
      PV psA[R][4][N], psB[R][4][N];
      load_spinor_p<CS, N, R>(psA, a, idx, int(slot));  // own inner site
      load_spinor_p<CS, N, R>(psB, a, nx, int(slot));   // x outer
      hc_hop_p<CS, N, 0, R>(acc, psA, blk8, e ? 4 : 0, sx);
      hc_hop_xchg_p<CS, N, 1, R>(acc, psA, blk8, (cn & 1u) ? 5 : 1, sy, int(Pl));
      hc_hop_xchg_p<CS, N, 2, R>(acc, psA, blk8, (cn & 2u) ? 6 : 2, sz, int(2 * Pl));
      hc_hop_xchg_p<CS, N, 3, R>(acc, psA, blk8, (cn & 4u) ? 7 : 3, stt, int(4 * Pl));
      load_spinor_p<CS, N, R>(psA, a, ny, int(slot));
      if (lx) hc_hop_p<CS, N, 0, R>(acc, psB, blk8, e ? 0 : 4, -sx);
      load_spinor_p<CS, N, R>(psB, a, nz, int(slot));
      if (ly) hc_hop_p<CS, N, 1, R>(acc, psA, blk8, (cn & 1u) ? 1 : 5, -sy);
      load_spinor_p<CS, N, R>(psA, a, nt, int(slot));
      if (lz) hc_hop_p<CS, N, 2, R>(acc, psB, blk8, (cn & 2u) ? 2 : 6, -sz);
      if (lt) hc_hop_p<CS, N, 3, R>(acc, psA, blk8, (cn & 4u) ? 3 : 7, -stt);
    } else {
      cpx<A> psA[R][4][N], psB[R][4][N];
      load_spinor<N, R, SL, CS, A>(psA, a, a.in, idx, int(slot));  // own inner site
      load_spinor<N, R, SL, CS, A>(psB, a, a.in, nx, int(slot));   // x outer
      hc_hop<N, 0, R, CG, A>(acc, psA, blk8, e ? 4 : 0, sx);
      hc_hop_xchg<N, 1, R, CG, A>(acc, psA, blk8, (cn & 1u) ? 5 : 1, sy, int(Pl));
      hc_hop_xchg<N, 2, R, CG, A>(acc, psA, blk8, (cn & 2u) ? 6 : 2, sz, int(2 * Pl));
      hc_hop_xchg<N, 3, R, CG, A>(acc, psA, blk8, (cn & 4u) ? 7 : 3, stt, int(4 * Pl));
      load_spinor<N, R, SL, CS, A>(psA, a, a.in, ny, int(slot));
      if (lx) hc_hop<N, 0, R, CG, A>(acc, psB, blk8, e ? 0 : 4, -sx);
      load_spinor<N, R, SL, CS, A>(psB, a, a.in, nz, int(slot));
      if (ly) hc_hop<N, 1, R, CG, A>(acc, psA, blk8, (cn & 1u) ? 1 : 5, -sy);
      load_spinor<N, R, SL, CS, A>(psA, a, a.in, nt, int(slot));
      if (lz) hc_hop<N, 2, R, CG, A>(acc, psB, blk8, (cn & 2u) ? 2 : 6, -sz);
      if (lt) hc_hop<N, 3, R, CG, A>(acc, psA, blk8, (cn & 4u) ? 3 : 7, -stt);
    }

    if constexpr (SL == SL_PACKED) epilogue_packed<N, R, CS, CO, A>(a, acc, idx, slot, ef, nrm);
    else epilogue_ptrs<N, SL, CS, CO, A>(a, acc, idx, slot, nrm);
  }
  workgroup_partials<R, A>(a, blk, active, slot, nrm);
}

// KF > 0: compiled for exactly KF right-hand sides (constant row offsets and
// lane shifts).
template <int N, int R, int SL, class CS, class CG, class CO, class A, int KF = 0>
__global__ void __launch_bounds__(256) dslash_hc_kernel(const DslashArgs a)
{
  static_assert(SL == SL_PACKED || (SL == SL_AOS_PTRS && R == 1 && KF == 0),
                "hypercube kernels: packed spinors, or user AoS pointers with one RHS per lane");
  if constexpr (KF > 0) {
    static_assert(KF % R == 0 && (KF / R) <= 32 && ((KF / R) & (KF / R - 1)) == 0, "k / R: a power of two <= 32");
    DslashArgs b = a;
    b.k = KF;
    b.hc_lp = uint32_t(__builtin_ctz(unsigned(KF / R)));
    dslash_hc_body<N, R, SL, CS, CG, CO, A>(b, blockIdx.x);
  } else {
    dslash_hc_body<N, R, SL, CS, CG, CO, A>(a, blockIdx.x);
  }
}

}  // namespace qmrhs

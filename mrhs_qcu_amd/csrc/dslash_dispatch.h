// Host-side launch interface of the dslash kernel family: a registry maps a
// kernel key (layouts, precisions, N, tile) to the instantiated launcher, and
// launch_dslash() validates a call, fills DslashArgs and launches.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "dslash_hcube.h"
#include "dslash_kernel.h"

namespace qmrhs {

struct KernelKey {
  int sl, gl, n, r, lds, site_minor, prec_s, prec_g, prec_o;
  int kf = 0;  // > 0: instantiated for exactly this RHS count
  int hc = 0;  // 1: hypercube work mapping (dslash_hcube.h)
  bool operator==(const KernelKey &o) const
  {
    return sl == o.sl && gl == o.gl && n == o.n && r == o.r && lds == o.lds && site_minor == o.site_minor &&
           prec_s == o.prec_s && prec_g == o.prec_g && prec_o == o.prec_o && kf == o.kf && hc == o.hc;
  }
};

using LaunchFn = hipError_t (*)(const DslashArgs &, dim3, dim3, size_t, hipStream_t);
using AttrFn = hipError_t (*)(hipFuncAttributes *);

void register_kernel(const KernelKey &key, LaunchFn fn, AttrFn attr);
LaunchFn find_kernel(const KernelKey &key);
AttrFn find_kernel_attr(const KernelKey &key);
int registry_size();

template <int N, int R, int SL, int GL, bool LDS, bool SM, class CS, class CG, class CO, class A, int KF>
hipError_t launch_impl(const DslashArgs &a, dim3 g, dim3 b, size_t shm, hipStream_t st)
{
  hipLaunchKernelGGL((dslash_kernel<N, R, SL, GL, LDS, SM, CS, CG, CO, A, KF>), g, b, shm, st, a);
  return hipGetLastError();
}
template <int N, int R, int SL, int GL, bool LDS, bool SM, class CS, class CG, class CO, class A, int KF>
hipError_t attr_impl(hipFuncAttributes *at)
{
  return hipFuncGetAttributes(at,
                              reinterpret_cast<const void *>(&dslash_kernel<N, R, SL, GL, LDS, SM, CS, CG, CO, A, KF>));
}

template <int N, int R, int SL, int GL, bool LDS, bool SM, class CS, class CG, class CO, class A, int KF = 0>
struct Reg {
  Reg()
  {
    KernelKey key{SL, GL, N, R, LDS ? 1 : 0, SM ? 1 : 0, StoreInfo<CS>::prec, StoreInfo<CG>::prec, StoreInfo<CO>::prec};
    key.kf = KF;
    register_kernel(key, &launch_impl<N, R, SL, GL, LDS, SM, CS, CG, CO, A, KF>,
                    &attr_impl<N, R, SL, GL, LDS, SM, CS, CG, CO, A, KF>);
  }
};

template <int N, int R, int SL, class CS, class CG, class CO, class A, int KF>
hipError_t launch_hc_impl(const DslashArgs &a, dim3 g, dim3 b, size_t shm, hipStream_t st)
{
  hipLaunchKernelGGL((dslash_hc_kernel<N, R, SL, CS, CG, CO, A, KF>), g, b, shm, st, a);
  return hipGetLastError();
}
template <int N, int R, int SL, class CS, class CG, class CO, class A, int KF>
hipError_t attr_hc_impl(hipFuncAttributes *at)
{
  return hipFuncGetAttributes(at, reinterpret_cast<const void *>(&dslash_hc_kernel<N, R, SL, CS, CG, CO, A, KF>));
}

// Hypercube kernels: packed layout (or user AoS pointers), native links
// staged in LDS.
template <int N, int R, class CS, class CG, class CO, class A, int KF = 0, int SL = SL_PACKED>
struct RegHc {
  RegHc()
  {
    KernelKey key{SL, GL_NATIVE, N, R, 1, 0, StoreInfo<CS>::prec, StoreInfo<CG>::prec, StoreInfo<CO>::prec};
    key.kf = KF;
    key.hc = 1;
    register_kernel(key, &launch_hc_impl<N, R, SL, CS, CG, CO, A, KF>,
                    &attr_hc_impl<N, R, SL, CS, CG, CO, A, KF>);
  }
};

// Dynamic LDS a launch may request: 64 KiB per workgroup minus the kernel's
// static 2 KiB per-RHS reduction buffer.
constexpr int kMaxDynLds = 65536 - 2048;
// LDS of one gfx950 compute unit.
constexpr int kLdsPerCu = 160 * 1024;

// Geometry of one lattice (local volume + process grid) and colour count.
struct Geom {
  int L[4];
  int grid[4];
  int N;
};

// Tile shape of a launch.
struct Tile {
  int R = 1;       // RHS per lane
  int block = 256; // lanes per workgroup
  int lds = 0;     // stage links in LDS
  int tblock = 1;  // traversal in blocks of tblock time slices ...
  int zblock = 0;  // ... and zblock z slices (0 = Lz: no z blocking)
  int hints = 0;   // bit 0: non-temporal output stores, bit 1: non-temporal link loads,
                   // bit 2: keep the hardware workgroup order (no XCD remap),
                   // bit 3: XCD-cyclic order over half y-streamed blocks (xcd_cycle_gran)
  int occ = 0;     // cap on resident workgroups per CU (LDS padding); 0 = no cap
  int ystream = 0; // 1: stream each (tblock, zblock) block along y
  int hcube = 0;   // 1: hypercube work mapping (2x2x2x2 blocks, one per 8 P lanes; tblock / zblock
                   // then count blocks)
  int trav_hi = 0; // traversal bits 2 and above: 1 = column order (tblock / zblock = a column's extent
                   // in y / z, in sites, or in blocks with hcube; XCD x walks columns 8 g + x along t),
                   // 3 = the same with x split in two halves
};

// The tile a site-major kernel runs for tile t (hcube cleared): a
// column-order tile (traversal 4, or 6 with its extents in blocks doubled to
// sites) keeps its columns when they fill whole workgroups of the packed
// layout, else it becomes y-streamed blocks of one t slice and z_block z
// slices in the XCD-cyclic order.
Tile site_major_tile(const Tile &t, const Geom &g, int k, int sl);

// The traversal field of qcuMrhsTile: bit 0 ystream, bit 1 hcube, bit 2 column order.
inline int tile_traversal(const Tile &t) { return t.ystream | (t.hcube << 1) | (t.trav_hi << 2); }

struct DslashCall {
  int sl = SL_PACKED, gl = GL_NATIVE;
  int prec_s = PREC_SINGLE, prec_g = PREC_SINGLE, prec_o = PREC_SINGLE;
  Geom geom{};
  Tile tile{};
  int k = 1, parity = 0, dagger = 0;
  void *out = nullptr;
  const void *in = nullptr, *gauge = nullptr, *x = nullptr;
  void *const *out_ptrs = nullptr;       // host arrays of k device pointers (pointer layouts)
  const void *const *in_ptrs = nullptr;
  const void *const *x_ptrs = nullptr;   // pointer layouts: per-RHS xpay operand (out = xa x + xb D in)
  const float *sc_in = nullptr, *sc_x = nullptr, *sc_out = nullptr;
  double xa = 0.0, xb = 1.0;
  double *partial = nullptr;  // [nblocks][k]; nblocks from dslash_num_blocks()
  // Fused CG update r -= alpha out (FP32 or FP64 packed output only).
  void *rx_r = nullptr;
  const double *rx_alpha = nullptr;
  // GL_NATIVE12: time links across the t boundary carry a factor -1.
  int t_flip = 0;
  hipStream_t stream = nullptr;
};

// Returns 0 on success; otherwise a message in *err (when non-null).
int launch_dslash(const DslashCall &c, std::string *err);
// The kernel arguments and dynamic LDS bytes launch_dslash would use for a
// packed-layout call (validated the same way), without launching; for
// work-list launches that run several calls' workgroups in one kernel.
int packed_launch_args(const DslashCall &c, DslashArgs *a, size_t *shm, std::string *err);
// Workgroups a packed-layout launch uses (sizing of the partial-sum buffer).
long dslash_num_blocks(const DslashCall &c);

// Workgroups per chunk of the XCD-cyclic order (cache hint bit 3) for a
// packed launch: half of one y-streamed (t_block x z_block) block, so the 8
// XCDs work side by side on adjacent blocks and share their halo sites
// through the memory-side cache. 0 when the tile does not ask for it or the
// lattice does not split into whole chunks (the launch then keeps the
// XCD-contiguous order).
unsigned xcd_cycle_gran(const Geom &g, const Tile &t, int k);
// Default tile for a packed launch (used when the autotuner has no entry).
Tile default_tile(int N, int k, int prec);
Tile default_tile(const Geom &g, int k, int prec);  // with a traversal for the lattice

}  // namespace qmrhs

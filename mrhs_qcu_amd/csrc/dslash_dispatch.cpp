// Validation and launch of the dslash kernel family.
#include "dslash_dispatch.h"

#include <cstdint>
#include <mutex>
#include <vector>

namespace qmrhs {

namespace {
struct Entry {
  KernelKey key;
  LaunchFn fn;
  AttrFn attr;
};
std::vector<Entry> &registry()
{
  static std::vector<Entry> r;
  return r;
}
std::mutex &registry_mutex()
{
  static std::mutex m;
  return m;
}

int bytes_of(int prec)
{
  switch (prec) {
    case PREC_DOUBLE: return 16;
    case PREC_SINGLE: return 8;
    case PREC_HALF: return 4;
    case PREC_BF16: return 4;
    case PREC_FP8: return 2;
    default: return 0;
  }
}
}  // namespace

void register_kernel(const KernelKey &key, LaunchFn fn, AttrFn attr)
{
  std::lock_guard<std::mutex> lk(registry_mutex());
  registry().push_back(Entry{key, fn, attr});
}

LaunchFn find_kernel(const KernelKey &key)
{
  std::lock_guard<std::mutex> lk(registry_mutex());
  for (const Entry &e : registry())
    if (e.key == key) return e.fn;
  return nullptr;
}

AttrFn find_kernel_attr(const KernelKey &key)
{
  std::lock_guard<std::mutex> lk(registry_mutex());
  for (const Entry &e : registry())
    if (e.key == key) return e.attr;
  return nullptr;
}

int registry_size()
{
  std::lock_guard<std::mutex> lk(registry_mutex());
  return int(registry().size());
}

Tile default_tile(int N, int k, int prec)
{
  // The shape the autotuner picks most often on gfx950 for N = 3: two RHS per
  // lane (four for 16-bit storage), links staged in LDS, 128-lane workgroups.
  Tile t;
  t.block = 128;
  t.tblock = 1;
  t.R = 1;
  t.lds = 0;
  const bool half = prec == PREC_HALF || prec == PREC_BF16;
  if (half && k % 4 == 0 && N <= 3) t.R = 4;
  else if (k % 2 == 0 && N <= 5) t.R = 2;
  if (k / t.R >= 2) t.lds = 1;
  if (prec == PREC_FP8) t.lds = 1;  // the FP8 kernels stage links in LDS
  return t;
}

Tile default_tile(const Geom &g, int k, int prec)
{
  Tile t = default_tile(g.N, k, prec);
  // y-streamed blocks of the largest t and z extents that divide the lattice.
  int tb = 0, zb = 0;
  for (int c : {6, 4, 3, 2})
    if (!tb && g.L[3] % c == 0 && c < g.L[3]) tb = c;
  for (int c : {4, 6, 3, 2})
    if (!zb && g.L[2] % c == 0 && c < g.L[2]) zb = c;
  if (tb && zb) {
    t.tblock = tb;
    t.zblock = zb;
    t.ystream = 1;
  }
  return t;
}

static bool fail(std::string *err, const std::string &m)
{
  if (err) *err = m;
  return false;
}

static bool check_geom(const Geom &g, std::string *err)
{
  for (int d = 0; d < 4; ++d) {
    if (g.L[d] <= 0 || (g.L[d] & 1)) return fail(err, "lattice extents must be positive and even");
    if (g.grid[d] <= 0) return fail(err, "process grid extents must be positive");
  }
  if (g.N < 2 || g.N > 8) return fail(err, "nColors must be in 2..8");
  const long long V = 1LL * g.L[0] * g.L[1] * g.L[2] * g.L[3];
  if (V / 2 >= (1LL << 30)) return fail(err, "local volume too large for 32-bit site indexing");
  return true;
}

static void fill_common(DslashArgs &a, const DslashCall &c)
{
  a.out = c.out;
  a.in = c.in;
  a.gauge = c.gauge;
  a.x = c.x;
  a.sc_in = c.sc_in;
  a.sc_x = c.sc_x;
  a.sc_out = c.sc_out;
  a.xa = c.xa;
  a.xb = c.xb;
  a.partial = c.partial;
  a.t_flip = c.t_flip ? 1 : 0;
  a.rx_r = c.rx_r;
  a.rx_alpha = c.rx_alpha;
  a.plain_out = (!c.x && c.xb == 1.0 && !c.sc_in && !c.sc_out && !c.partial && !c.rx_r) ? 1 : 0;
  a.sign = c.dagger ? -1.0f : 1.0f;
  a.parity = c.parity;
  a.subLx = c.geom.L[0] / 2;
  a.Ly = c.geom.L[1];
  a.Lz = c.geom.L[2];
  a.Lt = c.geom.L[3];
  a.S = a.subLx * a.Ly * a.Lz * a.Lt;
  for (int d = 0; d < 4; ++d) a.skip[d] = c.geom.grid[d] > 1 ? 1 : 0;
  a.tblock = c.tile.tblock < 1 ? 1 : c.tile.tblock;
  a.zblock = c.tile.zblock < 1 ? a.Lz : c.tile.zblock;
  a.nt_store = (c.tile.hints & 1) ? 1 : 0;
  a.nt_link = (c.tile.hints & 2) ? 1 : 0;
  a.xcd_map = (c.tile.hints & 4) ? 0 : 1;
  a.div_subLx = make_fastdiv(uint32_t(a.subLx));
  a.div_row = make_fastdiv(uint32_t(a.subLx * a.Ly));
  a.div_tb = make_fastdiv(uint32_t(a.tblock));
  a.div_zb = make_fastdiv(uint32_t(a.zblock));
  a.div_nzb = make_fastdiv(uint32_t(a.Lz / a.zblock));
  a.div_Ly = make_fastdiv(uint32_t(a.Ly));
  a.ystream = c.tile.ystream ? 1 : 0;
  a.div_S = make_fastdiv(uint32_t(a.S));
}

// A hypercube tile runs the hypercube kernel when one is instantiated for
// the call's layouts, precisions, N and R (else the site-major kernel of the
// same blocking runs).
static bool runs_hcube(const DslashCall &c)
{
  if (!c.tile.hcube || c.sl != SL_PACKED) return false;
  KernelKey kh{SL_PACKED, c.gl, c.geom.N, c.tile.R, c.tile.lds ? 1 : 0, 0, c.prec_s, c.prec_g, c.prec_o};
  kh.hc = 1;
  KernelKey kh_kf = kh;
  kh_kf.kf = c.k;
  return find_kernel(kh) || find_kernel(kh_kf);
}

long dslash_num_blocks(const DslashCall &c)
{
  const long S = long(c.geom.L[0]) * c.geom.L[1] * c.geom.L[2] * c.geom.L[3] / 2;
  const int R = c.sl == SL_PACKED ? c.tile.R : 1;
  const long lanes = S * (c.k / (R > 0 ? R : 1));
  const int B = c.tile.block > 0 ? c.tile.block : 256;
  return (lanes + B - 1) / B;
}

unsigned xcd_cycle_gran(const Geom &g, const Tile &t, int k)
{
  if (!(t.hints & 8) || (t.hints & 4) || !t.ystream || t.R < 1 || k % t.R) return 0;
  const int B = t.block > 0 ? t.block : 256;
  const long long tb = t.tblock < 1 ? 1 : t.tblock;
  const long long P = k / t.R;
  const long long lanes = 1LL * g.L[0] / 2 * g.L[1] * g.L[2] * g.L[3] * P;
  long long half;  // lanes per half block
  if (t.hcube) {
    // Block lattice: 8 P lanes per 2x2x2x2 block, Ly / 2 blocks along y
    // (a whole y line of blocks when that is odd).
    const long long zb = t.zblock < 1 ? g.L[2] / 2 : t.zblock;
    const long long by = g.L[1] / 2;
    half = tb * zb * (by % 2 ? by : by / 2) * (g.L[0] / 2) * 8 * P;
  } else {
    const long long zb = t.zblock < 1 ? g.L[2] : t.zblock;
    half = tb * zb * (g.L[1] / 2) * (g.L[0] / 2) * P;
  }
  if (half % B || lanes % B) return 0;
  const long long gran = half / B, nblocks = lanes / B;
  if (nblocks % (8 * gran)) return 0;
  return unsigned(gran);
}

int prepare_packed(const DslashCall &c, DslashArgs *ap, size_t *shm_out, LaunchFn *fn_out, std::string *err)
{
  DslashArgs &a = *ap;
  const int B = c.tile.block;
  const int R = c.tile.R;
  if (R != 1 && R != 2 && R != 4) return fail(err, "rhs_per_lane must be 1, 2 or 4"), 1;
  if (c.k % R) return fail(err, "nrhs must be a multiple of rhs_per_lane"), 1;
  if (!c.out || !c.in) return fail(err, "null spinor"), 1;
  const uintptr_t al = uintptr_t(16);
  if ((uintptr_t(c.out) | uintptr_t(c.in) | uintptr_t(c.gauge) | uintptr_t(c.x)) % al)
    return fail(err, "packed fields must be 16-byte aligned"), 1;
  if (c.partial && c.k > 256) return fail(err, "norm partials support at most 256 RHS"), 1;
  if (c.rx_r) {
    if (c.prec_o != PREC_SINGLE && c.prec_o != PREC_DOUBLE)
      return fail(err, "the fused CG update needs FP32 or FP64 output"), 1;
    if (!c.rx_alpha) return fail(err, "fused CG update: null alpha"), 1;
    if (uintptr_t(c.rx_r) % al) return fail(err, "packed fields must be 16-byte aligned"), 1;
  }
  if (c.gl != GL_NATIVE && c.gl != GL_NATIVE12 && c.gl != GL_NATIVE8)
    return fail(err, "packed spinors use the native (or two-row / eight-real SU(3)) gauge layout"), 1;
  if (c.gl == GL_NATIVE12 && c.geom.N != 3) return fail(err, "two-row links need nColors = 3"), 1;
  if (c.gl == GL_NATIVE8 && c.geom.N != 3) return fail(err, "eight-real links need nColors = 3"), 1;
  if (c.gl == GL_NATIVE8 && !c.tile.lds)
    return fail(err, "eight-real links are rebuilt while staging: the tile needs link_lds = 1"), 1;
  const int lds = c.tile.lds ? 1 : 0;
  // A hypercube tile runs the site-major kernel of the same shape where no
  // hypercube kernel is instantiated for these precisions / N / R (its
  // blocking divides the lattice either way).
  const int hc = runs_hcube(c) ? 1 : 0;
  if (hc) {
    // 2x2x2x2 blocks of 8 P lanes (P = k / R a power of two <= 32; the
    // workgroup holds whole blocks), native links staged in LDS; tblock /
    // zblock count blocks.
    const int P = c.k / R;
    if (P > 32 || (P & (P - 1)))
      return fail(err, "hypercube tiles need nrhs / rhs_per_lane in {1, 2, 4, 8, 16, 32}"), 1;
    if (B % (8 * P)) return fail(err, "hypercube tiles need a block that is a multiple of 8 * nrhs / rhs_per_lane"), 1;
    if (c.gl != GL_NATIVE || !lds) return fail(err, "hypercube tiles need native links staged in LDS"), 1;
    const int bz = c.geom.L[2] / 2;
    a.tblock = c.tile.tblock < 1 ? 1 : c.tile.tblock;
    a.zblock = c.tile.zblock < 1 ? bz : c.tile.zblock;
    a.bLy = c.geom.L[1] / 2;
    a.bLz = bz;
    a.div_tb = make_fastdiv(uint32_t(a.tblock));
    a.div_zb = make_fastdiv(uint32_t(a.zblock));
    a.hdiv_Ly = make_fastdiv(uint32_t(a.bLy));
    a.hdiv_row = make_fastdiv(uint32_t(a.subLx * a.bLy));
    a.hdiv_nzb = make_fastdiv(uint32_t(bz / a.zblock));
    a.hc_lp = uint32_t(__builtin_ctz(unsigned(P)));
  }
  KernelKey key{SL_PACKED, c.gl, c.geom.N, R, lds, 0, c.prec_s, c.prec_g, c.prec_o};
  key.hc = hc;
  // The same kernel compiled for this exact RHS count, when one exists
  // (bit-identical results, constant row offsets).
  KernelKey key_kf = key;
  key_kf.kf = c.k;
  LaunchFn fn = find_kernel(key_kf);
  if (!fn) fn = find_kernel(key);
  if (!fn) return fail(err, "no kernel instantiated for this precision / N / tile"), 1;
  a.k = c.k;
  a.P = c.k / R;
  a.div_P = make_fastdiv(uint32_t(a.P));
  const unsigned long long lanes = 1ULL * a.S * a.P;
  if (lanes >= (1ULL << 32)) return fail(err, "too many lanes for one launch"), 1;
  a.nlanes = uint32_t(lanes);
  a.nblocks = uint32_t((lanes + B - 1) / B);
  a.ngrid = a.nblocks;
  Tile tl = c.tile;
  tl.hcube = hc;
  if (c.tile.trav_hi) {
    // Column order: t_block / z_block are a column's extent in y / z (blocks
    // for the hypercube kernel, sites for the site-major one), x whole or
    // (traversal bit 3) halved; every column is a whole number of workgroups.
    const int cy = c.tile.tblock, cz = c.tile.zblock;
    const int nx = (c.tile.trav_hi & 2) ? 2 : 1;
    if (a.subLx % nx) return fail(err, "column order: splitting x needs Lx / 2 even"), 1;
    const int ey = hc ? c.geom.L[1] / 2 : c.geom.L[1], ez = hc ? c.geom.L[2] / 2 : c.geom.L[2];
    const uint64_t nT = uint64_t(hc ? c.geom.L[3] / 2 : c.geom.L[3]);
    const uint64_t slab = uint64_t(a.subLx / nx) * cy * cz;
    const uint64_t col_lanes = slab * nT * (hc ? 8ull * a.P : uint64_t(a.P));
    if (col_lanes % uint64_t(B)) return fail(err, "column order: a column must fill whole workgroups"), 1;
    a.col_order = 1;
    a.col_cy = uint32_t(cy);
    a.col_cz = uint32_t(cz);
    a.col_ncol = uint32_t((ey / cy) * (ez / cz) * nx);
    a.col_sx = uint32_t(a.subLx / nx);
    a.col_nx = uint32_t(nx);
    a.cdiv_sx = make_fastdiv(a.col_sx);
    a.cdiv_nx = make_fastdiv(a.col_nx);
    a.col_ncy = uint32_t(ey / cy);
    a.col_nT = uint32_t(nT);
    a.col_slab = uint32_t(slab);
    a.col_grp = uint32_t(col_lanes / uint64_t(B));
    a.cdiv_slab = make_fastdiv(a.col_slab);
    a.cdiv_nT = make_fastdiv(a.col_nT);
    a.cdiv_cy = make_fastdiv(a.col_cy);
    a.cdiv_ncy = make_fastdiv(a.col_ncy);
    a.cdiv_grp = make_fastdiv(a.col_grp);
    if (uint64_t(a.nblocks) != uint64_t(a.col_ncol) * a.col_grp)
      return fail(err, "column order: the columns do not tile the lattice"), 1;
    // The grid pads the last group of 8 columns with idle workgroups.
    a.ngrid = (a.col_ncol + 7u) / 8u * 8u * a.col_grp;
    a.xcd_map = 0;
  }
  a.xcd_gran = a.col_order ? 0u : xcd_cycle_gran(c.geom, tl, c.k);
  size_t shm = 0;
  if (lds) {
    const size_t nsite_max = size_t((B - 1) / a.P + 2);
    // Two-row and eight-real links are staged as full links (rebuilt while staging).
    shm = nsite_max * 8 * link_cplx(GL_NATIVE, c.geom.N) * size_t(bytes_of(c.prec_g));
    if (shm > kMaxDynLds) return fail(err, "LDS link staging exceeds the per-workgroup LDS budget"), 1;
    // Per-RHS epilogue factors staged beside the links when they are used
    // and fit, for 16-bit input (FP32 accumulation): at 32^3x64, 16 RHS,
    // the scaled FP16 operator takes 624 instead of 644 us; for FP32 input
    // it made no difference (1,094 / 1,097 us).
    const bool half_in = c.prec_s == PREC_HALF || c.prec_s == PREC_BF16 || c.prec_s == PREC_FP8;
    const size_t efac = size_t(4) * c.k * sizeof(float);
    if (half_in && !a.plain_out && shm + efac <= size_t(kMaxDynLds)) {
      a.efac_off = uint32_t(shm);
      shm += efac;
    }
  }
  if (c.tile.occ > 0) {
    // Pad the LDS request so at most `occ` workgroups fit on a CU.
    const size_t cap = size_t(kLdsPerCu / c.tile.occ) - 2048 - 512;
    if (cap > size_t(kMaxDynLds)) return fail(err, "occupancy cap below 3 workgroups per CU is not reachable"), 1;
    if (shm < cap) shm = cap;
  }
  *shm_out = shm;
  *fn_out = fn;
  return 0;
}

static int check_call(const DslashCall &c, std::string *err)
{
  if (!check_geom(c.geom, err)) return 1;
  if (c.k < 1) return fail(err, "nrhs must be >= 1"), 1;
  if (c.parity != 0 && c.parity != 1) return fail(err, "parity must be 0 or 1"), 1;
  if (!bytes_of(c.prec_s) || !bytes_of(c.prec_g) || !bytes_of(c.prec_o))
    return fail(err, "unsupported storage precision"), 1;
  if ((c.prec_s == PREC_FP8 || c.prec_o == PREC_FP8) &&
      (c.sl != SL_PACKED || c.geom.N != 3 || c.k % 2 || c.tile.R < 2 || !c.tile.lds))
    return fail(err, "FP8 storage is instantiated for the packed layout, nColors = 3, an even number of right-hand "
                     "sides and tiles with rhs_per_lane 2 or 4 and LDS-staged links"), 1;
  const int B = c.tile.block;
  if (B < 64 || B > 256 || B % 64) return fail(err, "block must be 64, 128, 192 or 256"), 1;
  if (c.tile.trav_hi > 3 || c.tile.trav_hi == 2)
    return fail(err, "traversal bits 3 and above are defined only as bit 3 with bit 2 (column order split in x)"), 1;
  if (c.tile.trav_hi) {
    // Column order of the packed layout: t_block x z_block blocks (traversal
    // 6) or sites (traversal 4) in (y, z).
    if (c.tile.ystream || c.sl != SL_PACKED)
      return fail(err, "column order (traversal bit 2) is a traversal of the packed layout (traversal 4 or 6)"), 1;
    const int f = c.tile.hcube ? 2 : 1;
    if (c.tile.tblock < 1 || (c.geom.L[1] / f) % c.tile.tblock)
      return fail(err, c.tile.hcube ? "column order: t_block (column extent in y blocks) must divide Ly / 2"
                                    : "column order: t_block (column extent in y) must divide Ly"), 1;
    if (c.tile.zblock < 1 || (c.geom.L[2] / f) % c.tile.zblock)
      return fail(err, c.tile.hcube ? "column order: z_block (column extent in z blocks) must divide Lz / 2"
                                    : "column order: z_block (column extent in z) must divide Lz"), 1;
  } else if (!c.tile.hcube) {
    if (c.tile.tblock > 1 && c.geom.L[3] % c.tile.tblock) return fail(err, "t_block must divide Lt"), 1;
    if (c.tile.zblock > 0 && c.geom.L[2] % c.tile.zblock) return fail(err, "z_block must divide Lz"), 1;
  } else {
    // Counted in 2x2x2x2 blocks (so they also divide Lt and Lz when the
    // site-major kernel runs the tile).
    if (c.sl == SL_SOA_PTRS) return fail(err, "hypercube tiles are for the packed and user AoS layouts"), 1;
    if (c.tile.tblock > 1 && (c.geom.L[3] / 2) % c.tile.tblock)
      return fail(err, "hypercube t_block must divide Lt / 2"), 1;
    if (c.tile.zblock > 0 && (c.geom.L[2] / 2) % c.tile.zblock)
      return fail(err, "hypercube z_block must divide Lz / 2"), 1;
  }
  if (!c.gauge) return fail(err, "null gauge"), 1;
  return 0;
}

Tile site_major_tile(const Tile &t0, const Geom &g, int k, int sl)
{
  Tile t = t0;
  t.hcube = 0;
  if (!t0.trav_hi) return t;
  const int R = t.R > 0 ? t.R : 1, B = t.block > 0 ? t.block : 256;
  const int f = t0.hcube ? 2 : 1;  // column extents in sites
  const int nx = (t0.trav_hi & 2) ? 2 : 1;
  const long long col_lanes = 1LL * (g.L[0] / 2 / nx) * (f * t0.tblock) * (f * t0.zblock) * g.L[3] * (k / R);
  if (sl == SL_PACKED && k % R == 0 && (g.L[0] / 2) % nx == 0 && col_lanes % B == 0) {
    // The same columns in sites (traversal 4).
    t.tblock = f * t0.tblock;
    t.zblock = f * t0.zblock;
    return t;
  }
  t.trav_hi = 0;
  t.ystream = 1;
  t.tblock = 1;
  t.hints = 8;
  return t;
}

// A column-order hypercube tile (traversal 6) on a call that no hypercube
// kernel serves (compressed links, precisions / N / rhs_per_lane without
// one) runs the site-major kernel in the same columns (site_major_tile).
static DslashCall effective_call(const DslashCall &c0)
{
  if (!c0.tile.trav_hi || runs_hcube(c0)) return c0;
  DslashCall c = c0;
  c.tile = site_major_tile(c0.tile, c0.geom, c0.k, c0.sl);
  return c;
}

int packed_launch_args(const DslashCall &c0, DslashArgs *a, size_t *shm, std::string *err)
{
  const DslashCall c = effective_call(c0);
  if (c.sl != SL_PACKED) return fail(err, "work lists hold packed-layout launches"), 1;
  if (check_call(c, err)) return 1;
  *a = DslashArgs{};
  fill_common(*a, c);
  LaunchFn fn = nullptr;
  return prepare_packed(c, a, shm, &fn, err);
}

int launch_dslash(const DslashCall &c0, std::string *err)
{
  const DslashCall c = effective_call(c0);
  if (check_call(c, err)) return 1;
  const int B = c.tile.block;

  DslashArgs a{};
  fill_common(a, c);

  if (c.sl == SL_PACKED) {
    size_t shm = 0;
    LaunchFn fn = nullptr;
    if (prepare_packed(c, &a, &shm, &fn, err)) return 1;
    hipError_t e = fn(a, dim3(a.ngrid), dim3(B), shm, c.stream);
    if (e != hipSuccess) return fail(err, std::string("launch failed: ") + hipGetErrorString(e)), 1;
    return 0;
  }

  // Pointer layouts: one RHS per lane, at most kMaxPtrs RHS per launch.
  if (!c.out_ptrs || !c.in_ptrs) return fail(err, "null pointer arrays"), 1;
  if (c.x || c.sc_out || c.sc_x || c.partial || c.rx_r)
    return fail(err, "xpay / scaling / norms need the packed layout"), 1;
  const bool sm = c.sl == SL_SOA_PTRS;
  if (c.sl == SL_AOS_PTRS && c.gl != GL_NATIVE) return fail(err, "AoS spinors use the native gauge layout"), 1;
  if (c.sl == SL_SOA_PTRS && c.gl != GL_REF_SOA) return fail(err, "SoA spinors use the reference SoA gauge"), 1;
  if (c.sl != SL_AOS_PTRS && c.sl != SL_SOA_PTRS) return fail(err, "unknown spinor layout"), 1;
  KernelKey key{c.sl, c.gl, c.geom.N, 1, 0, sm ? 1 : 0, c.prec_s, c.prec_g, c.prec_o};
  LaunchFn fn = find_kernel(key);
  if (!fn) return fail(err, "no kernel instantiated for this layout / precision / N"), 1;
  // AoS pointers: links staged in LDS when that variant exists and fits.
  KernelKey key_lds = key;
  key_lds.lds = 1;
  LaunchFn fn_lds = c.sl == SL_AOS_PTRS ? find_kernel(key_lds) : nullptr;
  for (int r0 = 0; r0 < c.k; r0 += kMaxPtrs) {
    const int kk = (c.k - r0) < kMaxPtrs ? (c.k - r0) : kMaxPtrs;
    DslashArgs b = a;
    for (int r = 0; r < kk; ++r) {
      if (!c.out_ptrs[r0 + r] || !c.in_ptrs[r0 + r]) return fail(err, "null spinor pointer"), 1;
      b.out_ptrs.p[r] = c.out_ptrs[r0 + r];
      b.in_ptrs.p[r] = const_cast<void *>(c.in_ptrs[r0 + r]);
      if (c.x_ptrs) {
        if (!c.x_ptrs[r0 + r]) return fail(err, "null xpay spinor pointer"), 1;
        b.x_ptrs.p[r] = const_cast<void *>(c.x_ptrs[r0 + r]);
      }
    }
    b.xptr = c.x_ptrs ? 1 : 0;
    b.k = kk;
    b.P = kk;
    b.div_P = make_fastdiv(uint32_t(kk));
    // AoS pointers: blocks of t_block x z_block slices streamed along y, the
    // XCDs side by side on adjacent halves (24^3x48 FP32, 16 RHS: 720 ->
    // 658 us); the natural order when the lattice does not divide.
    int tb = 1, zb = b.Lz;
    if (c.sl == SL_AOS_PTRS && c.tile.ystream && !c.tile.hcube && c.tile.tblock > 1 && c.tile.zblock > 0 &&
        b.Lt % c.tile.tblock == 0 && b.Lz % c.tile.zblock == 0) {
      // A y-streamed user-layout tile gives the blocking itself.
      tb = c.tile.tblock;
      zb = c.tile.zblock;
    } else if (c.sl == SL_AOS_PTRS) {
      for (int t : {3, 2, 4, 6})
        if (b.Lt % t == 0 && t < b.Lt) {
          tb = t;
          break;
        }
      for (int z : {6, 4, 3, 2})
        if (b.Lz % z == 0 && z < b.Lz) {
          zb = z;
          break;
        }
      if (tb == 1 || zb == b.Lz) {
        tb = 1;
        zb = b.Lz;
      }
    }
    b.tblock = tb;
    b.zblock = zb;
    b.div_tb = make_fastdiv(uint32_t(tb));
    b.div_zb = make_fastdiv(uint32_t(zb));
    b.div_nzb = make_fastdiv(uint32_t(b.Lz / zb));
    b.ystream = tb > 1 ? 1 : 0;
    const unsigned long long lanes = 1ULL * b.S * kk;
    if (lanes >= (1ULL << 32)) return fail(err, "too many lanes for one launch"), 1;
    b.nlanes = uint32_t(lanes);
    b.nblocks = uint32_t((lanes + B - 1) / B);
    b.xcd_gran = 0;
    if (b.ystream && b.xcd_map) {
      Tile t;
      t.R = 1;
      t.block = B;
      t.tblock = tb;
      t.zblock = zb;
      t.hints = 8;
      t.ystream = 1;
      b.xcd_gran = xcd_cycle_gran(c.geom, t, kk);
    }
    uintptr_t ptr_bits = 0;
    for (int r = 0; r < kk; ++r) ptr_bits |= uintptr_t(b.out_ptrs.p[r]) | uintptr_t(b.in_ptrs.p[r]);
    b.aos_vec = (ptr_bits % 16) == 0 ? 1 : 0;
    size_t shm = 0;
    LaunchFn f = fn;
    int Bl = B;
    if (fn_lds) {
      const size_t lds_bytes = size_t((B - 1) / kk + 2) * 8 * c.geom.N * c.geom.N * size_t(bytes_of(c.prec_g));
      if (lds_bytes <= size_t(kMaxDynLds)) {
        f = fn_lds;
        shm = lds_bytes;
      }
    }
    // Group reads and writes of the AoS records (SL_AOS_COOP: a site's lanes
    // load its neighbours' records together and exchange them in LDS) for a
    // power-of-two RHS count (a site's lanes inside one wavefront), 16-byte
    // aligned buffers, no xpay operand and whole 16-byte pieces per record,
    // with the exchange area beside the staged links (the workgroup halved
    // until it fits), where they were measured faster at 24^3x48: from 8 RHS
    // (FP32 N = 2 / 3 / 4 at 16 RHS: 420 -> 276, 671 -> 479, 869 -> 781 us;
    // FP16: 198 -> 180, 336 -> 289, 416 -> 407 us), from 16 RHS for FP16 at
    // N >= 4 (8 RHS: 217 -> 227 us); at 4 RHS about even. Instantiated for
    // FP32 and FP16. Cache hint bit 4 keeps each lane reading its own record.
    const bool coop_pays = kk >= 8 && !(c.prec_s == PREC_HALF && c.geom.N >= 4 && kk < 16);
    if (c.sl == SL_AOS_PTRS && f == fn_lds && !(kk & (kk - 1)) && coop_pays && b.aos_vec && !b.xptr &&
        !(c.tile.hints & 16)) {
      KernelKey kc = key_lds;
      kc.sl = SL_AOS_COOP;
      LaunchFn fc = find_kernel(kc);
      const size_t rec = size_t(4 * c.geom.N) * size_t(bytes_of(c.prec_s));
      const size_t slot_bytes = (rec / 16 | 1) * 16;  // a record's slot of the exchange area (AosRec::QS)
      auto need = [&](int bl, size_t *off) {
        const size_t links = size_t((bl - 1) / kk + 2) * 8 * c.geom.N * c.geom.N * size_t(bytes_of(c.prec_g));
        *off = (links + 15) / 16 * 16;
        return *off + size_t(bl) * slot_bytes;
      };
      int Bc = B;
      size_t off = 0;
      while (Bc > 64 && need(Bc, &off) > size_t(kMaxDynLds)) Bc /= 2;
      const size_t tot = need(Bc, &off);
      if (fc && rec % 16 == 0 && Bc % kk == 0 && tot <= size_t(kMaxDynLds)) {
        f = fc;
        shm = tot;
        b.coop_off = uint32_t(off);
        if (Bc != B) {
          Bl = Bc;
          b.nblocks = uint32_t((b.nlanes + uint32_t(Bc) - 1) / uint32_t(Bc));
          if (b.xcd_gran) {
            Tile t;
            t.R = 1;
            t.block = Bc;
            t.tblock = tb;
            t.zblock = zb;
            t.hints = 8;
            t.ystream = 1;
            b.xcd_gran = xcd_cycle_gran(c.geom, t, kk);
          }
        }
      }
    }
    // A hypercube tile (the caller's, from the user-layout entry of the tile
    // table) runs the AoS pointers through the hypercube kernel when this
    // chunk's RHS count is a power of two, the workgroup holds whole blocks
    // and the links fit in LDS; otherwise the launch above is kept.
    if (c.sl == SL_AOS_PTRS && c.tile.hcube && kk <= 32 && !(kk & (kk - 1)) && c.tile.block % (8 * kk) == 0 &&
        c.tile.block >= 64 && c.tile.block <= 256) {
      KernelKey kh = key_lds;
      kh.hc = 1;
      LaunchFn fh = find_kernel(kh);
      const int Bh = c.tile.block;
      const size_t lds_h = size_t(Bh / kk) * 8 * c.geom.N * c.geom.N * size_t(bytes_of(c.prec_g));
      if (fh && lds_h <= size_t(kMaxDynLds)) {
        const int bz = b.Lz / 2;
        b.tblock = c.tile.tblock < 1 ? 1 : c.tile.tblock;
        b.zblock = c.tile.zblock < 1 ? bz : c.tile.zblock;
        b.ystream = c.tile.ystream ? 1 : 0;
        b.bLy = b.Ly / 2;
        b.bLz = bz;
        b.div_tb = make_fastdiv(uint32_t(b.tblock));
        b.div_zb = make_fastdiv(uint32_t(b.zblock));
        b.hdiv_Ly = make_fastdiv(uint32_t(b.bLy));
        b.hdiv_row = make_fastdiv(uint32_t(b.subLx * b.bLy));
        b.hdiv_nzb = make_fastdiv(uint32_t(bz / b.zblock));
        b.hc_lp = uint32_t(__builtin_ctz(unsigned(kk)));
        b.nblocks = uint32_t((b.nlanes + uint32_t(Bh) - 1) / uint32_t(Bh));
        b.xcd_map = (c.tile.hints & 4) ? 0 : 1;
        Tile t = c.tile;
        t.R = 1;
        b.xcd_gran = xcd_cycle_gran(c.geom, t, kk);
        f = fh;
        shm = lds_h;
        Bl = Bh;
      }
    }
    hipError_t e = f(b, dim3(b.nblocks), dim3(Bl), shm, c.stream);
    if (e != hipSuccess) return fail(err, std::string("launch failed: ") + hipGetErrorString(e)), 1;
  }
  return 0;
}

}  // namespace qmrhs

// Extended C ABI (include/qcu_mrhs.h): packed-field dslash, storage layer and
// the tile autotuner.
#include "qcu_mrhs.h"

#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "api_internal.h"
#include "dslash_dispatch.h"
#include "solver_blas.h"
#include "storage_api.h"

namespace qmrhs {

static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }

int api_fail(const std::string &m)
{
  g_last_error = m;
  return 1;
}

bool geom_from_c(const qcuMrhsGeom *g, Geom *out, std::string *err)
{
  if (!g) {
    if (err) *err = "null geometry";
    return false;
  }
  for (int d = 0; d < 4; ++d) {
    out->L[d] = g->L[d];
    out->grid[d] = g->grid[d] <= 0 ? 1 : g->grid[d];
    if (g->L[d] <= 0 || (g->L[d] & 1)) {
      if (err) *err = "lattice extents must be positive and even";
      return false;
    }
  }
  out->N = g->nColors;
  if (out->N < 2 || out->N > 8) {
    if (err) *err = "nColors must be in 2..8";
    return false;
  }
  return true;
}

int prec_bytes(int prec)
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

long long sites_per_parity(const Geom &g) { return 1LL * g.L[0] / 2 * g.L[1] * g.L[2] * g.L[3]; }

// ---------------------------------------------------------------------------
// Tuning cache: (N, nrhs, prec, L, skip mask) -> tile. With the environment
// variable QCU_MRHS_TUNE_CACHE=<file>, entries for the current GPU
// architecture are read from that text file on first use and every new
// result is appended to it (one line per entry, later lines win):
//   arch N nrhs prec Lx Ly Lz Lt mask R block lds t_block z_block hints occ traversal
namespace {
using TuneKey = std::tuple<int, int, int, int, int, int, int, int>;

std::string device_arch()
{
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
  std::string a = p.gcnArchName;
  const size_t colon = a.find(':');
  return colon == std::string::npos ? a : a.substr(0, colon);
}

const char *tune_file() { return getenv("QCU_MRHS_TUNE_CACHE"); }

// The pinned tile table: QCU_MRHS_TILES, or tiles_<arch>.txt in the directory
// of this library (tiles the autotuner chose on that GPU for the benchmark and
// solver shapes, shipped with the library so that every process uses the same
// tile for them instead of re-tuning near-ties differently each run).
std::string pinned_tile_file(const std::string &arch)
{
  const char *env = getenv("QCU_MRHS_TILES");
  if (env) return env;
  Dl_info info;
  if (!dladdr(reinterpret_cast<void *>(&qcuMrhsVersion), &info) || !info.dli_fname) return "";
  std::string dir = info.dli_fname;
  const size_t slash = dir.rfind('/');
  dir = slash == std::string::npos ? "." : dir.substr(0, slash);
  return dir + "/tiles_" + arch + ".txt";
}

void load_tune_file(std::map<TuneKey, Tile> &m, const std::string &path_s)
{
  const char *path = path_s.c_str();
  if (!path[0]) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  const std::string arch = device_arch();
  char line[512], a[64];
  int v[17];
  // One shape per line; lines that do not parse (comments) are skipped.
  while (fgets(line, sizeof line, f)) {
    if (sscanf(line, "%63s %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d", a, &v[0], &v[1], &v[2], &v[3], &v[4],
               &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12], &v[13], &v[14], &v[15]) != 17)
      continue;
    if (arch != a) continue;
    Tile t;
    t.R = v[8];
    t.block = v[9];
    t.lds = v[10];
    t.tblock = v[11];
    t.zblock = v[12];
    t.hints = v[13];
    t.occ = v[14];
    t.ystream = v[15] & 1;
    t.hcube = (v[15] >> 1) & 1;
    t.trav_hi = v[15] >> 2;
    if (t.trav_hi > 3 || t.trav_hi == 2 || (t.trav_hi && t.ystream)) continue;  // not a defined traversal
    m[TuneKey{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]}] = t;
  }
  fclose(f);
}

void append_tune_file(const TuneKey &k, const Tile &t)
{
  const char *path = tune_file();
  if (!path || !path[0]) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d\n", device_arch().c_str(), std::get<0>(k),
          std::get<1>(k), std::get<2>(k), std::get<3>(k), std::get<4>(k), std::get<5>(k), std::get<6>(k),
          std::get<7>(k), t.R, t.block, t.lds, t.tblock, t.zblock, t.hints, t.occ, tile_traversal(t));
  fclose(f);
}

std::map<TuneKey, Tile> &tune_cache()
{
  static std::map<TuneKey, Tile> m;
  static bool loaded = false;
  if (!loaded) {
    loaded = true;
    load_tune_file(m, pinned_tile_file(device_arch()));  // later lines and the cache file win
    const char *cache = tune_file();
    if (cache) load_tune_file(m, cache);
  }
  return m;
}
std::mutex &tune_mutex()
{
  static std::mutex m;
  return m;
}
TuneKey tune_key(const Geom &g, int prec, int k)
{
  int mask = 0;
  for (int d = 0; d < 4; ++d) mask |= (g.grid[d] > 1 ? 1 : 0) << d;
  return TuneKey{g.N, k, prec, g.L[0], g.L[1], g.L[2], g.L[3], mask};
}
}  // namespace

Tile tile_for(const Geom &g, int prec, int k)
{
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_cache().find(tune_key(g, prec, k));
    if (it != tune_cache().end()) return it->second;
  }
  Tile t = default_tile(g, k, prec);
  // Fall back to a plain tile when the preferred one is not instantiated.
  KernelKey key{SL_PACKED, GL_NATIVE, g.N, t.R, t.lds, 0, prec, gauge_prec_for(prec), prec};
  if (k % t.R || !find_kernel(key)) {
    t.R = 1;
    t.lds = 0;
  }
  // LDS staging only within the per-workgroup budget.
  if (t.lds) {
    const int P = k / t.R;
    const size_t shm = size_t((t.block - 1) / P + 2) * 8 * g.N * g.N * prec_bytes(gauge_prec_for(prec));
    if (shm > size_t(kMaxDynLds)) t.lds = 0;
  }
  return t;
}

Tile user_layout_tile(const Geom &g, int prec, int k)
{
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_cache().find(tune_key(g, prec | QCU_MRHS_TILE_USER_LAYOUT, k));
    if (it != tune_cache().end()) return it->second;
  }
  return Tile{};
}

bool tuned(const Geom &g, int prec, int k)
{
  std::lock_guard<std::mutex> lk(tune_mutex());
  return tune_cache().count(tune_key(g, prec, k)) > 0;
}

void set_tile(const Geom &g, int prec, int k, const Tile &t)
{
  std::lock_guard<std::mutex> lk(tune_mutex());
  tune_cache()[tune_key(g, prec, k)] = t;
  append_tune_file(tune_key(g, prec, k), t);
}

std::vector<Tile> tile_candidates(const Geom &g, int prec, int k)
{
  std::vector<Tile> out;
  for (int R : {1, 2, 4}) {
    if (k % R) continue;
    for (int lds : {0, 1}) {
      KernelKey key{SL_PACKED, GL_NATIVE, g.N, R, lds, 0, prec, gauge_prec_for(prec), prec};
      if (!find_kernel(key)) continue;
      for (int block : {128, 256}) {
        if (lds) {
          const int P = k / R;
          const size_t shm = size_t((block - 1) / P + 2) * 8 * g.N * g.N * prec_bytes(gauge_prec_for(prec));
          if (shm > size_t(kMaxDynLds)) continue;
          // No link reuse inside a workgroup; FP8 has LDS-staging kernels only.
          if (P < 2 && prec != PREC_FP8) continue;
        }
        for (int tb : {1, 2, 4}) {
          if (g.L[3] % tb) continue;
          for (int zb : {0, 1, 2, 4}) {
            if (zb > 0 && (g.L[2] % zb || zb >= g.L[2])) continue;
            if (tb == 1 && zb != 0) continue;  // z blocking only inside time blocks
            for (int hints : {0, 3}) {
              Tile t;
              t.R = R;
              t.lds = lds;
              t.block = block;
              t.tblock = tb;
              t.zblock = zb;
              t.hints = hints;
              out.push_back(t);
            }
          }
        }
        // Blocks streamed along y (t_block x z_block cross-sections).
        for (int tb : {2, 3, 4, 6, 8, 12, 16}) {
          if (g.L[3] % tb || tb >= g.L[3]) continue;
          for (int zb : {2, 3, 4, 6, 8, 12, 16}) {
            if (g.L[2] % zb || zb >= g.L[2]) continue;
            Tile t;
            t.R = R;
            t.lds = lds;
            t.block = block;
            t.tblock = tb;
            t.zblock = zb;
            t.ystream = 1;
            out.push_back(t);
            // The same tile with the XCDs side by side on adjacent blocks.
            t.hints = 8;
            if (xcd_cycle_gran(g, t, k)) out.push_back(t);
          }
        }
        // Column order in sites (traversal 4): columns of cy x cz sites in
        // (y, z), XCD x walking columns 8 g + x along t; only column counts
        // that fill every XCD, each column whole workgroups.
        for (int cy : {2, 3, 4, 6, 8, 12}) {
          for (int cz : {2, 3, 4, 6, 8, 12}) {
            if (g.L[1] % cy || g.L[2] % cz || ((g.L[1] / cy) * (g.L[2] / cz)) % 8) continue;
            if ((1LL * (g.L[0] / 2) * cy * cz * g.L[3] * (k / R)) % block) continue;
            Tile t;
            t.R = R;
            t.lds = lds;
            t.block = block;
            t.tblock = cy;
            t.zblock = cz;
            t.trav_hi = 1;
            out.push_back(t);
          }
        }
      }
    }
  }
  // Hypercube tiles (one wavefront per 2x2x2x2 block of output sites):
  // y-streamed blocks of the block lattice, both XCD orders.
  for (int R : {1, 2, 4}) {
    const int P = k % R ? 0 : k / R;
    if (!P || P > 32 || (P & (P - 1))) continue;
    KernelKey key{SL_PACKED, GL_NATIVE, g.N, R, 1, 0, prec, gauge_prec_for(prec), prec};
    key.hc = 1;
    KernelKey key_kf = key;
    key_kf.kf = k;
    if (!find_kernel(key) && !find_kernel(key_kf)) continue;
    for (int block : {64, 128, 256}) {
      if (block % (8 * P)) continue;
      const size_t shm = size_t(block / P) * 8 * g.N * g.N * prec_bytes(gauge_prec_for(prec));
      if (shm > size_t(kMaxDynLds)) continue;
      for (int tb : {1, 2, 3, 4, 6, 8}) {
        if ((g.L[3] / 2) % tb || tb >= g.L[3] / 2) continue;
        for (int zb : {1, 2, 3, 4, 6}) {
          if ((g.L[2] / 2) % zb || zb >= g.L[2] / 2) continue;
          Tile t;
          t.R = R;
          t.lds = 1;
          t.block = block;
          t.tblock = tb;
          t.zblock = zb;
          t.ystream = 1;
          t.hcube = 1;
          out.push_back(t);
          t.hints = 8;
          if (xcd_cycle_gran(g, t, k)) out.push_back(t);
        }
      }
      // Column order: columns of cy x cz blocks in (y, z), XCD x walking
      // columns 8 g + x along t; only column counts that fill every XCD.
      for (int cy : {1, 2, 3, 4, 6, 8}) {
        for (int cz : {1, 2, 3, 4, 6, 8}) {
          const int by = g.L[1] / 2, bz = g.L[2] / 2;
          if (by % cy || bz % cz || ((by / cy) * (bz / cz)) % 8) continue;
          if ((g.L[0] / 2 * cy * cz) % (block / (8 * P))) continue;
          Tile t;
          t.R = R;
          t.lds = 1;
          t.block = block;
          t.tblock = cy;
          t.zblock = cz;
          t.hcube = 1;
          t.trav_hi = 1;
          out.push_back(t);
        }
      }
    }
  }
  return out;
}

Tile tile_from_c(const qcuMrhsTile &c)
{
  Tile t;
  t.R = c.rhs_per_lane;
  t.block = c.block;
  t.lds = c.link_lds;
  t.tblock = c.t_block;
  t.zblock = c.z_block;
  t.hints = c.cache_hints;
  t.occ = c.occupancy;
  t.ystream = c.traversal & 1;
  t.hcube = (c.traversal >> 1) & 1;
  t.trav_hi = c.traversal >> 2;
  return t;
}

qcuMrhsTile tile_to_c(const Tile &t)
{
  qcuMrhsTile c;
  c.rhs_per_lane = t.R;
  c.block = t.block;
  c.link_lds = t.lds;
  c.t_block = t.tblock;
  c.z_block = t.zblock;
  c.cache_hints = t.hints;
  c.occupancy = t.occ;
  c.traversal = tile_traversal(t);
  return c;
}

bool dslash_ex_call(DslashCall *cp, void *out, int out_prec, const void *in, const void *x, int in_prec, double a,
                    double b, const float *sc_in, const float *sc_x, const float *sc_out, const void *gauge, int nrhs,
                    int parity, int dagger, const qcuMrhsGeom *gc, void *stream, std::string *err)
{
  Geom g;
  if (!geom_from_c(gc, &g, err)) return false;
  DslashCall &c = *cp;
  c.sl = SL_PACKED;
  c.gl = GL_NATIVE;
  c.prec_s = in_prec;
  c.prec_g = gauge_prec_for(in_prec);
  c.prec_o = out_prec;
  c.geom = g;
  c.tile = tile_for(g, in_prec, nrhs);
  c.k = nrhs;
  c.parity = parity;
  c.dagger = dagger;
  c.out = out;
  c.in = in;
  c.x = x;
  c.xa = a;
  c.xb = b;
  c.gauge = gauge;
  c.sc_in = sc_in;
  c.sc_x = sc_x;
  c.sc_out = sc_out;
  c.stream = static_cast<hipStream_t>(stream);
  return true;
}

size_t dslash_ex_scratch_bytes(const DslashCall &c)
{
  // [nblocks][nrhs] per-workgroup partials followed by the reduction's stage buffer
  return (size_t(dslash_num_blocks(c)) + kReduceStages) * size_t(c.k) * sizeof(double);
}

int dslash_ex_launch(const DslashCall &call, double *norm2_out, double *scratch, std::string *err)
{
  DslashCall c = call;
  const long long nb = dslash_num_blocks(c);
  c.partial = norm2_out ? scratch : nullptr;
  if (launch_dslash(c, err)) return 1;
  if (norm2_out) {
    const hipError_t e = blas_reduce(scratch, nb, c.k, norm2_out, scratch + size_t(nb) * c.k, c.stream);
    if (e != hipSuccess) {
      if (err) *err = hipGetErrorString(e);
      return 1;
    }
  }
  return 0;
}

int run_packed_dslash(void *out, const void *in, const void *x, double xa, double xb, const void *gauge, int prec,
                      int k, int parity, int dagger, const Geom &g, const Tile *tile, hipStream_t st, std::string *err)
{
  DslashCall c;
  c.sl = SL_PACKED;
  c.gl = GL_NATIVE;
  c.prec_s = c.prec_o = prec;
  c.prec_g = gauge_prec_for(prec);
  c.geom = g;
  c.tile = tile ? *tile : tile_for(g, prec, k);
  c.k = k;
  c.parity = parity;
  c.dagger = dagger;
  c.out = out;
  c.in = in;
  c.gauge = gauge;
  c.x = x;
  c.xa = xa;
  c.xb = xb;
  c.stream = st;
  return launch_dslash(c, err);
}

}  // namespace qmrhs

using namespace qmrhs;

extern "C" {

const char *qcuMrhsLastError(void) { return g_last_error.c_str(); }

int qcuMrhsVersion(void) { return 10000; }

size_t qcuMrhsSpinorBytes(const qcuMrhsGeom *gc, int nrhs, int prec)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err) || nrhs < 1 || !prec_bytes(prec)) return 0;
  return size_t(sites_per_parity(g)) * 4 * g.N * nrhs * prec_bytes(prec);
}

size_t qcuMrhsGaugeBytes(const qcuMrhsGeom *gc, int prec)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err) || !prec_bytes(prec)) return 0;
  // FP8 spinors go with binary16 links: the same bytes and layout as HALF.
  return size_t(sites_per_parity(g)) * 2 * 8 * g.N * g.N * prec_bytes(gauge_prec_for(prec));
}

int qcuMrhsPackGauge(void *dst, int dst_prec, const void *src, int src_prec, const qcuMrhsGeom *gc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!dst || !src) return api_fail("null gauge pointer");
  if (pack_gauge(dst, gauge_prec_for(dst_prec), src, src_prec, g.N, g.L, static_cast<hipStream_t>(stream), &err))
    return api_fail(err);
  return 0;
}

int qcuMrhsPackSpinors(void *dst, int dst_prec, const void *const *src, int src_prec, int nrhs, const qcuMrhsGeom *gc,
                       void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!dst || !src || nrhs < 1) return api_fail("null pointer or nrhs < 1");
  for (int r = 0; r < nrhs; ++r)
    if (!src[r]) return api_fail("null source spinor");
  if (pack_spinors(dst, dst_prec, src, src_prec, nrhs, g.N, sites_per_parity(g), static_cast<hipStream_t>(stream),
                   &err))
    return api_fail(err);
  return 0;
}

int qcuMrhsUnpackSpinors(void *const *dst, int dst_prec, const void *src, int src_prec, int nrhs, const qcuMrhsGeom *gc,
                         void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!dst || !src || nrhs < 1) return api_fail("null pointer or nrhs < 1");
  for (int r = 0; r < nrhs; ++r)
    if (!dst[r]) return api_fail("null destination spinor");
  if (unpack_spinors(dst, dst_prec, src, src_prec, nrhs, g.N, sites_per_parity(g), static_cast<hipStream_t>(stream),
                     &err))
    return api_fail(err);
  return 0;
}

int qcuMrhsConvert(void *dst, int dst_prec, const void *src, int src_prec, size_t n, void *stream)
{
  std::string err;
  if (!dst || !src) return api_fail("null pointer");
  if (convert_complex(dst, dst_prec, src, src_prec, (long long)n, static_cast<hipStream_t>(stream), &err))
    return api_fail(err);
  return 0;
}

int qcuMrhsShiftSpinor(void *dst, const void *src, int direction, const qcuMrhsGeom *gc, int bpc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!dst || !src) return api_fail("null pointer");
  const int S = int(sites_per_parity(g)), M = 4 * g.N;
  // TO_COALESCE: [S][M] -> [M][S]; TO_NON_COALESCE: [M][S] -> [S][M].
  const int rows = direction == 0 ? S : M, cols = direction == 0 ? M : S;
  if (transpose_batched(dst, src, rows, cols, 1, bpc, static_cast<hipStream_t>(stream), &err)) return api_fail(err);
  return 0;
}

int qcuMrhsShiftGauge(void *dst, const void *src, int direction, const qcuMrhsGeom *gc, int bpc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!dst || !src) return api_fail("null pointer");
  const int S = int(sites_per_parity(g)), M = g.N * g.N;
  const int rows = direction == 0 ? S : M, cols = direction == 0 ? M : S;
  if (transpose_batched(dst, src, rows, cols, 8, bpc, static_cast<hipStream_t>(stream), &err)) return api_fail(err);
  return 0;
}

int qcuMrhsDslash(void *out, const void *in, const void *gauge, int prec, int nrhs, int parity, int dagger,
                  const qcuMrhsGeom *gc, const qcuMrhsTile *tc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  Tile t;
  const Tile *tp = nullptr;
  if (tc) {
    t = tile_from_c(*tc);
    tp = &t;
  }
  if (run_packed_dslash(out, in, nullptr, 0.0, 1.0, gauge, prec, nrhs, parity, dagger, g, tp,
                        static_cast<hipStream_t>(stream), &err))
    return api_fail(err);
  return 0;
}

int qcuMrhsDslashXpay(void *out, const void *in, const void *x, double a, double b, const void *gauge, int prec,
                      int nrhs, int parity, int dagger, const qcuMrhsGeom *gc, const qcuMrhsTile *tc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!x) return api_fail("null xpay operand");
  Tile t;
  const Tile *tp = nullptr;
  if (tc) {
    t = tile_from_c(*tc);
    tp = &t;
  }
  if (run_packed_dslash(out, in, x, a, b, gauge, prec, nrhs, parity, dagger, g, tp, static_cast<hipStream_t>(stream),
                        &err))
    return api_fail(err);
  return 0;
}

int qcuMrhsDslashEx(void *out, int out_prec, const void *in, const void *x, int in_prec, double a, double b,
                    const float *sc_in, const float *sc_x, const float *sc_out, double *norm2_out, const void *gauge,
                    int nrhs, int parity, int dagger, const qcuMrhsGeom *gc, void *stream)
{
  std::string err;
  DslashCall c;
  if (!dslash_ex_call(&c, out, out_prec, in, x, in_prec, a, b, sc_in, sc_x, sc_out, gauge, nrhs, parity, dagger, gc,
                      stream, &err))
    return api_fail(err);
  double *partial = nullptr;
  if (norm2_out && hipMalloc(&partial, dslash_ex_scratch_bytes(c)) != hipSuccess) return api_fail("allocation failed");
  int rc = dslash_ex_launch(c, norm2_out, partial, &err);
  // The partials are freed on return, so the call completes first.
  if (partial) {
    if (hipStreamSynchronize(c.stream) != hipSuccess && !rc) rc = 1, err = "synchronisation failed";
    (void)hipFree(partial);
  }
  if (rc) return api_fail(err);
  return 0;
}

int qcuMrhsDslashUser(void *const *out_user, const void *const *in_user, const void *gauge, int prec, int nrhs,
                      int parity, int dagger, const qcuMrhsGeom *gc, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  DslashCall c;
  c.sl = SL_AOS_PTRS;
  c.gl = GL_NATIVE;
  c.prec_s = c.prec_o = prec;
  c.prec_g = gauge_prec_for(prec);
  c.geom = g;
  c.k = nrhs;
  c.parity = parity;
  c.dagger = dagger;
  c.out_ptrs = out_user;
  c.in_ptrs = in_user;
  c.gauge = gauge;
  c.stream = static_cast<hipStream_t>(stream);
  c.tile = user_layout_tile(g, prec, nrhs);
  if (launch_dslash(c, &err)) return api_fail(err);
  return 0;
}

int qcuMrhsKernelCount(void) { return registry_size(); }

int qcuMrhsGetTile(const qcuMrhsGeom *gc, int prec, int nrhs, qcuMrhsTile *tile)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!tile) return api_fail("null tile");
  *tile = tile_to_c((prec & QCU_MRHS_TILE_USER_LAYOUT) ? user_layout_tile(g, prec & ~QCU_MRHS_TILE_USER_LAYOUT, nrhs)
                                                       : tile_for(g, prec, nrhs));
  return 0;
}

int qcuMrhsSetTile(const qcuMrhsGeom *gc, int prec, int nrhs, const qcuMrhsTile *tile)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!tile) return api_fail("null tile");
  if (nrhs < 1) return api_fail("nrhs must be >= 1");
  const Tile t = tile_from_c(*tile);
  if (t.R != 1 && t.R != 2 && t.R != 4) return api_fail("rhs_per_lane must be 1, 2 or 4");
  if (t.block < 64 || t.block > 256 || t.block % 64) return api_fail("block must be 64, 128, 192 or 256");
  if (t.trav_hi > 3 || t.trav_hi == 2)
    return api_fail("traversal bits 3 and above are defined only as bit 3 with bit 2 (column order split in x)");
  if (t.trav_hi && t.ystream) return api_fail("column order (traversal bit 2) is traversal 4 or 6");
  set_tile(g, prec, nrhs, t);
  return 0;
}

int qcuMrhsTune(void *out, const void *in, const void *gauge, int prec, int nrhs, const qcuMrhsGeom *gc,
                qcuMrhsTile *best, double *best_us, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<Tile> cands = tile_candidates(g, prec, nrhs);
  if (cands.empty()) return api_fail("no instantiated tile for this shape");
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return api_fail("hipEventCreate failed");
  // Time `reps` launches (alternating parity) of a tile; -1: the tile is not
  // valid here, -2: a launch failed on the device.
  auto time_tile = [&](const Tile &t, int reps) -> double {
    // The warm-up launch also validates the tile.
    if (run_packed_dslash(out, in, nullptr, 0.0, 1.0, gauge, prec, nrhs, 0, 0, g, &t, st, &err)) return -1.0;
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r)
      run_packed_dslash(out, in, nullptr, 0.0, 1.0, gauge, prec, nrhs, r & 1, 0, g, &t, st, &err);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) return -2.0;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return 1e3 * ms / reps;
  };
  // Two rounds: every candidate briefly, then the few fastest again with
  // more launches, so that near-ties are decided on a steadier measurement.
  std::vector<std::pair<double, Tile>> timed;
  const char *logenv = getenv("QCU_MRHS_TUNE_LOG");
  const bool log = logenv && logenv[0] && logenv[0] != '0';
  auto log_tile = [&](const char *round, const Tile &t, double us) {
    if (log)
      fprintf(stderr, "qcuMrhsTune %s N=%d k=%d prec=%d tile=%d,%d,%d,%d,%d,%d,%d,%d us=%.2f\n", round, g.N, nrhs,
              prec, t.R, t.block, t.lds, t.tblock, t.zblock, t.hints, t.occ, tile_traversal(t), us);
  };
  for (const Tile &t : cands) {
    const double us = time_tile(t, 5);
    log_tile("first", t, us);
    if (us == -2.0) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return api_fail("tuning launch failed");
    }
    if (us >= 0.0) timed.emplace_back(us, t);
  }
  if (timed.empty()) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return api_fail("no tile launched for this shape: " + err);
  }
  std::stable_sort(timed.begin(), timed.end(),
                   [](const std::pair<double, Tile> &x, const std::pair<double, Tile> &y) { return x.first < y.first; });
  double best_t = 1e30;
  Tile best_tile = timed[0].second;
  const size_t finalists = timed.size() < 6 ? timed.size() : 6;
  for (size_t i = 0; i < finalists; ++i) {
    const double us = time_tile(timed[i].second, 20);
    log_tile("final", timed[i].second, us);
    if (us == -2.0) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return api_fail("tuning launch failed");
    }
    if (us >= 0.0 && us < best_t) {
      best_t = us;
      best_tile = timed[i].second;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  set_tile(g, prec, nrhs, best_tile);
  if (best) *best = tile_to_c(best_tile);
  if (best_us) *best_us = best_t;
  return 0;
}

}  // extern "C"

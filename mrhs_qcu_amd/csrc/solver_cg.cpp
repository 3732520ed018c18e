// Multi-RHS even-odd preconditioned CG on the normal equations.
//
// M^ = 1 - kappa^2 D_eo D_oe on the even checkerboard (D_eo: output parity 0,
// D_oe: output parity 1, D the hopping term of the dslash kernels); M^+ is the
// same with the dagger flag set, because D(s=-1) = g5 D(s=+1) g5. The solver
// returns x with M^+ M^ x = M^+ b for every right-hand side.
//
// Precision scheme (sloppy_prec != DOUBLE):
//   * dslash storage in sloppy_prec (FP16 / BF16 / FP32) with FP32 arithmetic;
//     binary16 fields carry a per-RHS scale (true = stored * scale) chosen
//     from the CG scalars so no value can overflow;
//   * the residual r and M^+M^ p in FP32; the search direction only in its
//     (scaled) sloppy copy, which the operator reads;
//   * all dot products and norms accumulated in FP64 (per-workgroup partials,
//     deterministic reduction);
//   * the solution is FP64 only: the search directions of the last few steps
//     stay stored (a ring of kDirRing fields) with their step lengths, and
//     x += sum alpha_j p_j is applied in FP64 at each reliable update or when
//     the ring is full, so no per-iteration x field is read and rewritten;
//   * the true residual is recomputed from the FP64 operator when the
//     iterated residual fell by delta, every reliable_interval iterations and
//     whenever the iterated residual says converged.
// With sloppy_prec == DOUBLE every field is FP64 and the true residual is
// recomputed only when the iterated residual reaches the tolerance.
// All scalars live on the device; the host reads k flags per check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "api_internal.h"
#include "clover.h"
#include "dslash_dispatch.h"
#include "qcu.h"
#include "qcu_mrhs.h"
#include "solver_blas.h"
#include "storage_api.h"

namespace qmrhs {

// Search directions kept between x updates (each one sloppy field, FP64 in the
// FP64 solve): x += sum alpha_j p_j is applied once per ring instead of
// reading and rewriting an x field every iteration.
constexpr int kDirRing = 8;
static_assert(kDirRing <= kMaxDirSteps, "a flush applies up to a whole ring of steps");

int qcu_context_colors();
hipError_t cg_init_state(const CgState &s, int k, hipStream_t st);
hipError_t cg_check_true(const CgState &s, const double *rr_true, int k, double stop2, hipStream_t st);

namespace {

// Device work vectors kept between solves: a block is handed back to the
// pool at the end of a solve and reused by the next request of the same size.
struct Pool {
  std::vector<std::pair<size_t, void *>> free_blocks;
  std::mutex m;
  void *take(size_t bytes)
  {
    std::lock_guard<std::mutex> lk(m);
    for (size_t i = 0; i < free_blocks.size(); ++i)
      if (free_blocks[i].first == bytes) {
        void *p = free_blocks[i].second;
        free_blocks.erase(free_blocks.begin() + long(i));
        return p;
      }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
  }
  void give(size_t bytes, void *p)
  {
    std::lock_guard<std::mutex> lk(m);
    free_blocks.emplace_back(bytes, p);
  }
  void release()
  {
    std::lock_guard<std::mutex> lk(m);
    for (auto &b : free_blocks) (void)hipFree(b.second);
    free_blocks.clear();
  }
};
Pool &pool()
{
  static Pool *p = new Pool();  // never destroyed: process exit may outlive the HIP runtime
  return *p;
}

struct DevMem {
  std::vector<std::pair<size_t, void *>> blocks;
  ~DevMem()
  {
    for (auto &b : blocks) pool().give(b.first, b.second);
  }
  template <class T>
  T *alloc(size_t count, std::string *err, bool *ok)
  {
    const size_t bytes = count * sizeof(T);
    void *p = pool().take(bytes);
    if (!p) {
      *ok = false;
      if (err && err->empty()) *err = "device allocation failed in the solver";
      return nullptr;
    }
    blocks.emplace_back(bytes, p);
    return static_cast<T *>(p);
  }
  void *bytes(size_t n, std::string *err, bool *ok) { return alloc<unsigned char>(n, err, ok); }
};

struct Solver {
  Geom g;
  int k;
  double kappa;
  hipStream_t st;
  long long S, n;  // sites per parity, complex numbers per field
  double *partial = nullptr;
  size_t partial_cap = 0;  // doubles
  std::string *err;

  bool dslash(void *out, const void *in, const void *x, double xa, double xb, const void *gauge, int prec_in,
              int prec_out, int parity, int dagger, const float *sc_in, const float *sc_x, const float *sc_out,
              bool norm, const DslashCall *fuse = nullptr)
  {
    DslashCall c;
    if (fuse) {
      c.rx_r = fuse->rx_r;
      c.rx_alpha = fuse->rx_alpha;
    }
    c.sl = SL_PACKED;
    c.gl = GL_NATIVE;
    c.prec_s = prec_in;
    c.prec_g = gauge_prec_for(prec_in);
    c.prec_o = prec_out;
    c.geom = g;
    c.tile = tile_for(g, prec_in, k);
    c.k = k;
    c.parity = parity;
    c.dagger = dagger;
    c.out = out;
    c.in = in;
    c.x = x;
    c.xa = xa;
    c.xb = xb;
    c.gauge = gauge;
    c.sc_in = sc_in;
    c.sc_x = sc_x;
    c.sc_out = sc_out;
    c.stream = st;
    if (norm) {
      if (size_t(dslash_num_blocks(c)) * k > partial_cap) {
        *err = "partial buffer too small";
        return false;
      }
      c.partial = partial;
    }
    if (launch_dslash(c, err)) return false;
    if (norm && hipGetLastError() != hipSuccess) return false;
    return true;
  }

  long long dslash_blocks(int prec) const
  {
    DslashCall c;
    c.sl = SL_PACKED;
    c.geom = g;
    c.tile = tile_for(g, prec, k);
    c.k = k;
    return dslash_num_blocks(c);
  }

  // Clover term (both null: Wilson): FP64 fields A and A^-1, and two FP64
  // scratch fields for the clover form of M^.
  const void *cl64 = nullptr, *cli64 = nullptr;
  void *cl_u = nullptr, *cl_a = nullptr;

  bool clover(void *out, const void *in, const void *field, int prec, int parity)
  {
    return hip_ok_s(clover_apply(prec, out, in, field, parity, int(S), k, g.N, st), "clover apply");
  }
  bool hip_ok_s(hipError_t e, const char *what)
  {
    if (e == hipSuccess) return true;
    *err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
  }

  // out = M^ in (dagger 0) or M^+ in (dagger 1) in prec, using tmp (odd parity):
  // Wilson M^ = 1 - kappa^2 D_eo D_oe; with the clover term
  // M^ = A_ee - kappa^2 D_eo A_oo^-1 D_oe (gauge and clover fields in prec).
  bool apply_m(void *out, const void *in, void *tmp, const void *gauge, int dagger, int prec,
               const void *cla = nullptr, const void *clai = nullptr, void *u = nullptr, void *a = nullptr)
  {
    if (!dslash(tmp, in, nullptr, 0.0, 1.0, gauge, prec, prec, 1, dagger, nullptr, nullptr, nullptr, false)) return false;
    if (!cla)
      return dslash(out, tmp, in, 1.0, -kappa * kappa, gauge, prec, prec, 0, dagger, nullptr, nullptr, nullptr, false);
    return clover(u, tmp, clai, prec, 1) && clover(a, in, cla, prec, 0) &&
           dslash(out, u, a, 1.0, -kappa * kappa, gauge, prec, prec, 0, dagger, nullptr, nullptr, nullptr, false);
  }
  bool apply_m64(void *out, const void *in, void *tmp, const void *gauge, int dagger)
  {
    return apply_m(out, in, tmp, gauge, dagger, PREC_DOUBLE, cl64, cli64, cl_u, cl_a);
  }
};

bool hip_ok(hipError_t e, std::string *err, const char *what)
{
  if (e == hipSuccess) return true;
  *err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

}  // namespace


static bool tune_now(void *out, void *in, const void *gauge, int prec, int k, const Geom &g, hipStream_t st,
                     std::string *err)
{
  // Any finite input works for timing; zero it so no stale bits propagate.
  if (hipMemsetAsync(in, 0, size_t(sites_per_parity(g)) * 4 * g.N * k * prec_bytes(prec), st) != hipSuccess) {
    *err = "memset failed";
    return false;
  }
  qcuMrhsGeom gc;
  for (int d = 0; d < 4; ++d) {
    gc.L[d] = g.L[d];
    gc.grid[d] = g.grid[d];
  }
  gc.nColors = g.N;
  qcuMrhsTile best;
  double us = 0;
  if (qcuMrhsTune(out, in, gauge, prec, k, &gc, &best, &us, st) != 0) {
    *err = std::string("tuning failed: ") + qcuMrhsLastError();
    return false;
  }
  return true;
}

int cg_solve(void *x_out, const void *b, const void *gauge64, int k, const Geom &g, qcuMrhsCgParam *prm,
             double *true_res, hipStream_t st, std::string *err, const void *clover64 = nullptr,
             const void *clover_inv64 = nullptr)
{
  const auto t0 = std::chrono::steady_clock::now();
  if (k < 1 || k > kBlasMaxRhs) return *err = "nrhs must be in 1..256", 1;
  const int ps = prm->sloppy_prec;
  if (ps != PREC_DOUBLE && ps != PREC_SINGLE && ps != PREC_HALF && ps != PREC_BF16 && ps != PREC_FP8)
    return *err = "unsupported sloppy precision", 1;
  if (ps == PREC_FP8 && g.N != 3) return *err = "the FP8 sloppy operator is instantiated for nColors = 3", 1;
  if (ps == PREC_FP8 && clover64) return *err = "the clover solve has no FP8 sloppy precision", 1;
  if (!(prm->tol > 0.0)) return *err = "tolerance must be positive", 1;
  const bool clov = clover64 != nullptr;
  if (clov && (!clover_inv64 || g.N > 5)) return *err = "the clover solve needs A and A^-1 and nColors <= 5", 1;
  const bool mixed = ps != PREC_DOUBLE;
  const int pk = mixed ? PREC_SINGLE : PREC_DOUBLE;  // Krylov precision
  Solver s;
  s.g = g;
  s.k = k;
  s.kappa = prm->kappa;
  s.st = st;
  s.err = err;
  s.S = sites_per_parity(g);
  s.n = s.S * 4 * g.N * k;
  const long long n = s.n;
  const int interval = prm->reliable_interval > 0 ? prm->reliable_interval : 50;
  const int max_iter = prm->max_iter > 0 ? prm->max_iter : 10000;
  const double stop2 = prm->tol * prm->tol;

  DevMem mem;
  bool ok = true;
  const size_t b64 = size_t(n) * 16, bk = size_t(n) * prec_bytes(pk), bs = size_t(n) * prec_bytes(ps);
  void *bhat = mem.bytes(b64, err, &ok);
  void *rh = mem.bytes(b64, err, &ok);
  void *t64 = mem.bytes(b64, err, &ok);
  void *ax64 = mem.bytes(b64, err, &ok);
  void *r = mem.bytes(bk, err, &ok);
  // The search direction: FP64 in the FP64 solve; in a mixed solve it lives
  // only in its (scaled) sloppy copy, which the operator reads, so x and r
  // advance along exactly the direction the operator was applied to. It is
  // written out of place into a ring of stored directions: the x steps of the
  // directions since the last update are applied together, in FP64.
  const int dir_prec = mixed ? ps : pk;
  const size_t bdir = mixed ? bs : bk;
  int nring = kDirRing;
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) {
      // at most a quarter of the free memory, at least two directions
      const size_t fit = free_b / 4 / std::max<size_t>(bdir, 1);
      nring = int(std::max<size_t>(2, std::min<size_t>(size_t(kDirRing), fit)));
    }
  }
  std::vector<void *> ring(nring);
  for (int j = 0; j < nring; ++j) ring[j] = mem.bytes(bdir, err, &ok);
  double *ring_alpha = mem.alloc<double>(size_t(nring) * k, err, &ok);
  float *ring_sc = mem.alloc<float>(size_t(nring) * k, err, &ok);
  void *t_s = mixed ? mem.bytes(bs, err, &ok) : t64;
  void *q_s = mixed ? mem.bytes(bs, err, &ok) : ax64;
  void *t2_s = mixed ? mem.bytes(bs, err, &ok) : t64;
  const int pg = gauge_prec_for(ps);  // sloppy links (FP8 spinors: binary16 links)
  void *gs = mixed ? mem.bytes(size_t(s.S) * 2 * 8 * g.N * g.N * prec_bytes(pg), err, &ok) : const_cast<void *>(gauge64);
  const long long nb_blas = blas_num_blocks(n, k);
  // Partial sums for the largest workgroup count any tile can use
  // (one RHS per lane, 64-lane workgroups).
  const long long nb_max = std::max(nb_blas, (s.S * k + 63) / 64);
  s.partial_cap = size_t(nb_max) * k;
  s.partial = mem.alloc<double>(s.partial_cap, err, &ok);
  double *partial_p = mem.alloc<double>(size_t(nb_blas) * k, err, &ok);
  double *stage = mem.alloc<double>(size_t(kReduceStages) * k, err, &ok);
  CgState cs;
  cs.rr = mem.alloc<double>(k, err, &ok);
  cs.rr_new = mem.alloc<double>(k, err, &ok);
  cs.pap = mem.alloc<double>(k, err, &ok);
  cs.alpha = mem.alloc<double>(k, err, &ok);
  cs.beta = mem.alloc<double>(k, err, &ok);
  cs.pnorm2 = mem.alloc<double>(k, err, &ok);
  cs.pnorm = mem.alloc<double>(k, err, &ok);
  cs.b2 = mem.alloc<double>(k, err, &ok);
  cs.rdotp = mem.alloc<double>(k, err, &ok);
  cs.gamma = mem.alloc<double>(k, err, &ok);
  cs.done = mem.alloc<int>(k, err, &ok);
  cs.sloppy_conv = mem.alloc<int>(k, err, &ok);
  cs.sc_p = mem.alloc<float>(k, err, &ok);
  cs.sc_t = mem.alloc<float>(k, err, &ok);
  cs.sc_q = mem.alloc<float>(k, err, &ok);
  cs.sc_t2 = mem.alloc<float>(k, err, &ok);
  cs.sc_p_old = mem.alloc<float>(k, err, &ok);
  // Stored sloppy fields are sized from a bound of their 2-norm: at most
  // 1024 (binary16, max 65504) or 256 (FP8 E4M3, max 448) per RHS for p.
  cs.scale_div = ps == PREC_FP8 ? 256.0f : 1024.0f;
  double *rr_true = mem.alloc<double>(k, err, &ok);
  // Clover: FP64 scratch for M^ in FP64 (also the loop's fields when the
  // solve is all FP64), sloppy copies of A and A^-1 and two sloppy fields.
  const size_t clover_elems = size_t(s.S) * 2 * 8 * g.N * g.N;
  void *u_s = nullptr, *a_s = nullptr;
  const void *cl_s = nullptr, *cli_s = nullptr;
  if (clov) {
    s.cl64 = clover64;
    s.cli64 = clover_inv64;
    s.cl_u = mem.bytes(b64, err, &ok);
    s.cl_a = mem.bytes(b64, err, &ok);
    u_s = mixed ? mem.bytes(bs, err, &ok) : s.cl_u;
    a_s = mixed ? mem.bytes(bs, err, &ok) : s.cl_a;
    if (mixed) {
      void *c1 = mem.bytes(clover_elems * prec_bytes(ps), err, &ok);
      void *c2 = mem.bytes(clover_elems * prec_bytes(ps), err, &ok);
      if (!ok) return 1;
      std::string e2;
      if (convert_complex(c1, ps, clover64, PREC_DOUBLE, clover_elems, st, &e2) ||
          convert_complex(c2, ps, clover_inv64, PREC_DOUBLE, clover_elems, st, &e2))
        return *err = e2, 1;
      cl_s = c1;
      cli_s = c2;
    } else {
      cl_s = clover64;
      cli_s = clover_inv64;
    }
  }
  if (!ok) return 1;
  if (mixed) {
    std::string e2;
    if (convert_complex(gs, pg, gauge64, PREC_DOUBLE, s.S * 2 * 8 * g.N * g.N, st, &e2)) return *err = e2, 1;
  }
  // Tune the operator tiles for this shape once per process (the tuner
  // overwrites the scratch fields it times on; they are initialised below).
  if (!tuned(g, ps, k) && !tune_now(t_s, ring[0], gs, ps, k, g, st, err)) return 1;
  if (mixed && !tuned(g, PREC_DOUBLE, k) && !tune_now(t64, ax64, gauge64, PREC_DOUBLE, k, g, st, err)) return 1;
  // Reliable updates taken while every live residual is still far above FP32
  // resolution (relative residual > 1e-5; the FP32 operator's error is about
  // 1e-7 |b^|) apply M^+M^ in FP32; the later ones, and every convergence
  // decision near tol, use FP64. The FP32 vectors live in the FP64 scratch
  // t64 (x32, t32) and ax64 (rh32); M^+M^ x lands in r.
  const bool fp32_updates = mixed && !clov && prm->tol < 1e-6;
  void *x32 = t64, *t32 = static_cast<char *>(t64) + bk, *rh32 = ax64;
  void *g32 = nullptr;
  if (fp32_updates) {
    if (ps == PREC_SINGLE) {
      g32 = gs;
    } else {
      g32 = mem.bytes(size_t(s.S) * 2 * 8 * g.N * g.N * prec_bytes(PREC_SINGLE), err, &ok);
      if (!ok) return 1;
      std::string e2;
      if (convert_complex(g32, PREC_SINGLE, gauge64, PREC_DOUBLE, s.S * 2 * 8 * g.N * g.N, st, &e2))
        return *err = e2, 1;
    }
    if (!tuned(g, PREC_SINGLE, k) && !tune_now(t32, x32, g32, PREC_SINGLE, k, g, st, err)) return 1;
  }
  const bool scaled = mixed && (ps == PREC_HALF || ps == PREC_FP8);
  const float *sc_p = scaled ? cs.sc_p : nullptr, *sc_t = scaled ? cs.sc_t : nullptr;
  const float *sc_q = scaled ? cs.sc_q : nullptr, *sc_t2 = scaled ? cs.sc_t2 : nullptr;
  const float *sc_p_old = scaled ? cs.sc_p_old : nullptr;
  // The current search direction (ring slot cur), its precision and scale
  // (as read by the operator and by the residual of a reliable update,
  // before cg_beta); `pending` holds the slots whose x step is not applied.
  int cur = 0;
  std::vector<int> pending;
  const int pdir_prec = dir_prec;

#define QM_TRY(expr)                                     \
  do {                                                   \
    if (!(expr)) return (err->empty() ? (*err = "solver step failed: " #expr) : *err), 1; \
  } while (0)
#define QM_HIP(expr) QM_TRY(hip_ok((expr), err, #expr))

  // b^ = M^+ b ; |b^|^2
  QM_TRY(s.apply_m64(bhat, b, t64, gauge64, 1));
  QM_HIP(blas_norm2(PREC_DOUBLE, bhat, s.partial, n, k, st));
  QM_HIP(blas_reduce(s.partial, nb_blas, k, cs.b2, stage, st));
  QM_HIP(cg_init_state(cs, k, st));  // rr = b2, beta = 0, done = 0, scales from |b^|
  // x = 0, r = b^, p = r
  QM_HIP(blas_zero(PREC_DOUBLE, x_out, n, k, st));
  if (mixed) {
    std::string e2;
    if (convert_complex(r, pk, bhat, PREC_DOUBLE, n, st, &e2)) return *err = e2, 1;
  } else {
    QM_HIP(hipMemcpyAsync(r, bhat, b64, hipMemcpyDeviceToDevice, st));
  }
  QM_HIP(blas_zero(pdir_prec, ring[cur], n, k, st));
  QM_HIP(blas_update_p(pk, ps, mixed ? nullptr : ring[cur], mixed ? nullptr : ring[cur], mixed ? ring[cur] : nullptr,
                       mixed ? ring[cur] : nullptr, r, cs.beta, nullptr, nullptr, nullptr, sc_p, sc_p_old, partial_p, n,
                       k, st));
  // x += sum of the pending steps (and, with `step`, the step along the
  // current direction), in FP64.
  auto flush_steps = [&](bool step) -> bool {
    DirSteps ds;
    auto add = [&](int slot) {
      ds.p[ds.count] = ring[slot];
      ds.alpha[ds.count] = ring_alpha + size_t(slot) * k;
      ds.sc[ds.count] = scaled ? ring_sc + size_t(slot) * k : nullptr;
      ++ds.count;
    };
    for (int slot : pending) add(slot);
    if (step) add(cur);
    pending.clear();
    return hip_ok(blas_flush(pdir_prec, x_out, ds, n, k, st), err, "flush");
  };
  // p_new = gamma r + beta p into the next ring slot (gamma null: 1).
  auto next_direction = [&](const double *gamma) -> bool {
    const int nxt = (cur + 1) % nring;
    void *src = ring[cur], *dst = ring[nxt];
    if (!hip_ok(blas_update_p(pk, ps, mixed ? nullptr : dst, mixed ? nullptr : src, mixed ? dst : nullptr,
                              mixed ? src : nullptr, r, cs.beta, gamma, nullptr, nullptr, sc_p, sc_p_old, partial_p, n,
                              k, st),
                err, "update_p"))
      return false;
    cur = nxt;
    return true;
  };

  std::vector<int> done(k, 0);
  std::vector<double> rrt(k, 0.0), b2h(k, 0.0), rr_ru(k, 0.0);
  int iter = 0, updates = 0, since_update = 0;
  bool converged = false;
  const double kk2 = -prm->kappa * prm->kappa;
  // Reliable-update trigger: the iterated residual of some live RHS fell by
  // delta since that RHS's last true residual (the sloppy and the true
  // residual then cannot drift apart), or every live RHS reached the
  // tolerance, or reliable_interval iterations passed.
  const double delta = prm->reliable_delta > 0.0 ? prm->reliable_delta
                       : (ps == PREC_SINGLE ? 1e-3 : (ps == PREC_HALF ? 3e-3 : (ps == PREC_FP8 ? 3e-1 : 1e-1)));
  const double delta2 = mixed ? delta * delta : 0.0;
  // QCU_MRHS_CG_VERBOSE=1 prints the iterated and true residuals.
  const char *verb_env = getenv("QCU_MRHS_CG_VERBOSE");
  const bool verbose = verb_env && verb_env[0] == '1';

  // The iterated residual norms come back through pinned host memory once per
  // iteration; the host decides on a reliable update before the search
  // direction is formed, so an update can re-orthogonalise it.
  struct Pinned {
    void *p = nullptr;
    ~Pinned()
    {
      if (p) (void)hipHostFree(p);
    }
  } pin;
  QM_HIP(hipHostMalloc(&pin.p, size_t(k) * sizeof(double), hipHostMallocDefault));
  double *h_rr = static_cast<double *>(pin.p);

  // True residual: x_h += the pending steps and, with `step`, alpha p (the
  // step of the iteration just finished), r_h = b^ - M^+M^ x_h, |r_h|^2
  // replaces the iterated norm, done flags; the Krylov residual restarts
  // from r_h.
  auto reliable_update = [&](bool step, bool use32) -> bool {
    if (!flush_steps(step)) return false;
    void *pdir = ring[cur];
    // r = b^ - M^+M^ x_h straight into the Krylov vector, with |r|^2 and
    // <r, p> (for re-orthogonalising p) from the same pass.
    if (use32) {
      std::string e2;
      if (convert_complex(x32, PREC_SINGLE, x_out, PREC_DOUBLE, n, st, &e2)) return *err = e2, false;
      if (!s.apply_m(rh32, x32, t32, g32, 0, PREC_SINGLE)) return false;
      if (!s.apply_m(r, rh32, t32, g32, 1, PREC_SINGLE)) return false;
      if (!hip_ok(blas_residual(pk, r, bhat, r, PREC_SINGLE, pdir, pdir_prec, sc_p, s.partial, partial_p, n, k, st),
                  err, "residual"))
        return false;
    } else {
      if (!s.apply_m64(rh, x_out, t64, gauge64, 0)) return false;
      if (!s.apply_m64(ax64, rh, t64, gauge64, 1)) return false;
      if (!hip_ok(blas_residual(pk, r, bhat, ax64, PREC_DOUBLE, pdir, pdir_prec, sc_p, s.partial, partial_p, n, k,
                                st),
                  err, "residual"))
        return false;
    }
    if (!hip_ok(blas_reduce(s.partial, nb_blas, k, rr_true, stage, st), err, "reduce")) return false;
    if (!hip_ok(blas_reduce(partial_p, nb_blas, k, cs.rdotp, stage, st), err, "reduce")) return false;
    if (!hip_ok(cg_check_true(cs, rr_true, k, stop2, st), err, "check")) return false;
    ++updates;
    since_update = 0;
    if (!hip_ok(hipMemcpyAsync(done.data(), cs.done, k * sizeof(int), hipMemcpyDeviceToHost, st), err, "d2h"))
      return false;
    if (!hip_ok(hipMemcpyAsync(rrt.data(), rr_true, k * sizeof(double), hipMemcpyDeviceToHost, st), err, "d2h"))
      return false;
    if (!hip_ok(hipStreamSynchronize(st), err, "sync")) return false;
    converged = true;
    double mx = 0;
    for (int i = 0; i < k; ++i) {
      converged = converged && done[i];
      rr_ru[i] = rrt[i];
      mx = std::max(mx, b2h[i] > 0 ? std::sqrt(rrt[i] / b2h[i]) : 0.0);
    }
    if (verbose)
      fprintf(stderr, "cg iter %d: reliable update %d (%s operator), true relative residual max %.3e\n", iter, updates,
              use32 ? "FP32" : "FP64", mx);
    return true;
  };

  QM_HIP(hipMemcpyAsync(b2h.data(), cs.b2, k * sizeof(double), hipMemcpyDeviceToHost, st));
  QM_HIP(hipMemcpyAsync(done.data(), cs.done, k * sizeof(int), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  rr_ru = b2h;
  converged = true;
  for (int i = 0; i < k; ++i) converged = converged && done[i];

  // r -= alpha Ap is fused into the last operator launch of an iteration;
  // x += alpha p waits in the ring (the direction stays stored with its alpha
  // and scale) until a reliable update or a full ring applies it in FP64.
  DslashCall fuse;
  fuse.rx_r = r;
  fuse.rx_alpha = cs.alpha;
  while (!converged && iter < max_iter) {
    // t = D_oe p ; q = p - kappa^2 D_eo t (|q|^2 = <p, M^+M^ p>). Clover:
    // q = A_ee p - kappa^2 D_eo A_oo^-1 t (the clover applications keep the
    // per-RHS storage scale of their input, the operator being linear).
    void *p_cur = ring[cur];
    QM_TRY(s.dslash(t_s, p_cur, nullptr, 0.0, 1.0, gs, ps, ps, 1, 0, sc_p, nullptr, sc_t, false));
    if (clov) {
      QM_TRY(s.clover(u_s, t_s, cli_s, ps, 1) && s.clover(a_s, p_cur, cl_s, ps, 0));
      QM_TRY(s.dslash(q_s, u_s, a_s, 1.0, kk2, gs, ps, ps, 0, 0, sc_t, sc_p, sc_q, true));
    } else {
      QM_TRY(s.dslash(q_s, t_s, p_cur, 1.0, kk2, gs, ps, ps, 0, 0, sc_t, sc_p, sc_q, true));
    }
    QM_HIP(blas_reduce(s.partial, s.dslash_blocks(ps), k, cs.pap, stage, st));
    QM_HIP(blas_reduce(partial_p, nb_blas, k, cs.pnorm2, stage, st));
    QM_HIP(cg_alpha(cs, k, st));
    // this step's alpha and the direction's storage scale, kept with its slot
    QM_HIP(hipMemcpyAsync(ring_alpha + size_t(cur) * k, cs.alpha, k * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (scaled)
      QM_HIP(hipMemcpyAsync(ring_sc + size_t(cur) * k, cs.sc_p, k * sizeof(float), hipMemcpyDeviceToDevice, st));
    // Ap = q - kappa^2 D^+_eo D^+_oe q
    QM_TRY(s.dslash(t2_s, q_s, nullptr, 0.0, 1.0, gs, ps, ps, 1, 1, sc_q, nullptr, sc_t2, false));
    // Ap = q - kappa^2 D^+_eo t2 stays in registers: the same launch does
    // r -= alpha Ap and the per-RHS |r|^2 partials (out unused). Clover:
    // Ap = A_ee q - kappa^2 D^+_eo A_oo^-1 t2.
    if (clov) {
      QM_TRY(s.clover(u_s, t2_s, cli_s, ps, 1) && s.clover(a_s, q_s, cl_s, ps, 0));
      QM_TRY(s.dslash(r, u_s, a_s, 1.0, kk2, gs, ps, pk, 0, 1, sc_t2, sc_q, nullptr, true, &fuse));
    } else {
      QM_TRY(s.dslash(r, t2_s, q_s, 1.0, kk2, gs, ps, pk, 0, 1, sc_t2, sc_q, nullptr, true, &fuse));
    }
    QM_HIP(blas_reduce(s.partial, s.dslash_blocks(ps), k, cs.rr_new, stage, st));
    QM_HIP(hipMemcpyAsync(h_rr, cs.rr_new, k * sizeof(double), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    ++iter;
    ++since_update;
    bool all_conv = true, drop = false;
    double mx = 0, mn2 = 1e300;
    for (int i = 0; i < k; ++i) {
      if (done[i]) continue;
      all_conv = all_conv && h_rr[i] <= stop2 * b2h[i];
      drop = drop || (h_rr[i] < delta2 * rr_ru[i]);
      mx = std::max(mx, b2h[i] > 0 ? std::sqrt(h_rr[i] / b2h[i]) : 0.0);
      if (b2h[i] > 0) mn2 = std::min(mn2, h_rr[i] / b2h[i]);
    }
    if (verbose && iter % 10 == 0) fprintf(stderr, "cg iter %d: iterated relative residual max %.3e\n", iter, mx);
    if (all_conv || drop || (mixed && since_update >= interval) || iter == max_iter) {
      QM_TRY(reliable_update(true, fp32_updates && !all_conv && mn2 > 1e-10));
      if (converged || iter == max_iter) break;
      // beta = |r_true|^2 / |r_old|^2 ; p = gamma r + beta p with the old p
      // made orthogonal to r_true.
      QM_HIP(cg_beta(cs, k, stop2, st));
      QM_HIP(cg_orth(cs, k, st));
      QM_TRY(next_direction(cs.gamma));
    } else {
      // beta = |r|^2 / |r_old|^2 ; p = r + beta p into the next slot; this
      // step's x update waits in the ring (applied before the slot is reused)
      QM_HIP(cg_beta(cs, k, stop2, st));
      pending.push_back(cur);
      QM_TRY(next_direction(nullptr));
      if (int(pending.size()) >= nring - 1) QM_TRY(flush_steps(false));
    }
  }
  // Every iteration ends in a reliable update, which applies every pending x
  // step, or leaves its step pending in the ring: apply what is left.
  if (!pending.empty()) QM_TRY(flush_steps(false));
  if (!converged && since_update > 0) QM_TRY(reliable_update(false, false));
  QM_HIP(hipMemcpyAsync(rrt.data(), rr_true, k * sizeof(double), hipMemcpyDeviceToHost, st));
  QM_HIP(hipMemcpyAsync(b2h.data(), cs.b2, k * sizeof(double), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  if (true_res)
    for (int i = 0; i < k; ++i) true_res[i] = b2h[i] > 0 ? std::sqrt(rrt[i] / b2h[i]) : 0.0;
  prm->iterations = iter;
  prm->reliable_updates = updates;
  prm->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
#undef QM_TRY
#undef QM_HIP
  return 0;
}

}  // namespace qmrhs

using namespace qmrhs;

extern "C" {

// qcu.h cg_inverter (a stub in the reference, src/qcu.cu:173): solves the full
// Wilson system (1 - kappa D) x = b for one right-hand side. x and b are full
// spinors in the user layout, even checkerboard [S][4][N] followed by the odd
// one; gauge is the user AoS gauge [4][2][S][N][N]; all complex<double>.
// Schur complement on the even sites:
//   (1 - kappa^2 D_eo D_oe) x_e = b_e + kappa D_eo b_o,  x_o = b_o + kappa D_oe x_e,
// the first solved by the mixed-precision CG above to relative residual
// p_max_prec of its normal equations.
void cg_inverter(void *x_vector, void *b_vector, void *gauge, QcuParam *param, double p_max_prec, double p_kappa)
{
  auto die = [](const std::string &m) {
    fprintf(stderr, "cg_inverter: %s\n", m.c_str());
    exit(1);
  };
  if (!x_vector || !b_vector || !gauge || !param) die("null argument");
  Geom g;
  for (int d = 0; d < 4; ++d) {
    g.L[d] = param->lattice_size[d];
    g.grid[d] = 1;
    if (g.L[d] <= 0 || (g.L[d] & 1)) die("lattice sizes must be positive and even");
  }
  g.N = qcu_context_colors();
  const long long S = sites_per_parity(g);
  const size_t half = size_t(S) * 4 * g.N * 16;
  std::string err;
  void *gn = nullptr, *bh = nullptr;
  if (hipMalloc(&gn, size_t(S) * 2 * 8 * g.N * g.N * 16) != hipSuccess || hipMalloc(&bh, half) != hipSuccess)
    die("device allocation failed");
  hipStream_t st = nullptr;
  char *x = static_cast<char *>(x_vector), *b = static_cast<char *>(b_vector);
  if (pack_gauge(gn, PREC_DOUBLE, gauge, PREC_DOUBLE, g.N, g.L, st, &err)) die(err);
  // b^_e = b_e + kappa D_eo b_o
  if (run_packed_dslash(bh, b + half, b, 1.0, p_kappa, gn, PREC_DOUBLE, 1, 0, 0, g, nullptr, st, &err)) die(err);
  qcuMrhsCgParam prm{};
  prm.kappa = p_kappa;
  prm.tol = p_max_prec;
  prm.max_iter = 100000;
  prm.reliable_interval = 50;
  prm.sloppy_prec = PREC_HALF;
  double res = 0.0;
  if (cg_solve(x, bh, gn, 1, g, &prm, &res, st, &err)) die(err);
  // x_o = b_o + kappa D_oe x_e
  if (run_packed_dslash(x + half, x, b + half, 1.0, p_kappa, gn, PREC_DOUBLE, 1, 1, 0, g, nullptr, st, &err))
    die(err);
  if (hipStreamSynchronize(st) != hipSuccess) die("synchronisation failed");
  (void)hipFree(gn);
  (void)hipFree(bh);
  printf("cg_inverter: %d iterations, %d reliable updates, true relative residual %.3e, %.3f s\n", prm.iterations,
         prm.reliable_updates, res, prm.seconds);
}

int qcuMrhsSolverRelease(void)
{
  pool().release();
  return 0;
}

int qcuMrhsCgSolve(void *x, const void *b, const void *gauge_f64, int nrhs, const qcuMrhsGeom *gc, qcuMrhsCgParam *p,
                   double *true_res, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!x || !b || !gauge_f64 || !p) return api_fail("null argument");
  if (cg_solve(x, b, gauge_f64, nrhs, g, p, true_res, static_cast<hipStream_t>(stream), &err)) return api_fail(err);
  return 0;
}

int qcuMrhsCgSolveClover(void *x, const void *b, const void *gauge_f64, const void *clover_f64,
                         const void *clover_inv_f64, int nrhs, const qcuMrhsGeom *gc, qcuMrhsCgParam *p,
                         double *true_res, void *stream)
{
  Geom g;
  std::string err;
  if (!geom_from_c(gc, &g, &err)) return api_fail(err);
  if (!x || !b || !gauge_f64 || !clover_f64 || !clover_inv_f64 || !p) return api_fail("null argument");
  if (cg_solve(x, b, gauge_f64, nrhs, g, p, true_res, static_cast<hipStream_t>(stream), &err, clover_f64,
               clover_inv_f64))
    return api_fail(err);
  return 0;
}

}  // extern "C"

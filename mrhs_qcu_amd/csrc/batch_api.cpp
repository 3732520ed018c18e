// Batched hopping term: several independent problems (own lattice, colour
// count N, nrhs, parity, fields) applied by one replayable HIP graph.
//
// Create tunes every problem shape that has no tile yet (timing it on the
// problem's own fields, so `out` is scratch until the first Run) and freezes
// one tile per problem. When every problem has N = 2..5 (the colour counts
// the work-list kernel is compiled for), the graph holds ONE kernel node: a
// work-list launch whose workgroups cover all problems, largest first, each
// problem in its own range with its own tile traversal (batch_kernels.h), so
// the small problems run in the tail of the large ones instead of after it.
// Otherwise the graph chains one kernel node per problem, largest first.
// Run is a single hipGraphLaunch. The graph holds the field pointers given at
// Create: the caller keeps those buffers alive and unmoved until Destroy.
// Nothing in the graph clears memory (no memset nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "api_internal.h"
#include "batch_kernels.h"
#include "qcu_mrhs.h"

namespace qmrhs {
namespace {

struct Batch {
  std::vector<qcuMrhsProblem> probs;
  std::vector<Tile> tiles;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BatchItem *items = nullptr;  // device copy of the work list (work-list graphs)
  ~Batch()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (items) (void)hipFree(items);
  }
};

// The capture stream exists only while the graph is captured.
struct CaptureScratch {
  hipStream_t cap = nullptr;
  bool capturing = false;
  ~CaptureScratch()
  {
    if (capturing) {
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(cap, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    }
    if (cap) (void)hipStreamDestroy(cap);
  }
};

DslashCall problem_call(const qcuMrhsProblem &p, const Geom &g, const Tile &t, hipStream_t st)
{
  DslashCall c;
  c.sl = SL_PACKED;
  c.gl = GL_NATIVE;
  c.prec_s = c.prec_g = c.prec_o = PREC_SINGLE;
  c.geom = g;
  c.tile = t;
  c.k = p.nrhs;
  c.parity = p.parity;
  c.dagger = p.dagger;
  c.out = p.out;
  c.in = p.in;
  c.gauge = p.gauge;
  c.stream = st;
  return c;
}

}  // namespace
}  // namespace qmrhs

using namespace qmrhs;

extern "C" {

int qcuMrhsBatchCreate(const qcuMrhsProblem *probs, int nprob, void **handle, void *stream)
{
  if (!handle) return api_fail("null handle pointer");
  *handle = nullptr;
  if (!probs || nprob < 1) return api_fail("a batch needs at least one problem");
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto b = new Batch();
  std::string err;
  std::vector<Geom> geoms(nprob);
  for (int i = 0; i < nprob; ++i) {
    const qcuMrhsProblem &p = probs[i];
    if (!geom_from_c(&p.geom, &geoms[i], &err)) {
      delete b;
      return api_fail("problem " + std::to_string(i) + ": " + err);
    }
    if (p.nrhs < 1 || !p.out || !p.in || !p.gauge) {
      delete b;
      return api_fail("problem " + std::to_string(i) + ": null field or nrhs < 1");
    }
    // Tune shapes seen for the first time (on this problem's own fields).
    if (!tuned(geoms[i], PREC_SINGLE, p.nrhs)) {
      qcuMrhsTile best;
      double us = 0;
      if (qcuMrhsTune(p.out, p.in, p.gauge, PREC_SINGLE, p.nrhs, &p.geom, &best, &us, stream)) {
        const std::string m = qcuMrhsLastError();
        delete b;
        return api_fail("problem " + std::to_string(i) + ": tuning failed: " + m);
      }
    }
    b->probs.push_back(p);
    b->tiles.push_back(tile_for(geoms[i], PREC_SINGLE, p.nrhs));
  }
  // The work list needs every problem in the compiled tile family (RHS per
  // lane and link staging per N) with one workgroup size; the traversal of
  // each problem's tuned tile is kept.
  bool worklist = true;
  for (int i = 0; i < nprob && worklist; ++i) {
    int R = 0, lds = 0;
    worklist = batch_body_tile(geoms[i].N, &R, &lds) && b->probs[i].nrhs % R == 0;
    if (worklist) {
      Tile &t = b->tiles[i];
      t.R = R;
      t.lds = lds;
      t.block = kBatchBlock;
      t.occ = 0;
      // The work-list body is the site-major kernel: a column-order
      // hypercube tile runs its columns in sites (else y-streamed blocks);
      // other blockings still divide the lattice.
      t = site_major_tile(t, geoms[i], b->probs[i].nrhs, SL_PACKED);
    }
  }
  if (!worklist)
    for (int i = 0; i < nprob; ++i) b->tiles[i] = tile_for(geoms[i], PREC_SINGLE, b->probs[i].nrhs);
  // Every launch is validated eagerly first, so a bad problem fails here and
  // not inside the capture.
  for (int i = 0; i < nprob; ++i) {
    if (launch_dslash(problem_call(b->probs[i], geoms[i], b->tiles[i], st), &err)) {
      delete b;
      return api_fail("problem " + std::to_string(i) + ": " + err);
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    delete b;
    return api_fail("batch validation launches failed");
  }
  // Launch order: largest problem (sites x RHS x N^2) first.
  std::vector<int> order(nprob);
  for (int i = 0; i < nprob; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const double wa = double(sites_per_parity(geoms[a])) * b->probs[a].nrhs * geoms[a].N * geoms[a].N;
    const double wc = double(sites_per_parity(geoms[c])) * b->probs[c].nrhs * geoms[c].N * geoms[c].N;
    return wa > wc;
  });
  {
    CaptureScratch cs;
    bool ok = hipStreamCreateWithFlags(&cs.cap, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
      delete b;
      return api_fail("batch: stream creation failed");
    }
    // Work list: the problems' ranges in launch order, each starting at a
    // multiple of the XCD count; the dynamic LDS is the largest request.
    std::vector<BatchItem> items;
    uint32_t total = 0;
    size_t shm_max = 0;
    for (int j = 0; worklist && ok && j < nprob; ++j) {
      const int i = order[j];
      BatchItem it{};
      size_t shm = 0;
      if (packed_launch_args(problem_call(b->probs[i], geoms[i], b->tiles[i], cs.cap), &it.a, &shm, &err)) {
        ok = false;
        break;
      }
      it.first = total;
      it.n = geoms[i].N;
      total += (it.a.ngrid + kNumXcd - 1) / kNumXcd * kNumXcd;
      shm_max = std::max(shm_max, shm);
      items.push_back(it);
    }
    if (worklist && ok) {
      ok = hipMalloc(&b->items, items.size() * sizeof(BatchItem)) == hipSuccess &&
           hipMemcpy(b->items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice) == hipSuccess;
      if (!ok) err = "work-list allocation failed";
    }
    ok = ok && hipStreamBeginCapture(cs.cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
    cs.capturing = ok;
    if (worklist) {
      if (ok) ok = launch_batch_worklist(b->items, int(items.size()), total, shm_max, cs.cap) == hipSuccess;
    } else {
      for (int j = 0; ok && j < nprob; ++j) {
        const int i = order[j];
        ok = !launch_dslash(problem_call(b->probs[i], geoms[i], b->tiles[i], cs.cap), &err);
      }
    }
    hipGraph_t g = nullptr;
    if (cs.capturing) {
      const bool ended = hipStreamEndCapture(cs.cap, &g) == hipSuccess;
      cs.capturing = false;
      ok = ok && ended && g;
      if (!ok && g) (void)hipGraphDestroy(g);
    }
    if (!ok) {
      delete b;
      return api_fail("batch: graph capture failed" + (err.empty() ? std::string() : ": " + err));
    }
    b->graph = g;
  }
  if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
    delete b;
    return api_fail("batch: graph instantiation failed");
  }
  *handle = b;
  return 0;
}

int qcuMrhsBatchRun(void *handle, void *stream)
{
  if (!handle) return api_fail("null batch handle");
  Batch *b = static_cast<Batch *>(handle);
  const hipError_t e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return api_fail(std::string("batch launch failed: ") + hipGetErrorString(e));
  return 0;
}

int qcuMrhsBatchGetTile(void *handle, int i, qcuMrhsTile *tile)
{
  if (!handle || !tile) return api_fail("null batch handle or tile");
  Batch *b = static_cast<Batch *>(handle);
  if (i < 0 || i >= int(b->tiles.size())) return api_fail("problem index out of range");
  *tile = tile_to_c(b->tiles[i]);
  return 0;
}

int qcuMrhsBatchDestroy(void *handle)
{
  if (!handle) return api_fail("null batch handle");
  delete static_cast<Batch *>(handle);
  return 0;
}

}  // extern "C"

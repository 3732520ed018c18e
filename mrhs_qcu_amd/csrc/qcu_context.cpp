// The qcu.h boundary: process-wide context, Dirac operator, runtime dispatch
// and the reference-semantics layout shifts.
//
// Call protocol and semantics follow the reference (src/qcu.cu:17-181):
//   initGridSize -> loadQcuGauge -> getDslash -> pushBackFermions x k -> startDslash
// with the input of each pushBackFermions consumed at push time and all k
// outputs complete when startDslash returns. Misuse prints a message and
// exits with status 1 (reference include/qcu_macro.cuh:35-40).
//
// What differs by design: startDslash runs ONE launch for all k right-hand
// sides directly on the user buffers (user AoS layout, native gauge), with no
// per-call allocation, no layout shift passes and one stream synchronisation,
// where the reference allocates, shifts and synchronises per RHS
// (src/qcu.cu:58-101). Any k >= 1 and N in 2..8 is accepted.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "api_internal.h"
#include "dslash_dispatch.h"
#include "qcu.h"
#include "qcuDesc/qcu_desc.h"
#include "qcu_dirac.h"
#include "qcu_enum.h"
#include "qcu_kernel_param.cuh"
#include "qcu_macro.cuh"
#include "storage/qcu_storage.h"
#include "storage_api.h"

namespace qcu {

namespace {

int to_prec(QCU_PRECISION p)
{
  switch (p) {
    case QCU_HALF_PRECISION: return qmrhs::PREC_HALF;
    case QCU_SINGLE_PRECISION: return qmrhs::PREC_SINGLE;
    case QCU_DOUBLE_PRECISION: return qmrhs::PREC_DOUBLE;
    default: return -1;
  }
}

void fatal(const std::string &m)
{
  fprintf(stderr, "%s\n", m.c_str());
  exit(1);
}

}  // namespace

class Qcu {
  QcuLattDesc<4> lattDesc_;
  QcuProcDesc<4> procDesc_;
  Dirac *dirac_;
  int nColors_;
  int diracPrec_;  // storage precision of the current Dirac operator (-1: none)
  // Arguments of the last getDslash, to rebuild the operator when the gauge
  // copies it points to are replaced by a new loadQcuGauge.
  DSLASH_TYPE diracType_ = DSLASH_WILSON;
  double diracMass_ = 0.0;
  QCU_PRECISION diracFloatPrec_ = QCU_DOUBLE_PRECISION;
  std::unique_ptr<ExecutionStreams> streams_;

  void *gaugeSoA64_ = nullptr;              // reference SoA copy, FP64 (the reference's localGauge_)
  std::map<int, void *> gaugeSoAByPrec_;    // SoA copies in other precisions (Dirac operators)
  std::map<int, void *> gaugeNativeByPrec_; // native layout per precision (startDslash)

  std::vector<void *> outQueue_;
  std::vector<void *> stagingQueue_;
  std::vector<void *> stagingPool_;
  size_t stagingBytes_ = 0;

 public:
  Qcu(QcuParam *pLattDesc, QcuGrid *pProcDesc, int nColors)
      : lattDesc_(pLattDesc), procDesc_(pProcDesc), dirac_(nullptr), nColors_(nColors), diracPrec_(-1)
  {
  }

  ~Qcu()
  {
    // Also runs from the library destructor at process exit, possibly after
    // the HIP runtime is gone: failures are ignored here.
    delete dirac_;
    dirac_ = nullptr;
    auto release = [](void *p) {
      if (p) (void)hipFree(p);
    };
    release(gaugeSoA64_);
    for (auto &kv : gaugeSoAByPrec_) release(kv.second);
    for (auto &kv : gaugeNativeByPrec_) release(kv.second);
    for (void *p : stagingQueue_) release(p);
    for (void *p : stagingPool_) release(p);
  }

  ExecutionStreams &streams()
  {
    if (!streams_) streams_.reset(new ExecutionStreams());
    return *streams_;
  }

  qmrhs::Geom geom() const
  {
    qmrhs::Geom g;
    for (int d = 0; d < 4; ++d) {
      g.L[d] = lattDesc_.dim[d];
      g.grid[d] = procDesc_.dim[d];
    }
    g.N = nColors_;
    return g;
  }
  long long halfVol() const { return 1LL * lattDesc_.dim[0] / 2 * lattDesc_.dim[1] * lattDesc_.dim[2] * lattDesc_.dim[3]; }

  void loadGauge(void *gauge)
  {
    if (!gauge) errorQcu("gauge pointer is null.\n");
    hipStream_t st = streams().stream;
    const long long S = halfVol();
    const int NN = nColors_ * nColors_;
    if (!gaugeSoA64_) CHECK_HIP(hipMalloc(&gaugeSoA64_, size_t(2 * S) * 4 * NN * 16));
    std::string err;
    if (qmrhs::transpose_batched(gaugeSoA64_, gauge, int(S), NN, 8, 16, st, &err)) fatal(err);
    // Converted copies are rebuilt lazily from the new field.
    for (auto &kv : gaugeSoAByPrec_) (void)hipFree(kv.second);
    gaugeSoAByPrec_.clear();
    void *native = nativeBuffer(qmrhs::PREC_DOUBLE);
    if (qmrhs::pack_gauge(native, qmrhs::PREC_DOUBLE, gauge, qmrhs::PREC_DOUBLE, nColors_, geom().L, st, &err))
      fatal(err);
    for (auto &kv : gaugeNativeByPrec_)
      if (kv.first != qmrhs::PREC_DOUBLE) (void)hipFree(kv.second);
    void *keep = gaugeNativeByPrec_[qmrhs::PREC_DOUBLE];
    gaugeNativeByPrec_.clear();
    gaugeNativeByPrec_[qmrhs::PREC_DOUBLE] = keep;
    CHECK_HIP(hipStreamSynchronize(st));
    // An operator made by an earlier getDslash points to a converted copy
    // freed above: rebuild it on the new field.
    if (dirac_) makeDirac();
  }

  void makeDirac()
  {
    delete dirac_;
    dirac_ = nullptr;
    dirac_ = new DiracWilson(soaGauge(diracPrec_), diracMass_, nColors_, diracFloatPrec_, lattDesc_, procDesc_,
                             diracType_, streams());
  }

  // The current operator (null before getDslash).
  Dirac *dirac() { return dirac_; }

  void *nativeBuffer(int prec)
  {
    auto it = gaugeNativeByPrec_.find(prec);
    if (it != gaugeNativeByPrec_.end()) return it->second;
    void *p = nullptr;
    CHECK_HIP(hipMalloc(&p, size_t(halfVol()) * 2 * 8 * nColors_ * nColors_ * qmrhs::prec_bytes(prec)));
    gaugeNativeByPrec_[prec] = p;
    return p;
  }

  void *nativeGauge(int prec)
  {
    auto it = gaugeNativeByPrec_.find(prec);
    if (it != gaugeNativeByPrec_.end()) return it->second;
    void *src = gaugeNativeByPrec_.at(qmrhs::PREC_DOUBLE);
    void *dst = nativeBuffer(prec);
    std::string err;
    const long long n = halfVol() * 2 * 8 * nColors_ * nColors_;
    if (qmrhs::convert_complex(dst, prec, src, qmrhs::PREC_DOUBLE, n, streams().stream, &err)) fatal(err);
    return dst;
  }

  void *soaGauge(int prec)
  {
    if (prec == qmrhs::PREC_DOUBLE) return gaugeSoA64_;
    auto it = gaugeSoAByPrec_.find(prec);
    if (it != gaugeSoAByPrec_.end()) return it->second;
    void *p = nullptr;
    const long long n = halfVol() * 2 * 4 * nColors_ * nColors_;
    CHECK_HIP(hipMalloc(&p, size_t(n) * qmrhs::prec_bytes(prec)));
    std::string err;
    if (qmrhs::convert_complex(p, prec, gaugeSoA64_, qmrhs::PREC_DOUBLE, n, streams().stream, &err)) fatal(err);
    CHECK_HIP(hipStreamSynchronize(streams().stream));
    gaugeSoAByPrec_[prec] = p;
    return p;
  }

  void getDslash(DSLASH_TYPE dslashType, double mass, int nColors, int nInputs, QCU_PRECISION floatPrecision)
  {
    (void)nInputs;  // the RHS count is the number of pushed fermions (reference src/qcu.cu:64)
    if (gaugeSoA64_ == nullptr) errorQcu("localGauge_ is nullptr, not loaded.\n");
    if (dslashType != DSLASH_WILSON) errorQcu("Now we only support Wilson dslash.\n");
    if (nColors != nColors_) errorQcu("nColors differs from the one given to initGridSize.\n");
    const int prec = to_prec(floatPrecision);
    if (prec < 0) errorQcu("Now we only support half, single and double precision.\n");
    if (!stagingQueue_.empty() && prec != diracPrec_)
      errorQcu("precision changed while right-hand sides are queued.\n");
    diracPrec_ = prec;
    diracType_ = dslashType;
    diracMass_ = mass;
    diracFloatPrec_ = floatPrecision;
    makeDirac();
  }

  void pushFermion(void *fermionOut, void *fermionIn)
  {
    if (!fermionOut || !fermionIn) errorQcu("fermion pointer is null.\n");
    const int prec = diracPrec_ >= 0 ? diracPrec_ : qmrhs::PREC_DOUBLE;
    const size_t bytes = size_t(halfVol()) * Ns * nColors_ * qmrhs::prec_bytes(prec);
    if (bytes != stagingBytes_) {
      for (void *p : stagingPool_) CHECK_HIP(hipFree(p));
      stagingPool_.clear();
      stagingBytes_ = bytes;
    }
    void *buf;
    if (!stagingPool_.empty()) {
      buf = stagingPool_.back();
      stagingPool_.pop_back();
    } else {
      CHECK_HIP(hipMalloc(&buf, bytes));
    }
    hipStream_t st = streams().stream;
    std::string err;
    if (qmrhs::copy_device(buf, fermionIn, bytes, st, &err)) fatal(err);
    CHECK_HIP(hipStreamSynchronize(st));
    stagingQueue_.push_back(buf);
    outQueue_.push_back(fermionOut);
  }

  void startDslash(int parity, int daggerFlag)
  {
    if (dirac_ == nullptr) errorQcu("Dslash operator not created: call getDslash first.\n");
    if (outQueue_.empty()) errorQcu("No fermion pushed: call pushBackFermions first.\n");
    if (parity != 0 && parity != 1) errorQcu("parity must be 0 or 1.\n");
    hipStream_t st = streams().stream;
    qmrhs::DslashCall c;
    c.sl = qmrhs::SL_AOS_PTRS;
    c.gl = qmrhs::GL_NATIVE;
    c.prec_s = c.prec_g = c.prec_o = diracPrec_;
    c.geom = geom();
    c.k = int(outQueue_.size());
    c.parity = parity;
    c.dagger = daggerFlag ? 1 : 0;
    c.out_ptrs = outQueue_.data();
    std::vector<const void *> ins(stagingQueue_.begin(), stagingQueue_.end());
    c.in_ptrs = ins.data();
    c.gauge = nativeGauge(diracPrec_);
    c.stream = st;
    c.tile = qmrhs::user_layout_tile(c.geom, diracPrec_, c.k);
    std::string err;
    if (qmrhs::launch_dslash(c, &err)) fatal(err);
    CHECK_HIP(hipStreamSynchronize(st));
    for (void *p : stagingQueue_) stagingPool_.push_back(p);
    stagingQueue_.clear();
    outQueue_.clear();
  }
};

// ---------------------------------------------------------------------------
// Dirac operator and runtime dispatch (reference src/qcu_dirac.cu:7-14,
// src/qcd/qcu_wilson.cu:159-173).
void DiracWilson::Dslash(void *outputMRHS, void *inputMRHS, int parity, int daggerFlag, int nInputs)
{
  instantiateDslash(lattDesc_, procDesc_, parity, daggerFlag, kappa_, inputMRHS, outputMRHS, gauge_, dslashType_,
                    floatPrecision_, nColors_, nInputs, streams_);
}

void DiracWilson::DslashXpay(void *outputMRHS, void *inputMRHS, void *xMRHS, double a, double b, int parity,
                             int daggerFlag, int nInputs)
{
  instantiateDslashXpay(lattDesc_, procDesc_, parity, daggerFlag, a, b, inputMRHS, xMRHS, outputMRHS, gauge_,
                        dslashType_, floatPrecision_, nColors_, nInputs, streams_);
}

namespace {

// Shared by instantiateDslash and instantiateDslashXpay: reference SoA
// pointer packs, synchronous on the operator's stream.
void run_soa_dslash(const QcuLattDesc<Nd> &lattDesc, const QcuProcDesc<Nd> &procDesc, int parity, int daggerFlag,
                    const void *const *x, double a, double b, void *mrhsFermionIn, void *mrhsFermionOut, void *gauge,
                    DSLASH_TYPE dtype, QCU_PRECISION floatPrecision, int nColors, int nInputs,
                    ExecutionStreams &streams)
{
  if (dtype != DSLASH_WILSON) errorQcu("Now we only support Wilson dslash.\n");
  const int prec = to_prec(floatPrecision);
  if (prec < 0) errorQcu("Now we only support half, single and double precision.\n");
  if (nColors < 2 || nColors > 8) errorQcu("Now we only support 2-8 colors.\n");
  if (nInputs < 1) errorQcu("At least one input is needed.\n");
  qmrhs::DslashCall c;
  c.sl = qmrhs::SL_SOA_PTRS;
  c.gl = qmrhs::GL_REF_SOA;
  c.prec_s = c.prec_g = c.prec_o = prec;
  for (int d = 0; d < 4; ++d) {
    c.geom.L[d] = lattDesc.dim[d];
    c.geom.grid[d] = procDesc.dim[d];
  }
  c.geom.N = nColors;
  c.k = nInputs;
  c.parity = parity;
  c.dagger = daggerFlag ? 1 : 0;
  c.out_ptrs = static_cast<void *const *>(mrhsFermionOut);
  c.in_ptrs = static_cast<const void *const *>(mrhsFermionIn);
  c.x_ptrs = x;
  c.xa = a;
  c.xb = b;
  c.gauge = gauge;
  c.stream = streams.stream;
  std::string err;
  if (qmrhs::launch_dslash(c, &err)) fatal(err);
  CHECK_HIP(hipStreamSynchronize(streams.stream));
}

}  // namespace

void instantiateDslashXpay(const QcuLattDesc<Nd> &lattDesc, const QcuProcDesc<Nd> &procDesc, int parity,
                           int daggerFlag, double a, double b, void *mrhsFermionIn, void *mrhsFermionX,
                           void *mrhsFermionOut, void *gauge, DSLASH_TYPE dtype, QCU_PRECISION floatPrecision,
                           int nColors, int nInputs, ExecutionStreams &streams)
{
  if (!mrhsFermionX) errorQcu("DslashXpay: null xpay operand.\n");
  run_soa_dslash(lattDesc, procDesc, parity, daggerFlag, static_cast<const void *const *>(mrhsFermionX), a, b,
                 mrhsFermionIn, mrhsFermionOut, gauge, dtype, floatPrecision, nColors, nInputs, streams);
}

void instantiateDslash(const QcuLattDesc<Nd> &lattDesc, const QcuProcDesc<Nd> &procDesc, int parity, int daggerFlag,
                       double kappa, void *mrhsFermionIn, void *mrhsFermionOut, void *gauge, DSLASH_TYPE dtype,
                       QCU_PRECISION floatPrecision, int nColors, int nInputs, ExecutionStreams &streams)
{
  (void)kappa;  // the hopping term carries no kappa (reference qcu_wilson.cu:21-45)
  run_soa_dslash(lattDesc, procDesc, parity, daggerFlag, nullptr, 0.0, 1.0, mrhsFermionIn, mrhsFermionOut, gauge,
                 dtype, floatPrecision, nColors, nInputs, streams);
}


template <typename F> struct PrecOf;
template <> struct PrecOf<double> { static constexpr QCU_PRECISION v = QCU_DOUBLE_PRECISION; };
template <> struct PrecOf<float> { static constexpr QCU_PRECISION v = QCU_SINGLE_PRECISION; };

template <typename _Float, int _nColor, int _nInput>
void WilsonDslash<_Float, _nColor, _nInput>::apply()
{
  const DslashParam &p = this->dslashParam_;
  instantiateDslash(p.lattDesc, p.procDesc, p.parity, p.daggerFlag, p.kappa, p.fermionIn_MRHS, p.fermionOut_MRHS,
                    p.gauge, DSLASH_WILSON, PrecOf<_Float>::v, _nColor, _nInput, p.streams);
}
template <typename _Float, int _nColor, int _nInput>
void WilsonDslash<_Float, _nColor, _nInput>::preApply()
{
}
template <typename _Float, int _nColor, int _nInput>
void WilsonDslash<_Float, _nColor, _nInput>::postApply()
{
}

// The instantiations the reference's dispatch creates (qcu_wilson.cu:61-128).
#define QCU_WD(F, N)                   \
  template class WilsonDslash<F, N, 1>; \
  template class WilsonDslash<F, N, 2>; \
  template class WilsonDslash<F, N, 3>; \
  template class WilsonDslash<F, N, 4>; \
  template class WilsonDslash<F, N, 5>;
QCU_WD(float, 3)
QCU_WD(float, 4)
QCU_WD(float, 5)
QCU_WD(double, 3)
QCU_WD(double, 4)
QCU_WD(double, 5)
#undef QCU_WD

// ---------------------------------------------------------------------------
// Layout shifts (reference src/storage/qcu_storage.cu:9-35): complex<double>,
// N = 3, on the null stream, complete on return.
namespace storage {

void shiftVectorStorage(void *dstVec, void *srcVec, ShiftDirection dir, int subLx, int Ly, int Lz, int Lt, int nColors,
                        int bytesPerComplex)
{
  const int S = subLx * Ly * Lz * Lt, M = Ns * nColors;
  const int rows = dir == TO_COALESCE ? S : M, cols = dir == TO_COALESCE ? M : S;
  std::string err;
  if (qmrhs::transpose_batched(dstVec, srcVec, rows, cols, 1, bytesPerComplex, nullptr, &err)) fatal(err);
  CHECK_HIP(hipDeviceSynchronize());
}

void shiftGaugeStorage(void *dst, void *src, ShiftDirection dir, int Lx, int Ly, int Lz, int Lt, int nColors,
                       int bytesPerComplex)
{
  const int S = Lx / 2 * Ly * Lz * Lt, M = nColors * nColors;
  const int rows = dir == TO_COALESCE ? S : M, cols = dir == TO_COALESCE ? M : S;
  std::string err;
  if (qmrhs::transpose_batched(dst, src, rows, cols, Nd * 2, bytesPerComplex, nullptr, &err)) fatal(err);
  CHECK_HIP(hipDeviceSynchronize());
}

void shiftVectorStorage(void *dstVec, void *srcVec, ShiftDirection dir, int subLx, int Ly, int Lz, int Lt)
{
  shiftVectorStorage(dstVec, srcVec, dir, subLx, Ly, Lz, Lt, 3, 16);
}

void shiftGaugeStorage(void *dst, void *src, ShiftDirection dir, int Lx, int Ly, int Lz, int Lt)
{
  shiftGaugeStorage(dst, src, dir, Lx, Ly, Lz, Lt, 3, 16);
}

}  // namespace storage
}  // namespace qcu

// ---------------------------------------------------------------------------
// C ABI (reference src/qcu.cu:125-181).
static qcu::Qcu *qcu_ptr = nullptr;

namespace qmrhs {
// Colour count of the current context (3 when none was created).
int qcu_context_colors() { return qcu_ptr ? qcu_ptr->geom().N : 3; }
}  // namespace qmrhs

extern "C" {

void initGridSize(QcuGrid_t *grid, QcuParam *p_param, int nColors)
{
  if (!grid || !p_param) errorQcu("initGridSize: null descriptor.\n");
  for (int d = 0; d < 4; ++d) {
    const int L = p_param->lattice_size[d];
    if (L <= 0 || (L & 1)) errorQcu("lattice sizes must be positive and even.\n");
    if (grid->grid_size[d] <= 0) errorQcu("grid sizes must be positive.\n");
  }
  if (nColors < 2 || nColors > 8) errorQcu("Now we only support 2-8 colors.\n");
  delete qcu_ptr;
  qcu_ptr = new qcu::Qcu(p_param, grid, nColors);
}

void pushBackFermions(void *fermionOut, void *fermionIn)
{
  if (qcu_ptr == nullptr) errorQcu("Qcu is not initialized.\n");
  qcu_ptr->pushFermion(fermionOut, fermionIn);
}

void startDslash(int parity, int daggerFlag)
{
  if (qcu_ptr == nullptr) errorQcu("Qcu is not initialized.\n");
  qcu_ptr->startDslash(parity, daggerFlag);
}

void getDslash(int dslashType, double mass, int nColors, int nInputs, int floatPrecision)
{
  if (qcu_ptr == nullptr) errorQcu("Qcu is not initialized.\n");
  qcu_ptr->getDslash(static_cast<DSLASH_TYPE>(dslashType), mass, nColors, nInputs,
                     static_cast<QCU_PRECISION>(floatPrecision));
}

void loadQcuGauge(void *gauge)
{
  if (qcu_ptr == nullptr) errorQcu("Qcu is not initialized.\n");
  qcu_ptr->loadGauge(gauge);
}

void dslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity)
{
  (void)fermion_out, (void)fermion_in, (void)gauge, (void)param, (void)parity;
}

void fullDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int dagger_flag)
{
  (void)fermion_out, (void)fermion_in, (void)gauge, (void)param, (void)dagger_flag;
}

int qcuMrhsDslashReferenceLayout(void *const *out_soa, void *const *in_soa, void *gauge_soa, int prec, int nrhs,
                                 int parity, int dagger, const qcuMrhsGeom *g)
{
  if (!g || !out_soa || !in_soa || !gauge_soa) return qmrhs::api_fail("null argument");
  if (prec != qmrhs::PREC_HALF && prec != qmrhs::PREC_SINGLE && prec != qmrhs::PREC_DOUBLE)
    return qmrhs::api_fail("reference layout supports half, single and double");
  QcuLattDesc<4> latt(g->L[0], g->L[1], g->L[2], g->L[3]);
  QcuProcDesc<4> proc(g->grid[0] > 0 ? g->grid[0] : 1, g->grid[1] > 0 ? g->grid[1] : 1,
                      g->grid[2] > 0 ? g->grid[2] : 1, g->grid[3] > 0 ? g->grid[3] : 1);
  // Created once and kept for the life of the process (never destroyed at
  // exit, when the HIP runtime may already be gone).
  static qcu::ExecutionStreams *streams = new qcu::ExecutionStreams();
  qcu::instantiateDslash(latt, proc, parity, dagger, 0.0, const_cast<void **>(in_soa), const_cast<void **>(out_soa),
                         gauge_soa, DSLASH_WILSON, static_cast<QCU_PRECISION>(prec), g->nColors, nrhs, *streams);
  return 0;
}

}  // extern "C"

// The reference releases its context from a library destructor
// (src/qcu.cu:176-181). Here the context is released by the next initGridSize
// only: at process exit the HIP runtime may already be torn down, and device
// memory is returned with the process anyway.

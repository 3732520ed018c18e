// Dslash parameter block, operator templates and the runtime dispatch entry
// (interface of the reference's include/qcd/qcu_dslash.h:13-64).
//
// instantiateDslash applies the bare hopping term to nInputs right-hand sides
// held in the reference's SoA ("coalesced") layout [4*N][V/2] per RHS, with the
// gauge field in the SoA layout [4][2][N*N][V/2]; mrhsFermionIn/Out point to
// HOST arrays of nInputs device pointers.
#pragma once

#include "qcu.h"
#include "qcuDesc/mrhs_desc.h"
#include "qcuDesc/qcu_desc.h"
#include "qcu_enum.h"
#include "qcu_kernel_param.cuh"

namespace qcu {

static constexpr int Ndim = 4;

struct DslashParam {
  const QcuProcDesc<Ndim> &procDesc;
  const QcuLattDesc<Ndim> &lattDesc;
  int parity;
  int daggerFlag;
  double kappa;

  void *fermionIn_MRHS;
  void *fermionOut_MRHS;

  void *gauge;
  ExecutionStreams &streams;
  DslashParam(const QcuLattDesc<Ndim> &pLattDesc, const QcuProcDesc<Ndim> &pProcDesc, int pParity, int pDaggerFlag,
              double pKappa, void *pFermionIn, void *pFermionOut, void *pGauge, ExecutionStreams &pStreams)
      : procDesc(pProcDesc),
        lattDesc(pLattDesc),
        parity(pParity),
        daggerFlag(pDaggerFlag),
        kappa(pKappa),
        fermionIn_MRHS(pFermionIn),
        fermionOut_MRHS(pFermionOut),
        gauge(pGauge),
        streams(pStreams)
  {
  }
};

template <typename _Float, int _nColor, int _nInput>
class Dslash {
 protected:
  const DslashParam &dslashParam_;

 public:
  Dslash(const DslashParam &dslashParam) : dslashParam_(dslashParam) {}
  virtual ~Dslash() {}
  virtual void apply() = 0;
  virtual void preApply() = 0;
  virtual void postApply() = 0;
};

template <typename _Float, int _nColor, int _nInput>
class WilsonDslash : public Dslash<_Float, _nColor, _nInput> {
 public:
  WilsonDslash(const DslashParam &dslashParam) : Dslash<_Float, _nColor, _nInput>(dslashParam) {}
  virtual ~WilsonDslash() {}
  virtual void apply();
  virtual void preApply();
  virtual void postApply();
};

void instantiateDslash(const QcuLattDesc<Nd> &lattDesc, const QcuProcDesc<Nd> &procDesc, int parity, int daggerFlag,
                       double kappa, void *mrhsFermionIn, void *mrhsFermionOut, void *gauge, DSLASH_TYPE dtype,
                       QCU_PRECISION floatPrecision, int nColors, int nInputs, ExecutionStreams &streams);

// Not in the reference: out_r = a * x_r + b * Dslash(in_r) for the same
// pointer packs and layouts as instantiateDslash (mrhsFermionX: host array of
// nInputs device pointers, one per RHS, in the input's layout), one launch.
void instantiateDslashXpay(const QcuLattDesc<Nd> &lattDesc, const QcuProcDesc<Nd> &procDesc, int parity,
                           int daggerFlag, double a, double b, void *mrhsFermionIn, void *mrhsFermionX,
                           void *mrhsFermionOut, void *gauge, DSLASH_TYPE dtype, QCU_PRECISION floatPrecision,
                           int nColors, int nInputs, ExecutionStreams &streams);

}  // namespace qcu

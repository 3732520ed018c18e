// Dirac operator interface (the reference's include/qcu_dirac.h:7-50: same
// members, constructor and virtuals). kappa = 1/(2(4+mass)) is kept for the
// preconditioned operator; Dslash applies the bare hopping term.
#pragma once
#include "qcd/qcu_dslash.h"
#include "qcuDesc/qcu_desc.h"
#include "qcu_kernel_param.cuh"
namespace qcu {

class Dirac {
 protected:
  void *gauge_;
  double mass_;
  double kappa_;
  int nColors_;
  QCU_PRECISION floatPrecision_;
  const QcuLattDesc<Nd> &lattDesc_;
  const QcuProcDesc<Nd> &procDesc_;
  DSLASH_TYPE dslashType_;

  ExecutionStreams &streams_;

 public:
  Dirac(void *gauge, double mass, int nColors, QCU_PRECISION floatPrecision, const QcuLattDesc<Nd> &lattDesc,
        const QcuProcDesc<Nd> &procDesc, DSLASH_TYPE dslashType, ExecutionStreams &streams)
      : gauge_(gauge),
        mass_(mass),
        kappa_(1.0 / (2.0 * (4.0 + mass))),
        nColors_(nColors),
        floatPrecision_(floatPrecision),
        lattDesc_(lattDesc),
        procDesc_(procDesc),
        dslashType_(dslashType),
        streams_(streams)
  {
  }

  virtual ~Dirac() {}
  virtual void Dslash(void *outputMRHS, void *inputMRHS, int parity, int daggerFlag, int nInputs) = 0;
  virtual void DslashXpay() = 0;
  double kappa() const { return kappa_; }
};

class DiracWilson : public Dirac {
 public:
  DiracWilson(void *gauge, double mass, int nColors, QCU_PRECISION floatPrecision, const QcuLattDesc<Nd> &lattDesc,
              const QcuProcDesc<Nd> &procDesc, DSLASH_TYPE dslashType, ExecutionStreams &streams)
      : Dirac(gauge, mass, nColors, floatPrecision, lattDesc, procDesc, dslashType, streams)
  {
  }
  ~DiracWilson() {}
  // outputMRHS / inputMRHS: host arrays of nInputs device pointers, each a
  // single-parity spinor in the reference SoA layout [4*N][V/2].
  virtual void Dslash(void *outputMRHS, void *inputMRHS, int parity, int daggerFlag, int nInputs);
  // Empty in the reference (qcu_dirac.h:49); the fused x + a*D y lives in the
  // extended C ABI (qcuMrhsDslashXpay) where its operands can be named.
  virtual void DslashXpay() {}
};

}  // namespace qcu

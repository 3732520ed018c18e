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
  // Not in the reference (whose DslashXpay takes no operands): out_r = a * x_r
  // + b * Dslash(in_r), same pointer packs as Dslash, xMRHS one pointer per RHS.
  virtual void DslashXpay(void *outputMRHS, void *inputMRHS, void *xMRHS, double a, double b, int parity,
                          int daggerFlag, int nInputs)
  {
    (void)outputMRHS, (void)inputMRHS, (void)xMRHS, (void)a, (void)b, (void)parity, (void)daggerFlag, (void)nInputs;
    errorQcu("DslashXpay with operands is not implemented for this operator.\n");
  }
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
  // Empty in the reference (qcu_dirac.h:49), and kept empty: it names no operands.
  virtual void DslashXpay() {}
  // out_r = a * x_r + b * Dslash(in_r) in one launch (reference SoA layout,
  // like Dslash).
  virtual void DslashXpay(void *outputMRHS, void *inputMRHS, void *xMRHS, double a, double b, int parity,
                          int daggerFlag, int nInputs);
};

}  // namespace qcu

// A C++ caller of the reference's C++ interface, built against libqcu.so:
// include/qcu_dirac.h (DiracWilson), include/qcd/qcu_dslash.h (WilsonDslash,
// DslashParam), include/storage/qcu_storage.h (layout shifts) and include/qcu.h
// (the C entry points), in the way the reference's own driver uses them
// (reference src/test.cu:9-54, src/qcu_dirac.cu:7-14, src/qcd/qcu_wilson.cu:61-128).
//
// Usage: dirac_caller <dir>. <dir>/case.txt holds
//   Lx Ly Lz Lt nColors nInputs parity dagger precision a b mass
// (precision 1 = single, 2 = double; a, b the DslashXpay coefficients), and
// the binary inputs, all complex<double>:
//   gauge.bin, gauge_other.bin  user gauge [4][2][V/2][N][N]
//   in.bin, x.bin               nInputs user spinors [V/2][4][N] each
// Outputs (complex of the case's precision, nInputs user spinors each):
//   out_dirac.bin   DiracWilson::Dslash on reference SoA buffers
//   out_xpay.bin    DiracWilson::DslashXpay: a x + b Dslash(in)
//   out_wilson.bin  WilsonDslash<F, N, nInputs>::apply (nColors 3..5, nInputs 1..5)
//   out_qcuh.bin    initGridSize, loadQcuGauge(gauge_other), getDslash,
//                   loadQcuGauge(gauge), pushBackFermions x nInputs, startDslash
// Exit status 0 on success; 2 on a usage or input error.
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qcd/qcu_dslash.h"
#include "qcu.h"
#include "qcuDesc/qcu_desc.h"
#include "qcu_dirac.h"
#include "qcu_kernel_param.cuh"
#include "storage/qcu_storage.h"

namespace {

struct Case {
  int L[4], N, k, parity, dagger, prec;
  double a, b, mass;
};

bool read_file(const std::string &path, std::vector<char> &buf, size_t bytes)
{
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  buf.resize(bytes);
  const size_t got = fread(buf.data(), 1, bytes, f);
  fclose(f);
  return got == bytes;
}

bool write_file(const std::string &path, const void *p, size_t bytes)
{
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  const size_t put = fwrite(p, 1, bytes, f);
  fclose(f);
  return put == bytes;
}

// complex<double> host data -> the case's precision
std::vector<char> to_prec(const std::vector<char> &d64, int prec)
{
  if (prec == QCU_DOUBLE_PRECISION) return d64;
  const size_t n = d64.size() / sizeof(double);
  std::vector<char> out(n * sizeof(float));
  const double *s = reinterpret_cast<const double *>(d64.data());
  float *t = reinterpret_cast<float *>(out.data());
  for (size_t i = 0; i < n; ++i) t[i] = float(s[i]);
  return out;
}

void *to_device(const std::vector<char> &h)
{
  void *d = nullptr;
  CHECK_HIP(hipMalloc(&d, h.size()));
  CHECK_HIP(hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice));
  return d;
}

template <typename F, int N, int K>
void run_wilson(qcu::DslashParam &p)
{
  qcu::WilsonDslash<F, N, K> op(p);
  op.preApply();
  op.apply();
  op.postApply();
}

template <typename F, int N>
bool wilson_k(qcu::DslashParam &p, int k)
{
  switch (k) {
    case 1: run_wilson<F, N, 1>(p); return true;
    case 2: run_wilson<F, N, 2>(p); return true;
    case 3: run_wilson<F, N, 3>(p); return true;
    case 4: run_wilson<F, N, 4>(p); return true;
    case 5: run_wilson<F, N, 5>(p); return true;
    default: return false;
  }
}

template <typename F>
bool wilson_nk(qcu::DslashParam &p, int N, int k)
{
  switch (N) {
    case 3: return wilson_k<F, 3>(p, k);
    case 4: return wilson_k<F, 4>(p, k);
    case 5: return wilson_k<F, 5>(p, k);
    default: return false;
  }
}

}  // namespace

int main(int argc, char **argv)
{
  if (argc != 2) {
    fprintf(stderr, "usage: %s <case directory>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  Case c;
  {
    FILE *f = fopen((dir + "/case.txt").c_str(), "r");
    if (!f) {
      fprintf(stderr, "cannot open %s/case.txt\n", dir.c_str());
      return 2;
    }
    const int n = fscanf(f, "%d %d %d %d %d %d %d %d %d %lf %lf %lf", &c.L[0], &c.L[1], &c.L[2], &c.L[3], &c.N, &c.k,
                         &c.parity, &c.dagger, &c.prec, &c.a, &c.b, &c.mass);
    fclose(f);
    if (n != 12 || (c.prec != 1 && c.prec != 2) || c.k < 1 || c.N < 2) {
      fprintf(stderr, "bad case.txt\n");
      return 2;
    }
  }
  const int subLx = c.L[0] / 2, Ly = c.L[1], Lz = c.L[2], Lt = c.L[3];
  const size_t S = size_t(subLx) * Ly * Lz * Lt;
  const size_t spin64 = S * Ns * c.N * 16, gauge64 = 8 * S * c.N * c.N * 16;
  const int bpc = c.prec == 2 ? 16 : 8;  // bytes per stored complex
  const size_t spinB = S * Ns * c.N * bpc, gaugeB = 8 * S * c.N * c.N * bpc;
  const QCU_PRECISION prec = static_cast<QCU_PRECISION>(c.prec);

  std::vector<char> g64, gother64, in64, x64;
  if (!read_file(dir + "/gauge.bin", g64, gauge64) || !read_file(dir + "/gauge_other.bin", gother64, gauge64) ||
      !read_file(dir + "/in.bin", in64, spin64 * c.k) || !read_file(dir + "/x.bin", x64, spin64 * c.k)) {
    fprintf(stderr, "missing or short input file\n");
    return 2;
  }

  // ---- reference C++ layer: shifts + DiracWilson on SoA buffers ----------
  void *gaugeUser = to_device(to_prec(g64, c.prec));
  void *gaugeSoA = nullptr;
  CHECK_HIP(hipMalloc(&gaugeSoA, gaugeB));
  if (c.N == 3 && c.prec == 2)  // the reference signature: complex<double>, N = 3
    qcu::storage::shiftGaugeStorage(gaugeSoA, gaugeUser, TO_COALESCE, c.L[0], Ly, Lz, Lt);
  else
    qcu::storage::shiftGaugeStorage(gaugeSoA, gaugeUser, TO_COALESCE, c.L[0], Ly, Lz, Lt, c.N, bpc);

  const std::vector<char> inP = to_prec(in64, c.prec), xP = to_prec(x64, c.prec);
  std::vector<void *> inUser(c.k), xUser(c.k), inSoA(c.k), xSoA(c.k), outSoA(c.k), outUser(c.k);
  for (int r = 0; r < c.k; ++r) {
    std::vector<char> one(inP.begin() + r * spinB, inP.begin() + (r + 1) * spinB);
    std::vector<char> onex(xP.begin() + r * spinB, xP.begin() + (r + 1) * spinB);
    inUser[r] = to_device(one);
    xUser[r] = to_device(onex);
    CHECK_HIP(hipMalloc(&inSoA[r], spinB));
    CHECK_HIP(hipMalloc(&xSoA[r], spinB));
    CHECK_HIP(hipMalloc(&outSoA[r], spinB));
    CHECK_HIP(hipMalloc(&outUser[r], spinB));
    if (c.N == 3 && c.prec == 2) {
      qcu::storage::shiftVectorStorage(inSoA[r], inUser[r], TO_COALESCE, subLx, Ly, Lz, Lt);
      qcu::storage::shiftVectorStorage(xSoA[r], xUser[r], TO_COALESCE, subLx, Ly, Lz, Lt);
    } else {
      qcu::storage::shiftVectorStorage(inSoA[r], inUser[r], TO_COALESCE, subLx, Ly, Lz, Lt, c.N, bpc);
      qcu::storage::shiftVectorStorage(xSoA[r], xUser[r], TO_COALESCE, subLx, Ly, Lz, Lt, c.N, bpc);
    }
  }

  QcuLattDesc<Nd> latt(c.L[0], Ly, Lz, Lt);
  QcuProcDesc<Nd> proc(1, 1, 1, 1);
  qcu::ExecutionStreams streams;
  std::vector<char> host(spinB * c.k);
  auto gather = [&](const char *name) -> bool {
    for (int r = 0; r < c.k; ++r) {
      qcu::storage::shiftVectorStorage(outUser[r], outSoA[r], TO_NON_COALESCE, subLx, Ly, Lz, Lt, c.N, bpc);
      CHECK_HIP(hipMemcpy(host.data() + r * spinB, outUser[r], spinB, hipMemcpyDeviceToHost));
      CHECK_HIP(hipMemset(outSoA[r], 0, spinB));
    }
    return write_file(dir + "/" + name, host.data(), host.size());
  };
  {
    qcu::DiracWilson dirac(gaugeSoA, c.mass, c.N, prec, latt, proc, DSLASH_WILSON, streams);
    qcu::Dirac &op = dirac;  // through the base interface, as a solver would hold it
    op.Dslash(outSoA.data(), inSoA.data(), c.parity, c.dagger, c.k);
    if (!gather("out_dirac.bin")) return 2;
    op.DslashXpay(outSoA.data(), inSoA.data(), xSoA.data(), c.a, c.b, c.parity, c.dagger, c.k);
    if (!gather("out_xpay.bin")) return 2;
    op.DslashXpay();  // the reference's operand-free hook: a no-op
  }
  if (c.N >= 3 && c.N <= 5 && c.k <= 5) {
    qcu::DslashParam p(latt, proc, c.parity, c.dagger, 1.0 / (2.0 * (4.0 + c.mass)), inSoA.data(), outSoA.data(),
                       gaugeSoA, streams);
    const bool ok = c.prec == 2 ? wilson_nk<double>(p, c.N, c.k) : wilson_nk<float>(p, c.N, c.k);
    if (!ok || !gather("out_wilson.bin")) return 2;
  }

  // ---- qcu.h: the reference's call protocol, gauge loaded twice ----------
  {
    QcuGrid_t grid;
    QcuParam param;
    for (int d = 0; d < 4; ++d) {
      grid.grid_size[d] = 1;
      param.lattice_size[d] = c.L[d];
    }
    void *gOther = to_device(gother64), *gReal = to_device(g64);  // qcu.h takes complex<double> links
    initGridSize(&grid, &param, c.N);
    loadQcuGauge(gOther);
    getDslash(DSLASH_WILSON, c.mass, c.N, c.k, c.prec);
    loadQcuGauge(gReal);  // replaces the field under the operator made above
    for (int r = 0; r < c.k; ++r) pushBackFermions(outUser[r], inUser[r]);
    startDslash(c.parity, c.dagger);
    for (int r = 0; r < c.k; ++r) CHECK_HIP(hipMemcpy(host.data() + r * spinB, outUser[r], spinB, hipMemcpyDeviceToHost));
    if (!write_file(dir + "/out_qcuh.bin", host.data(), host.size())) return 2;
    CHECK_HIP(hipFree(gOther));
    CHECK_HIP(hipFree(gReal));
  }

  for (int r = 0; r < c.k; ++r) {
    CHECK_HIP(hipFree(inUser[r]));
    CHECK_HIP(hipFree(xUser[r]));
    CHECK_HIP(hipFree(inSoA[r]));
    CHECK_HIP(hipFree(xSoA[r]));
    CHECK_HIP(hipFree(outSoA[r]));
    CHECK_HIP(hipFree(outUser[r]));
  }
  CHECK_HIP(hipFree(gaugeSoA));
  CHECK_HIP(hipFree(gaugeUser));
  printf("dirac_caller: ok N=%d k=%d prec=%d\n", c.N, c.k, c.prec);
  return 0;
}

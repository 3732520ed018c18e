// Launcher for the REFERENCE's own kernels, compiled by hipcc directly from
// the reference sources (include/kernels/wilson_dslash_4d.cuh,
// include/kernels/storage/storage_kernels.cuh) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: GPU tests use it to check the CPU oracle (and the
// library) against the reference's dslash on identical inputs. It mirrors the
// reference launch (src/qcd/qcu_wilson.cu:9-18, src/storage/qcu_storage.cu:9-35)
// except that the block size is chosen to divide the site count, because the
// reference kernels carry no bounds check.
#include <hip/hip_runtime.h>

#include "kernels/storage/storage_kernels.cuh"
#include "kernels/wilson_dslash_4d.cuh"

namespace {

int block_for(int S)
{
  for (int b = 256; b >= 64; b >>= 1)
    if (S % b == 0) return b;
  return 0;
}

template <typename F, int N, int K>
int launch_k(void **out, void **in, void *gauge, F flag, const int *L, const int *grid, int parity)
{
  qcu::MrhsInput<K> o, i;
  for (int r = 0; r < K; ++r) {
    o.fermion[r] = out[r];
    i.fermion[r] = in[r];
  }
  const int S = L[0] * L[1] * L[2] * L[3] / 2;
  const int B = block_for(S);
  if (!B) return 2;
  hipLaunchKernelGGL((qcu::kernel::dslashKernel<F, N, K>), dim3(S / B), dim3(B), 0, 0, o, i, gauge, flag, L[0], L[1],
                     L[2], L[3], grid[0], grid[1], grid[2], grid[3], parity);
  if (hipGetLastError() != hipSuccess) return 3;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

template <typename F, int N>
int launch_n(void **out, void **in, void *gauge, int nrhs, F flag, const int *L, const int *grid, int parity)
{
  switch (nrhs) {
    case 1: return launch_k<F, N, 1>(out, in, gauge, flag, L, grid, parity);
    case 2: return launch_k<F, N, 2>(out, in, gauge, flag, L, grid, parity);
    case 3: return launch_k<F, N, 3>(out, in, gauge, flag, L, grid, parity);
    case 4: return launch_k<F, N, 4>(out, in, gauge, flag, L, grid, parity);
    case 5: return launch_k<F, N, 5>(out, in, gauge, flag, L, grid, parity);
    default: return 1;
  }
}

template <typename F>
int launch_f(void **out, void **in, void *gauge, int nrhs, int N, int dagger, const int *L, const int *grid,
             int parity)
{
  const F flag = dagger ? F(-1) : F(1);
  switch (N) {
    case 3: return launch_n<F, 3>(out, in, gauge, nrhs, flag, L, grid, parity);
    case 4: return launch_n<F, 4>(out, in, gauge, nrhs, flag, L, grid, parity);
    case 5: return launch_n<F, 5>(out, in, gauge, nrhs, flag, L, grid, parity);
    default: return 1;
  }
}

}  // namespace

extern "C" {

// Reference dslash on reference SoA spinors [4N][S] (host arrays of nrhs
// device pointers) and SoA gauge [4][2][N*N][S]. prec: 1 = float, 2 = double.
// nrhs 1..5 and N 3..5 (the reference's instantiation range).
int ref_dslash(void **out, void **in, void *gauge, int nrhs, int N, int prec, const int *L, const int *grid,
               int parity, int dagger)
{
  if (prec == 2) return launch_f<double>(out, in, gauge, nrhs, N, dagger, L, grid, parity);
  if (prec == 1) return launch_f<float>(out, in, gauge, nrhs, N, dagger, L, grid, parity);
  return 1;
}

// Reference layout shifts, complex<double>, N = 3 (the only instantiation the
// reference makes). dir 0 = to SoA ("coalesce"), 1 = back.
int ref_shift_vector(void *dst, void *src, int dir, int subLx, int Ly, int Lz, int Lt)
{
  const int S = subLx * Ly * Lz * Lt, B = block_for(S);
  if (!B) return 2;
  if (dir == 0)
    hipLaunchKernelGGL((qcu::kernel::storeVectorCoalesced<double, 3>), dim3(S / B), dim3(B), 0, 0, dst, src, subLx,
                       Ly, Lz, Lt);
  else
    hipLaunchKernelGGL((qcu::kernel::storeVectorNonCoalesced<double, 3>), dim3(S / B), dim3(B), 0, 0, dst, src,
                       subLx, Ly, Lz, Lt);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

int ref_shift_gauge(void *dst, void *src, int dir, int Lx, int Ly, int Lz, int Lt)
{
  const int S = Lx / 2 * Ly * Lz * Lt, B = block_for(S);
  if (!B) return 2;
  if (dir == 0)
    hipLaunchKernelGGL((qcu::kernel::storeGaugeCoalesced<double, 3>), dim3(S / B), dim3(B), 0, 0, dst, src, Lx, Ly,
                       Lz, Lt);
  else
    hipLaunchKernelGGL((qcu::kernel::storeGaugeNonCoalesced<double, 3>), dim3(S / B), dim3(B), 0, 0, dst, src, Lx, Ly,
                       Lz, Lt);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

}  // extern "C"

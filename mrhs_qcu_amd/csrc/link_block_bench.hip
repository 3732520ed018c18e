// Link x half-spinor block products on the vector ALUs and on the matrix
// cores: the arithmetic of one hop of the SU(N) operator,
//   Y (N x C complex) = U (N x N complex) H (N x C complex),  C = 2k
// (two half-spinor spins times k right-hand sides), for a batch of blocks
// whose operands are cache resident (nb blocks reused cyclically: the
// products and their operand loads, not HBM traffic) or, with `resident`,
// held in registers (one block per lane group repeated: the products alone).
// Used to decide whether an MFMA path pays for the N >= 4 operators.
//
// MFMA form: v_mfma_f32_16x16x4_f32 on the real embedding
//   [Yr; Yi] = [[Ur, -Ui], [Ui, Ur]] [Hr; Hi]
// (2N x 2N times 2N x C, zero-padded to 16 rows, K a multiple of 4 and
// columns a multiple of 16), one block per wave and iteration. Operand
// layout of the instruction (64 lanes): A element (row l % 16, k l / 16),
// B element (k l / 16, column l % 16), D rows 4 (l / 16) + v, v = 0..3, of
// column l % 16. The self-check below compares both forms with an FP64
// host product, so a wrong layout shows as a large error, not as a fault.
//
// VALU form: the layout of the production kernels, each lane one link times
// four (or two) columns of one block, several blocks per wave side by side.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "api_internal.h"

namespace qmrhs {
namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kWaves = 4;  // waves per workgroup
constexpr int kTables = 1024;  // blocks in the operand tables

struct BlockArgs {
  const float2 *U;  // [nb][N][N]
  const float2 *H;  // [nb][N][C]
  float2 *Y;        // check: [nb][N][C]; timing: [waves][N][C] (kept live)
  int N, C, nb, reps, check;
};

__device__ __forceinline__ uint32_t mix(uint32_t x)
{
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__global__ void fill_kernel(float2 *p, size_t n, uint32_t seed)
{
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t a = mix(uint32_t(i) * 2u + seed), b = mix(uint32_t(i) * 2u + 1u + seed);
  p[i] = make_float2(float(a >> 8) * (1.0f / 16777216.0f) - 0.5f, float(b >> 8) * (1.0f / 16777216.0f) - 0.5f);
}

// VALU: as in the production kernels, a lane multiplies one link into CPL
// columns (two spins times R right-hand sides there): lane (q, g) takes
// columns g CPL .. g CPL + CPL - 1 of the q-th of the wave's Q = 64 / (C / CPL)
// blocks, loading the link once for all of them.
template <int N, int CPL, bool RES>
__global__ void __launch_bounds__(64 * kWaves) valu_kernel(const BlockArgs a)
{
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int C = a.C, G = C / CPL, Q = 64 / G;
  const int q = lane / G, c0 = (lane % G) * CPL;
  if (q >= Q) return;
  float2 y[N][CPL];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int m = 0; m < CPL; ++m) y[i][m] = make_float2(0.f, 0.f);
  const int reps = a.check ? 1 : a.reps;
  float2 ur[RES ? N * N : 1], hr[RES ? N : 1][CPL];
  if constexpr (RES) {  // one block, loaded once: the loop below is arithmetic only
    const int b = (wave * Q + q) % a.nb;
#pragma unroll
    for (int e = 0; e < N * N; ++e) ur[e] = a.U[size_t(b) * N * N + e];
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int m = 0; m < CPL; ++m) hr[j][m] = a.H[size_t(b) * N * C + c0 + j * C + m];
  }
  for (int it = 0; it < reps; ++it) {
    const int blk = a.check ? wave * Q + q : (wave * reps + it) * Q + q;
    if (a.check && blk >= a.nb) return;
    const int b = blk % a.nb;
    const float2 *U = a.U + size_t(b) * N * N;
    const float2 *H = a.H + size_t(b) * N * C + c0;
    float2 h[N][CPL];
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int m = 0; m < CPL; ++m) h[j][m] = RES ? hr[RES ? j : 0][m] : H[j * C + m];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const float2 u = RES ? ur[RES ? i * N + j : 0] : U[i * N + j];
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
          y[i][m].x = fmaf(u.x, h[j][m].x, fmaf(-u.y, h[j][m].y, y[i][m].x));
          y[i][m].y = fmaf(u.x, h[j][m].y, fmaf(u.y, h[j][m].x, y[i][m].y));
        }
      }
    if (a.check) {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int m = 0; m < CPL; ++m) a.Y[(size_t(b) * N + i) * C + c0 + m] = y[i][m];
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int m = 0; m < CPL; ++m) a.Y[(size_t(wave) * N + i) * C + c0 + m] = y[i][m];
}

// MFMA: one block per wave and iteration; T = ceil(C / 16) column tiles,
// KS = ceil(2N / 4) k-steps.
template <int N, int T, bool RES>
__global__ void __launch_bounds__(64 * kWaves) mfma_kernel(const BlockArgs a)
{
  constexpr int KS = (2 * N + 3) / 4;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int C = a.C;
  const int row = lane & 15, kk = lane >> 4, col = lane & 15;
  f4v acc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
  const int reps = a.check ? 1 : a.reps;
  // Operands of block b for this lane: A = [[Ur, -Ui], [Ui, Ur]] element
  // (row, 4 s + kk), B = [Hr; Hi] element (4 s + kk, 16 t + col).
  auto load_ops = [&](int b, float (&av)[KS], float (&bv)[KS][T]) {
    const float2 *U = a.U + size_t(b) * N * N;
    const float2 *H = a.H + size_t(b) * N * C;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int j = 4 * s + kk;
      av[s] = 0.f;
      if (row < 2 * N && j < 2 * N) {
        const float2 u = U[(row % N) * N + (j % N)];
        const bool top = row < N, left = j < N;
        av[s] = top == left ? u.x : (top ? -u.y : u.y);
      }
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int cc = 16 * t + col;
        bv[s][t] = 0.f;
        if (j < 2 * N && cc < C) {
          const float2 h = H[(j % N) * C + cc];
          bv[s][t] = j < N ? h.x : h.y;
        }
      }
    }
  };
  float av[KS], bv[KS][T];
  if constexpr (RES) load_ops(wave % a.nb, av, bv);  // one block, loaded once
  for (int it = 0; it < reps; ++it) {
    const int blk = a.check ? wave : wave * reps + it;
    if (a.check && blk >= a.nb) return;
    if constexpr (!RES) load_ops(blk % a.nb, av, bv);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], bv[s][t], acc[t], 0, 0, 0);
    if (a.check) break;
  }
  const int bo = wave;  // check: the block index; timing: the wave's own slot
  if (a.check && bo >= a.nb) return;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int cc = 16 * t + col;
    if (cc >= C) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 4 * kk + v;  // D row
      if (r >= 2 * N) continue;
      float2 *y = a.Y + (size_t(bo) * N + (r % N)) * C + cc;
      // rows 0..N-1 hold Re Y, N..2N-1 Im Y; the two halves sit in
      // different lanes, so each writes its own component.
      if (r < N) reinterpret_cast<float *>(y)[0] = acc[t][v];
      else reinterpret_cast<float *>(y)[1] = acc[t][v];
    }
  }
}

template <int N, bool RES>
hipError_t launch_mfma(const BlockArgs &a, int grid, hipStream_t st)
{
  const int T = (a.C + 15) / 16;
  switch (T) {
    case 1: hipLaunchKernelGGL((mfma_kernel<N, 1, RES>), dim3(grid), dim3(64 * kWaves), 0, st, a); break;
    case 2: hipLaunchKernelGGL((mfma_kernel<N, 2, RES>), dim3(grid), dim3(64 * kWaves), 0, st, a); break;
    case 3: hipLaunchKernelGGL((mfma_kernel<N, 3, RES>), dim3(grid), dim3(64 * kWaves), 0, st, a); break;
    case 4: hipLaunchKernelGGL((mfma_kernel<N, 4, RES>), dim3(grid), dim3(64 * kWaves), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Columns per VALU lane: 4 (two spins x two right-hand sides, the
// production kernels' common tile) when C allows, else 2.
inline int valu_cpl(int C) { return C % 4 == 0 ? 4 : 2; }

template <int N, bool RES>
hipError_t launch_valu(const BlockArgs &a, int grid, hipStream_t st)
{
  if (valu_cpl(a.C) == 4) hipLaunchKernelGGL((valu_kernel<N, 4, RES>), dim3(grid), dim3(64 * kWaves), 0, st, a);
  else hipLaunchKernelGGL((valu_kernel<N, 2, RES>), dim3(grid), dim3(64 * kWaves), 0, st, a);
  return hipGetLastError();
}

template <int N, bool RES>
hipError_t launch_n(int mode, const BlockArgs &a, int grid, hipStream_t st)
{
  return mode == 0 ? launch_valu<N, RES>(a, grid, st) : launch_mfma<N, RES>(a, grid, st);
}

template <bool RES>
hipError_t launch_block(int mode, const BlockArgs &a, int grid, hipStream_t st)
{
  switch (a.N) {
    case 2: return launch_n<2, RES>(mode, a, grid, st);
    case 3: return launch_n<3, RES>(mode, a, grid, st);
    case 4: return launch_n<4, RES>(mode, a, grid, st);
    case 5: return launch_n<5, RES>(mode, a, grid, st);
    case 6: return launch_n<6, RES>(mode, a, grid, st);
    case 7: return launch_n<7, RES>(mode, a, grid, st);
    case 8: return launch_n<8, RES>(mode, a, grid, st);
    default: return hipErrorInvalidValue;
  }
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf()
  {
    if (p) (void)hipFree(p);
  }
};
struct Ev {
  hipEvent_t e = nullptr;
  ~Ev()
  {
    if (e) (void)hipEventDestroy(e);
  }
};

}  // namespace
}  // namespace qmrhs

using namespace qmrhs;

extern "C" int qcuMrhsLinkBlockBench(int N, int nrhs, int reps, int resident, double *tflops, double *max_rel_err,
                                     void *stream)
{
  if (N < 2 || N > 8 || nrhs < 1 || nrhs > 32 || reps < 1 || !tflops || !max_rel_err)
    return api_fail("qcuMrhsLinkBlockBench: needs 2 <= N <= 8, 1 <= nrhs <= 32, reps >= 1 and output pointers");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int C = 2 * nrhs, nb = kTables;
  const size_t nu = size_t(nb) * N * N, nh = size_t(nb) * N * C;
  // Timing grid: 8 waves per SIMD (8192 waves) for both forms. An MFMA wave
  // takes one block per iteration and runs `reps` iterations; a VALU wave
  // takes Q = 64 / (C / CPL) blocks side by side, so it runs reps / Q.
  const int Q = 64 / (C / valu_cpl(C));  // blocks per VALU wave
  const int waves = 8192;
  const size_t ny = std::max(nh, size_t(waves) * N * C);
  DevBuf U, H, Yv, Ym;
#define QM_HIP(x)                                                                           \
  do {                                                                                      \
    const hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return api_fail(std::string("qcuMrhsLinkBlockBench: ") + hipGetErrorString(e_)); \
  } while (0)
  QM_HIP(hipMalloc(&U.p, nu * sizeof(float2)));
  QM_HIP(hipMalloc(&H.p, nh * sizeof(float2)));
  QM_HIP(hipMalloc(&Yv.p, ny * sizeof(float2)));
  QM_HIP(hipMalloc(&Ym.p, ny * sizeof(float2)));
  hipLaunchKernelGGL(fill_kernel, dim3((nu + 255) / 256), dim3(256), 0, st, static_cast<float2 *>(U.p), nu, 17u);
  hipLaunchKernelGGL(fill_kernel, dim3((nh + 255) / 256), dim3(256), 0, st, static_cast<float2 *>(H.p), nh, 91u);
  QM_HIP(hipGetLastError());

  // Self-check: every table block once, both forms, against an FP64 host product.
  BlockArgs a{static_cast<const float2 *>(U.p), static_cast<const float2 *>(H.p), nullptr, N, C, nb, 1, 1};
  a.Y = static_cast<float2 *>(Yv.p);
  QM_HIP(launch_block<false>(0, a, (nb + Q * kWaves - 1) / (Q * kWaves), st));
  a.Y = static_cast<float2 *>(Ym.p);
  QM_HIP(launch_block<false>(1, a, (nb + kWaves - 1) / kWaves, st));
  std::vector<float2> hu(nu), hh(nh), hv(nh), hm(nh);
  QM_HIP(hipMemcpyAsync(hu.data(), U.p, nu * sizeof(float2), hipMemcpyDeviceToHost, st));
  QM_HIP(hipMemcpyAsync(hh.data(), H.p, nh * sizeof(float2), hipMemcpyDeviceToHost, st));
  QM_HIP(hipMemcpyAsync(hv.data(), Yv.p, nh * sizeof(float2), hipMemcpyDeviceToHost, st));
  QM_HIP(hipMemcpyAsync(hm.data(), Ym.p, nh * sizeof(float2), hipMemcpyDeviceToHost, st));
  QM_HIP(hipStreamSynchronize(st));
  double ev = 0, em = 0, ref = 0;
  for (int b = 0; b < nb; ++b)
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < C; ++c) {
        double re = 0, im = 0;
        for (int j = 0; j < N; ++j) {
          const float2 u = hu[(size_t(b) * N + i) * N + j], h = hh[(size_t(b) * N + j) * C + c];
          re += double(u.x) * h.x - double(u.y) * h.y;
          im += double(u.x) * h.y + double(u.y) * h.x;
        }
        const size_t o = (size_t(b) * N + i) * C + c;
        ref = std::max(ref, std::hypot(re, im));
        ev = std::max(ev, std::hypot(hv[o].x - re, hv[o].y - im));
        em = std::max(em, std::hypot(hm[o].x - re, hm[o].y - im));
      }
  max_rel_err[0] = ref > 0 ? ev / ref : ev;
  max_rel_err[1] = ref > 0 ? em / ref : em;

  // Timing: one warm-up launch, then three, per form.
  Ev e0, e1;
  QM_HIP(hipEventCreate(&e0.e));
  QM_HIP(hipEventCreate(&e1.e));
  a.check = 0;
  const int grid = waves / kWaves;
  for (int mode = 0; mode < 2; ++mode) {
    a.Y = static_cast<float2 *>(mode ? Ym.p : Yv.p);
    a.reps = mode ? reps : std::max(1, reps / Q);
    auto go = [&] { return resident ? launch_block<true>(mode, a, grid, st) : launch_block<false>(mode, a, grid, st); };
    QM_HIP(go());
    QM_HIP(hipEventRecord(e0.e, st));
    for (int r = 0; r < 3; ++r) QM_HIP(go());
    QM_HIP(hipEventRecord(e1.e, st));
    QM_HIP(hipEventSynchronize(e1.e));
    float ms = 0;
    QM_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
    // useful real flops: 8 per complex multiply-add, N^2 C of them per block
    const double blk = double(waves) * a.reps * (mode ? 1 : Q);
    tflops[mode] = blk * 8.0 * N * N * C / (ms / 3 * 1e-3) / 1e12;
  }
#undef QM_HIP
  return 0;
}

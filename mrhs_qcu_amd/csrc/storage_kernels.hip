// Storage layer: layout changes between the user layouts, the reference SoA
// layout and the native RHS-innermost layouts, with precision conversion.
//
//  * shift kernels (reference src/storage/qcu_storage.cu:9-35 and
//    include/kernels/storage/storage_kernels.cuh:14-101): pure permutations,
//    spinor [S][4N] <-> [4N][S] and gauge [4][2][S][N*N] <-> [4][2][N*N][S],
//    done here as one LDS-tiled batched transpose of E-byte elements so both
//    the reads and the writes are coalesced.
//  * pack / unpack: k user spinors [S][4][N] (one pointer each) <-> the native
//    packed field [S][4][N][k].
//  * gauge pack: user gauge [4][2][S][N][N] -> native [2][S][8][N][N] holding,
//    per site x of parity p, U_mu(x) (slots 0..3) and U_mu(x-mu)^+ (slots 4..7).
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "qcu_device.h"
#include "storage_api.h"

namespace qmrhs {

// ---------------------------------------------------------------------------
// Batched transpose: for b < nb: dst[b][c][r] = src[b][r][c], rows x cols
// elements of type T (T = the element's bytes: uint32_t / uint2 / uint4).
template <class T>
__global__ void __launch_bounds__(256) transpose_kernel(T *__restrict__ dst, const T *__restrict__ src, int rows,
                                                        int cols, int nb)
{
  __shared__ T tile[32][33];
  const int b = blockIdx.z;
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const T *s = src + size_t(b) * rows * cols;
  T *d = dst + size_t(b) * rows * cols;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    if (r < rows && c < cols) tile[j][tx] = s[size_t(r) * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + tx;
    if (r < rows && c < cols) d[size_t(c) * rows + r] = tile[tx][j];
  }
  (void)nb;
}

int transpose_batched(void *dst, const void *src, int rows, int cols, int nb, int elem_bytes, hipStream_t st,
                      std::string *err)
{
  if (rows <= 0 || cols <= 0 || nb <= 0) {
    if (err) *err = "empty transpose";
    return 1;
  }
  if (nb > 65535) {
    if (err) *err = "too many transpose batches";
    return 1;
  }
  dim3 grid((rows + 31) / 32, (cols + 31) / 32, nb), block(256);
  if (grid.y > 65535) {
    if (err) *err = "transpose too wide";
    return 1;
  }
  switch (elem_bytes) {
    case 4:
      hipLaunchKernelGGL(transpose_kernel<uint32_t>, grid, block, 0, st, static_cast<uint32_t *>(dst),
                         static_cast<const uint32_t *>(src), rows, cols, nb);
      break;
    case 8:
      hipLaunchKernelGGL(transpose_kernel<uint2>, grid, block, 0, st, static_cast<uint2 *>(dst),
                         static_cast<const uint2 *>(src), rows, cols, nb);
      break;
    case 16:
      hipLaunchKernelGGL(transpose_kernel<uint4>, grid, block, 0, st, static_cast<uint4 *>(dst),
                         static_cast<const uint4 *>(src), rows, cols, nb);
      break;
    default:
      if (err) *err = "element size must be 4, 8 or 16 bytes";
      return 1;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string("transpose launch failed: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// FP64 -> FP32 rounded to odd (truncate toward zero, then set the last
// mantissa bit if the result is inexact). A second rounding of that value to
// a format with at least two fewer significand bits (FP8 e4m3, bfloat16) then
// equals the direct correctly rounded conversion, which plain nearest
// rounding to FP32 first does not guarantee (a double just past an FP8
// midpoint can round to the midpoint in FP32 and then tie the wrong way).
__device__ __forceinline__ float f64_to_f32_odd(double d)
{
  float f = float(d);
  if (double(f) != d && d == d) {
    // one step toward zero (|f| > 0 here; sign-magnitude bits: magnitude - 1)
    if (__builtin_fabs(double(f)) > __builtin_fabs(d)) f = __uint_as_float(__float_as_uint(f) - 1u);
    f = __uint_as_float(__float_as_uint(f) | 1u);
  }
  return f;
}

// Precision-converting element copy helpers.
template <class CD, class CSrc>
__device__ __forceinline__ CD cvt_elem(const CSrc &v)
{
  if constexpr (sizeof(CD) == 16 && sizeof(CSrc) == 16) {
    return v;  // FP64 -> FP64: bit copy
  } else {
    using A = typename std::conditional<sizeof(CD) == 16, double, float>::type;
    if constexpr ((std::is_same<CD, c8>::value || std::is_same<CD, cb16>::value) && std::is_same<CSrc, c64>::value) {
      return StCvt<CD>::go(cpx<float>{f64_to_f32_odd(v.x), f64_to_f32_odd(v.y)});
    } else if constexpr (std::is_same<CD, c16>::value && std::is_same<CSrc, c64>::value) {
      // Direct double -> binary16 rounding (one rounding step).
      c16 r;
      r.x = _Float16(v.x);
      r.y = _Float16(v.y);
      return r;
    } else {
      return StCvt<CD>::go(ld_cvt<A>(v));
    }
  }
}

struct PtrTable {
  void *p[32];
};

// Pack by rows: one lane per (site, spin-colour) row gathers the row's kk
// RHS values (for each r, consecutive lanes read consecutive elements of
// buffer r: coalesced) and writes the kk-complex row of the packed field
// contiguously (16-byte stores when the row is 16-byte aligned). kk <= 32,
// known at compile time up to the unrolled bound so that the pointer table
// is indexed by constants.
template <class CD, class CSrc>
__global__ void __launch_bounds__(256) pack_rows_kernel(CD *__restrict__ dst, PtrTable src, int r0, int kk, int k,
                                                        long long nrows /* S*M */)
{
  const long long se = blockIdx.x * 256LL + threadIdx.x;
  if (se >= nrows) return;
  CD row[32];
#pragma unroll
  for (int r = 0; r < 32; ++r)
    if (r < kk) row[r] = cvt_elem<CD, CSrc>(static_cast<const CSrc *>(src.p[r])[se]);
  CD *d = dst + se * k + r0;
  constexpr int V = 16 / int(sizeof(CD)) > 0 ? 16 / int(sizeof(CD)) : 1;  // complex per 16-byte store
  if (V > 1 && kk % V == 0 && (reinterpret_cast<uintptr_t>(d) % 16) == 0) {
#pragma unroll
    for (int q = 0; q < 32 / V; ++q)
      if (q * V < kk) {
        CVec<CD, V> v;
#pragma unroll
        for (int j = 0; j < V; ++j) v.v[j] = row[q * V + j];
        reinterpret_cast<CVec<CD, V> *>(d)[q] = v;
      }
  } else {
#pragma unroll
    for (int r = 0; r < 32; ++r)
      if (r < kk) d[r] = row[r];
  }
}

// Unpack by rows: one lane per row reads the packed kk-complex row
// contiguously and writes element r to buffer r (for each r, consecutive
// lanes write consecutive elements).
template <class CD, class CSrc>
__global__ void __launch_bounds__(256) unpack_rows_kernel(PtrTable dst, const CSrc *__restrict__ src, int r0, int kk,
                                                          int k, long long nrows)
{
  const long long se = blockIdx.x * 256LL + threadIdx.x;
  if (se >= nrows) return;
  const CSrc *sp = src + se * k + r0;
  CSrc row[32];
  constexpr int V = 16 / int(sizeof(CSrc)) > 0 ? 16 / int(sizeof(CSrc)) : 1;
  if (V > 1 && kk % V == 0 && (reinterpret_cast<uintptr_t>(sp) % 16) == 0) {
#pragma unroll
    for (int q = 0; q < 32 / V; ++q)
      if (q * V < kk) {
        const CVec<CSrc, V> v = reinterpret_cast<const CVec<CSrc, V> *>(sp)[q];
#pragma unroll
        for (int j = 0; j < V; ++j) row[q * V + j] = v.v[j];
      }
  } else {
#pragma unroll
    for (int r = 0; r < 32; ++r)
      if (r < kk) row[r] = sp[r];
  }
#pragma unroll
  for (int r = 0; r < 32; ++r)
    if (r < kk) static_cast<CD *>(dst.p[r])[se] = cvt_elem<CD, CSrc>(row[r]);
}

// Pack through LDS, for 2^lk RHS per launch that are the whole row (kk = k):
// a workgroup takes 64 consecutive rows; each lane reads consecutive
// elements of one buffer (coalesced), the tile is transposed in LDS, and the
// 64 x kk complex destination block, contiguous, is written with
// consecutive lanes on consecutive elements.
template <class CD, class CSrc>
__global__ void __launch_bounds__(256) pack_tile_kernel(CD *__restrict__ dst, PtrTable src, int lk,
                                                        long long nrows)
{
  constexpr int TR = 64;  // rows per workgroup
  __shared__ CD tile[TR * 33];
  const int kk = 1 << lk;
  const long long row0 = blockIdx.x * (long long)TR;
  const int nr = (nrows - row0) < TR ? int(nrows - row0) : TR;
  // read: element e = r * 64 + row
  for (int e = threadIdx.x; e < TR * kk; e += 256) {
    const int r = e >> 6, row = e & 63;
    if (row < nr) tile[row * 33 + r] = cvt_elem<CD, CSrc>(static_cast<const CSrc *>(src.p[r])[row0 + row]);
  }
  __syncthreads();
  // write: element e = row * kk + r of the contiguous block
  CD *d = dst + row0 * kk;
  for (int e = threadIdx.x; e < nr * kk; e += 256) d[e] = tile[(e >> lk) * 33 + (e & (kk - 1))];
}

template <class CD, class CSrc>
static hipError_t launch_pack(bool unpack, void *packed, void *const *ptrs, int k, int M, long long S, hipStream_t st)
{
  if (!unpack && k <= 32 && !(k & (k - 1))) {
    PtrTable t{};
    for (int r = 0; r < k; ++r) t.p[r] = ptrs[r];
    const long long nrows = S * M;
    const unsigned nblk = unsigned((nrows + 63) / 64);
    hipLaunchKernelGGL((pack_tile_kernel<CD, CSrc>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(packed), t,
                       __builtin_ctz(unsigned(k)), nrows);
    return hipGetLastError();
  }
  if (!unpack) {
    for (int r0 = 0; r0 < k; r0 += 32) {
      const int kk = (k - r0) < 32 ? (k - r0) : 32;
      PtrTable t{};
      for (int r = 0; r < kk; ++r) t.p[r] = ptrs[r0 + r];
      const long long nrows = S * M;
      const unsigned nblk = unsigned((nrows + 255) / 256);
      hipLaunchKernelGGL((pack_rows_kernel<CD, CSrc>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(packed), t,
                         r0, kk, k, nrows);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  for (int r0 = 0; r0 < k; r0 += 32) {
    const int kk = (k - r0) < 32 ? (k - r0) : 32;
    PtrTable t{};
    for (int r = 0; r < kk; ++r) t.p[r] = ptrs[r0 + r];
    const long long nrows = S * M;
    const unsigned nblk = unsigned((nrows + 255) / 256);
    hipLaunchKernelGGL((unpack_rows_kernel<CD, CSrc>), dim3(nblk), dim3(256), 0, st, t,
                       static_cast<const CSrc *>(packed), r0, kk, k, nrows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <class CD>
static hipError_t pack_dispatch_src(bool unpack, int src_prec, void *packed, void *const *ptrs, int k, int M,
                                    long long S, hipStream_t st, bool *ok)
{
  *ok = true;
  // For pack: CD is the packed (destination) type, the source is the user type.
  // For unpack: CD is the user (destination) type, the source is the packed type.
  switch (src_prec) {
    case PREC_DOUBLE: return launch_pack<CD, c64>(unpack, packed, ptrs, k, M, S, st);
    case PREC_SINGLE: return launch_pack<CD, c32>(unpack, packed, ptrs, k, M, S, st);
    case PREC_HALF: return launch_pack<CD, c16>(unpack, packed, ptrs, k, M, S, st);
    case PREC_BF16: return launch_pack<CD, cb16>(unpack, packed, ptrs, k, M, S, st);
    case PREC_FP8: return launch_pack<CD, c8>(unpack, packed, ptrs, k, M, S, st);
    default: *ok = false; return hipSuccess;
  }
}

int pack_spinors(void *dst, int dst_prec, const void *const *src, int src_prec, int k, int N, long long S,
                 hipStream_t st, std::string *err)
{
  bool ok = false;
  hipError_t e = hipSuccess;
  void *const *ptrs = const_cast<void *const *>(src);
  const int M = 4 * N;
  switch (dst_prec) {
    case PREC_DOUBLE: e = pack_dispatch_src<c64>(false, src_prec, dst, ptrs, k, M, S, st, &ok); break;
    case PREC_SINGLE: e = pack_dispatch_src<c32>(false, src_prec, dst, ptrs, k, M, S, st, &ok); break;
    case PREC_HALF: e = pack_dispatch_src<c16>(false, src_prec, dst, ptrs, k, M, S, st, &ok); break;
    case PREC_BF16: e = pack_dispatch_src<cb16>(false, src_prec, dst, ptrs, k, M, S, st, &ok); break;
    case PREC_FP8: e = pack_dispatch_src<c8>(false, src_prec, dst, ptrs, k, M, S, st, &ok); break;
    default: ok = false;
  }
  if (!ok) {
    if (err) *err = "unsupported precision";
    return 1;
  }
  if (e != hipSuccess) {
    if (err) *err = std::string("pack launch failed: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

int unpack_spinors(void *const *dst, int dst_prec, const void *src, int src_prec, int k, int N, long long S,
                   hipStream_t st, std::string *err)
{
  bool ok = false;
  hipError_t e = hipSuccess;
  void *packed = const_cast<void *>(src);
  const int M = 4 * N;
  switch (dst_prec) {
    case PREC_DOUBLE: e = pack_dispatch_src<c64>(true, src_prec, packed, dst, k, M, S, st, &ok); break;
    case PREC_SINGLE: e = pack_dispatch_src<c32>(true, src_prec, packed, dst, k, M, S, st, &ok); break;
    case PREC_HALF: e = pack_dispatch_src<c16>(true, src_prec, packed, dst, k, M, S, st, &ok); break;
    case PREC_BF16: e = pack_dispatch_src<cb16>(true, src_prec, packed, dst, k, M, S, st, &ok); break;
    case PREC_FP8: e = pack_dispatch_src<c8>(true, src_prec, packed, dst, k, M, S, st, &ok); break;
    default: ok = false;
  }
  if (!ok) {
    if (err) *err = "unsupported precision";
    return 1;
  }
  if (e != hipSuccess) {
    if (err) *err = std::string("unpack launch failed: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Gauge: user [4][2][S][N][N] -> native [2][S][8][N][N]. One lane per
// (parity, site, slot, element).
template <class CD, class CSrc>
__global__ void __launch_bounds__(256) gauge_pack_kernel(CD *__restrict__ dst, const CSrc *__restrict__ src, int N,
                                                         int subLx, int Ly, int Lz, int Lt, long long nelem)
{
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= nelem) return;
  const int NN = N * N;
  const long long S = 1LL * subLx * Ly * Lz * Lt;
  const int e = int(i % NN);
  const long long q = i / NN;
  const int slot = int(q % 8);
  const long long ps = q / 8;
  const int p = int(ps / S);
  const int s = int(ps - 1LL * p * S);
  const int row = e / N, col = e % N;
  if (slot < 4) {
    dst[i] = cvt_elem<CD, CSrc>(src[((size_t(slot) * 2 + p) * S + s) * NN + e]);
    return;
  }
  // Backward link of direction mu: U_mu(x - mu)^+, stored at x - mu (parity 1-p).
  const int mu = slot - 4;
  int xh = s % subLx;
  int r = s / subLx;
  const int y = r % Ly;
  r /= Ly;
  const int z = r % Lz;
  const int t = r / Lz;
  const int eo = (y + z + t) & 1;
  int nx = xh, ny = y, nz = z, nt = t;
  if (mu == 0) nx = (eo == p) ? (xh == 0 ? subLx - 1 : xh - 1) : xh;
  if (mu == 1) ny = (y == 0) ? Ly - 1 : y - 1;
  if (mu == 2) nz = (z == 0) ? Lz - 1 : z - 1;
  if (mu == 3) nt = (t == 0) ? Lt - 1 : t - 1;
  const long long nidx = ((1LL * nt * Lz + nz) * Ly + ny) * subLx + nx;
  CSrc u = src[((size_t(mu) * 2 + (1 - p)) * S + nidx) * NN + col * N + row];
  u.y = -u.y;
  dst[i] = cvt_elem<CD, CSrc>(u);
}

template <class CD>
static hipError_t gauge_pack_src(int src_prec, void *dst, const void *src, int N, const int *L, hipStream_t st,
                                 bool *ok)
{
  const long long S = 1LL * L[0] / 2 * L[1] * L[2] * L[3];
  const long long nelem = 2LL * S * 8 * N * N;
  const unsigned nblk = unsigned((nelem + 255) / 256);
  *ok = true;
  switch (src_prec) {
    case PREC_DOUBLE:
      hipLaunchKernelGGL((gauge_pack_kernel<CD, c64>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst),
                         static_cast<const c64 *>(src), N, L[0] / 2, L[1], L[2], L[3], nelem);
      break;
    case PREC_SINGLE:
      hipLaunchKernelGGL((gauge_pack_kernel<CD, c32>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst),
                         static_cast<const c32 *>(src), N, L[0] / 2, L[1], L[2], L[3], nelem);
      break;
    default: *ok = false; return hipSuccess;
  }
  return hipGetLastError();
}

int pack_gauge(void *dst, int dst_prec, const void *src, int src_prec, int N, const int *L, hipStream_t st,
               std::string *err)
{
  bool ok = false;
  hipError_t e = hipSuccess;
  switch (dst_prec) {
    case PREC_DOUBLE: e = gauge_pack_src<c64>(src_prec, dst, src, N, L, st, &ok); break;
    case PREC_SINGLE: e = gauge_pack_src<c32>(src_prec, dst, src, N, L, st, &ok); break;
    case PREC_HALF: e = gauge_pack_src<c16>(src_prec, dst, src, N, L, st, &ok); break;
    case PREC_BF16: e = gauge_pack_src<cb16>(src_prec, dst, src, N, L, st, &ok); break;
    default: ok = false;
  }
  if (!ok) {
    if (err) *err = "unsupported gauge precision (source must be FP64 or FP32)";
    return 1;
  }
  if (e != hipSuccess) {
    if (err) *err = std::string("gauge pack launch failed: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// SU(3) two-row compression: native FP64 [2][S][8][3][3] -> [2][S][8][2][3] in
// the destination precision. One lane per (parity, site, slot). A link that
// crosses the t boundary (slot 3 at t = Lt - 1, slot 7 at t = 0) is divided by
// t_sign first. Every link must then be in SU(3) to within tol: rows 0 and 1
// orthonormal and row 2 = conj(row 0 x row 1); each violation counts into *bad.
// Link i of a native FP64 SU(3) gauge [2][S][8][3][3], divided by t_sign when
// it crosses the t boundary (slot 3 at t = Lt - 1, slot 7 at t = 0).
__device__ __forceinline__ void load_su3_link(double (&re)[3][3], double (&im)[3][3], const c64 *src, long long i,
                                              int subLx, int Ly, int Lz, int Lt, int t_sign)
{
  const long long S = 1LL * subLx * Ly * Lz * Lt;
  const int slot = int(i % 8);
  const long long s = (i / 8) % S;
  const int t = int(s / (1LL * subLx * Ly * Lz));
  const bool crosses = (slot == 3 && t == Lt - 1) || (slot == 7 && t == 0);
  const double sg = (crosses && t_sign < 0) ? -1.0 : 1.0;
  const c64 *u = src + size_t(i) * 9;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      re[r][c] = sg * u[r * 3 + c].x;
      im[r][c] = sg * u[r * 3 + c].y;
    }
}

// Distance from SU(3): rows 0 and 1 orthonormal, row 2 = conj(row 0 x row 1).
__device__ __forceinline__ double su3_distance(const double (&re)[3][3], const double (&im)[3][3])
{
  double err = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = a; b < 2; ++b) {
      double dr = 0.0, di = 0.0;  // <row a, row b>
      for (int c = 0; c < 3; ++c) {
        dr += re[a][c] * re[b][c] + im[a][c] * im[b][c];
        di += im[a][c] * re[b][c] - re[a][c] * im[b][c];
      }
      err = fmax(err, fabs(dr - (a == b ? 1.0 : 0.0)) + fabs(di));
    }
  for (int j = 0; j < 3; ++j) {
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    const double cr = re[0][j1] * re[1][j2] - im[0][j1] * im[1][j2] - (re[0][j2] * re[1][j1] - im[0][j2] * im[1][j1]);
    const double ci = re[0][j1] * im[1][j2] + im[0][j1] * re[1][j2] - (re[0][j2] * im[1][j1] + im[0][j2] * re[1][j1]);
    err = fmax(err, fabs(cr - re[2][j]) + fabs(-ci - im[2][j]));
  }
  return err;
}

template <class CD>
__global__ void __launch_bounds__(256) gauge_compress12_kernel(CD *__restrict__ dst, const c64 *__restrict__ src,
                                                               int subLx, int Ly, int Lz, int Lt, int t_sign,
                                                               double tol, long long nlinks, int *bad)
{
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= nlinks) return;
  double re[3][3], im[3][3];
  load_su3_link(re, im, src, i, subLx, Ly, Lz, Lt, t_sign);
  if (!(su3_distance(re, im) <= tol)) atomicAdd(bad, 1);
  CD *d = dst + size_t(i) * 6;
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) d[r * 3 + c] = StCvt<CD>::go(cpx<double>{re[r][c], im[r][c]});
}

int compress_gauge12(void *dst, int dst_prec, const void *native64, const int *L, int t_sign, double tol, int *bad,
                     hipStream_t st, std::string *err)
{
  const long long S = 1LL * L[0] / 2 * L[1] * L[2] * L[3];
  const long long nlinks = 2LL * S * 8;
  const unsigned nblk = unsigned((nlinks + 255) / 256);
  const c64 *src = static_cast<const c64 *>(native64);
#define QM_C12(CD)                                                                                               \
  hipLaunchKernelGGL((gauge_compress12_kernel<CD>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst), src, \
                     L[0] / 2, L[1], L[2], L[3], t_sign, tol, nlinks, bad)
  switch (dst_prec) {
    case PREC_DOUBLE: QM_C12(c64); break;
    case PREC_SINGLE: QM_C12(c32); break;
    case PREC_HALF: QM_C12(c16); break;
    case PREC_BF16: QM_C12(cb16); break;
    default:
      if (err) *err = "unsupported gauge precision";
      return 1;
  }
#undef QM_C12
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string("gauge compression launch failed: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

// Eight-real layout: U01, U02, U10 and (arg U00, arg U20) per link.
template <class CD>
__global__ void __launch_bounds__(256) gauge_compress8_kernel(CD *__restrict__ dst, const c64 *__restrict__ src,
                                                              int subLx, int Ly, int Lz, int Lt, int t_sign,
                                                              double tol, double min_n, long long nlinks, int *bad)
{
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= nlinks) return;
  double re[3][3], im[3][3];
  load_su3_link(re, im, src, i, subLx, Ly, Lz, Lt, t_sign);
  if (!(su3_distance(re, im) <= tol)) atomicAdd(bad, 1);
  const double n = re[0][1] * re[0][1] + im[0][1] * im[0][1] + re[0][2] * re[0][2] + im[0][2] * im[0][2];
  if (!(n >= min_n)) atomicAdd(bad + 1, 1);
  CD *d = dst + size_t(i) * 4;
  d[0] = StCvt<CD>::go(cpx<double>{re[0][1], im[0][1]});
  d[1] = StCvt<CD>::go(cpx<double>{re[0][2], im[0][2]});
  d[2] = StCvt<CD>::go(cpx<double>{re[1][0], im[1][0]});
  d[3] = StCvt<CD>::go(cpx<double>{atan2(im[0][0], re[0][0]), atan2(im[2][0], re[2][0])});
}

int compress_gauge8(void *dst, int dst_prec, const void *native64, const int *L, int t_sign, double tol,
                    double min_n, int *bad, hipStream_t st, std::string *err)
{
  const long long S = 1LL * L[0] / 2 * L[1] * L[2] * L[3];
  const long long nlinks = 2LL * S * 8;
  const unsigned nblk = unsigned((nlinks + 255) / 256);
  const c64 *src = static_cast<const c64 *>(native64);
#define QM_C8(CD)                                                                                               \
  hipLaunchKernelGGL((gauge_compress8_kernel<CD>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst), src, \
                     L[0] / 2, L[1], L[2], L[3], t_sign, tol, min_n, nlinks, bad)
  switch (dst_prec) {
    case PREC_DOUBLE: QM_C8(c64); break;
    case PREC_SINGLE: QM_C8(c32); break;
    case PREC_HALF: QM_C8(c16); break;
    case PREC_BF16: QM_C8(cb16); break;
    default:
      if (err) *err = "unsupported gauge precision";
      return 1;
  }
#undef QM_C8
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string("gauge compression launch failed: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Element-wise precision conversion of a contiguous array of complex numbers.
template <class CD, class CSrc>
__global__ void __launch_bounds__(256) convert_kernel(CD *__restrict__ dst, const CSrc *__restrict__ src,
                                                      long long n)
{
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i < n) dst[i] = cvt_elem<CD, CSrc>(src[i]);
}

template <class CD>
static hipError_t convert_src(int src_prec, void *dst, const void *src, long long n, hipStream_t st, bool *ok)
{
  const unsigned nblk = unsigned((n + 255) / 256);
  *ok = true;
  switch (src_prec) {
    case PREC_DOUBLE:
      hipLaunchKernelGGL((convert_kernel<CD, c64>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst),
                         static_cast<const c64 *>(src), n);
      break;
    case PREC_SINGLE:
      hipLaunchKernelGGL((convert_kernel<CD, c32>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst),
                         static_cast<const c32 *>(src), n);
      break;
    case PREC_HALF:
      hipLaunchKernelGGL((convert_kernel<CD, c16>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst),
                         static_cast<const c16 *>(src), n);
      break;
    case PREC_BF16:
      hipLaunchKernelGGL((convert_kernel<CD, cb16>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst),
                         static_cast<const cb16 *>(src), n);
      break;
    case PREC_FP8:
      hipLaunchKernelGGL((convert_kernel<CD, c8>), dim3(nblk), dim3(256), 0, st, static_cast<CD *>(dst),
                         static_cast<const c8 *>(src), n);
      break;
    default: *ok = false; return hipSuccess;
  }
  return hipGetLastError();
}

int convert_complex(void *dst, int dst_prec, const void *src, int src_prec, long long n, hipStream_t st,
                    std::string *err)
{
  bool ok = false;
  hipError_t e = hipSuccess;
  if (n <= 0) return 0;
  switch (dst_prec) {
    case PREC_DOUBLE: e = convert_src<c64>(src_prec, dst, src, n, st, &ok); break;
    case PREC_SINGLE: e = convert_src<c32>(src_prec, dst, src, n, st, &ok); break;
    case PREC_HALF: e = convert_src<c16>(src_prec, dst, src, n, st, &ok); break;
    case PREC_BF16: e = convert_src<cb16>(src_prec, dst, src, n, st, &ok); break;
    case PREC_FP8: e = convert_src<c8>(src_prec, dst, src, n, st, &ok); break;
    default: ok = false;
  }
  if (!ok) {
    if (err) *err = "unsupported precision";
    return 1;
  }
  if (e != hipSuccess) {
    if (err) *err = std::string("convert launch failed: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

}  // namespace qmrhs

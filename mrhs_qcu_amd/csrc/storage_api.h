// Host entry points of the storage kernels (all asynchronous on `st`;
// 0 on success, otherwise 1 with a message in *err).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace qmrhs {

// dst[b][c][r] = src[b][r][c] for b < nb, rows x cols elements of elem_bytes.
int transpose_batched(void *dst, const void *src, int rows, int cols, int nb, int elem_bytes, hipStream_t st,
                      std::string *err);
// k user spinors [S][4][N] (host array of device pointers) -> packed [S][4][N][k].
int pack_spinors(void *dst, int dst_prec, const void *const *src, int src_prec, int k, int N, long long S,
                 hipStream_t st, std::string *err);
// packed [S][4][N][k] -> k user spinors.
int unpack_spinors(void *const *dst, int dst_prec, const void *src, int src_prec, int k, int N, long long S,
                   hipStream_t st, std::string *err);
// user gauge [4][2][S][N][N] (FP64 or FP32) -> native [2][S][8][N][N] in dst_prec.
int pack_gauge(void *dst, int dst_prec, const void *src, int src_prec, int N, const int *L, hipStream_t st,
               std::string *err);
// native FP64 SU(3) gauge [2][S][8][3][3] -> two-row layout [2][S][8][2][3] in
// dst_prec; links crossing the t boundary divided by t_sign; *bad (device int)
// += the number of links further than tol from SU(3).
int compress_gauge12(void *dst, int dst_prec, const void *native64, const int *L, int t_sign, double tol, int *bad,
                     hipStream_t st, std::string *err);
// Same input -> eight-real layout [2][S][8][4] (U01, U02, U10, (arg U00, arg U20)); bad[0] += links further
// than tol from SU(3), bad[1] += links with |U01|^2 + |U02|^2 < min_n (the rebuild divides by it).
int compress_gauge8(void *dst, int dst_prec, const void *native64, const int *L, int t_sign, double tol,
                    double min_n, int *bad, hipStream_t st, std::string *err);
// Device-to-device copy of `bytes` (pushBackFermions): a grid-stride copy
// kernel of 1,024 workgroups for 16-byte-aligned buffers and sizes (one
// 16-RHS-field RHS at 24^3x48 FP32, 31.9 MB: 22.6 us per copy + stream
// synchronisation against 24.5 us through hipMemcpyAsync), hipMemcpyAsync
// otherwise.
int copy_device(void *dst, const void *src, size_t bytes, hipStream_t st, std::string *err);
// n complex numbers, element-wise precision conversion.
int convert_complex(void *dst, int dst_prec, const void *src, int src_prec, long long n, hipStream_t st,
                    std::string *err);

}  // namespace qmrhs

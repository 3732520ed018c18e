/* CPU oracle for the MRHS even-odd Wilson hopping term.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke
 * check of __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it. The library never calls it.
 *
 * It restates the reference kernel operation for operation:
 *   - site decoding and neighbours: include/basic_data/qcu_point.cuh:33-78 and
 *     include/kernels/wilson_dslash_4d.cuh:35-43;
 *   - link and spinor addressing: include/kernels/storageLoadStore/loadAndStore.cuh:6-62
 *     (here on the user AoS layouts the reference's shift kernels produce them
 *     from, include/kernels/storage/storage_kernels.cuh:14-76);
 *   - the eight direction blocks, in the reference order X+, X-, Y+, Y-, Z+, Z-,
 *     T+, T-: wilson_dslash_4d.cuh:59-293, each link loaded once and reused
 *     for every right-hand side (wilson_dslash_4d.cuh:61-66);
 *   - complex arithmetic: include/basic_data/qcu_complex.cuh:44-46,90-91,105;
 *   - flag = +1 for daggerFlag 0, -1 otherwise: src/qcd/qcu_wilson.cu:41.
 *
 * Layouts (complex interleaved):
 *   spinors: nrhs consecutive user spinors, each [S][4][N];
 *   gauge:   user AoS [4][2][S][N][N].
 * The loop over output sites is OpenMP-parallel.
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_DEFINE(T, NAME)                                                                                      \
  typedef struct {                                                                                                  \
    T re, im;                                                                                                       \
  } NAME##_c;                                                                                                       \
  static inline NAME##_c NAME##_mul(NAME##_c a, NAME##_c b)                                                         \
  {                                                                                                                 \
    NAME##_c r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};                                            \
    return r;                                                                                                       \
  }                                                                                                                 \
  static inline NAME##_c NAME##_add(NAME##_c a, NAME##_c b)                                                         \
  {                                                                                                                 \
    NAME##_c r = {a.re + b.re, a.im + b.im};                                                                        \
    return r;                                                                                                       \
  }                                                                                                                 \
  static inline NAME##_c NAME##_sub(NAME##_c a, NAME##_c b)                                                         \
  {                                                                                                                 \
    NAME##_c r = {a.re - b.re, a.im - b.im};                                                                        \
    return r;                                                                                                       \
  }                                                                                                                 \
  static inline NAME##_c NAME##_scale(NAME##_c a, T s)                                                              \
  {                                                                                                                 \
    NAME##_c r = {a.re * s, a.im * s};                                                                              \
    return r;                                                                                                       \
  }                                                                                                                 \
  /* multiply_i: (re, im) -> (-im, re); multiply_minus_i: (re, im) -> (im, -re) */                                  \
  static inline NAME##_c NAME##_mi(NAME##_c a)                                                                      \
  {                                                                                                                 \
    NAME##_c r = {-a.im, a.re};                                                                                     \
    return r;                                                                                                       \
  }                                                                                                                 \
  static inline NAME##_c NAME##_mmi(NAME##_c a)                                                                     \
  {                                                                                                                 \
    NAME##_c r = {a.im, -a.re};                                                                                     \
    return r;                                                                                                       \
  }                                                                                                                 \
  static inline NAME##_c NAME##_conj(NAME##_c a)                                                                    \
  {                                                                                                                 \
    NAME##_c r = {a.re, -a.im};                                                                                     \
    return r;                                                                                                       \
  }                                                                                                                 \
                                                                                                                    \
  void NAME(T *out_, const T *in_, const T *gauge_, int nrhs, int N, const int *L, const int *grid, int parity,    \
            int dagger)                                                                                             \
  {                                                                                                                 \
    const int subLx = L[0] / 2, Ly = L[1], Lz = L[2], Lt = L[3];                                                    \
    const long S = (long)subLx * Ly * Lz * Lt;                                                                      \
    const T flag = dagger ? (T)-1 : (T)1;                                                                           \
    NAME##_c *out = (NAME##_c *)out_;                                                                               \
    const NAME##_c *in = (const NAME##_c *)in_;                                                                     \
    const NAME##_c *gauge = (const NAME##_c *)gauge_;                                                               \
    const int Nc = N;                                                                                               \
    _Pragma("omp parallel for schedule(static)") for (long tid = 0; tid < S; ++tid)                                 \
    {                                                                                                               \
      NAME##_c U[64], src[32], dst_all[32 * 64];                                                                    \
      const int t = (int)(tid / ((long)Lz * Ly * subLx));                                                           \
      const int z = (int)(tid % ((long)Lz * Ly * subLx) / ((long)Ly * subLx));                                      \
      const int y = (int)(tid % ((long)Ly * subLx) / subLx);                                                        \
      const int x = (int)(tid % subLx);                                                                             \
      const int eo = (y + z + t) & 1;                                                                               \
      memset(dst_all, 0, sizeof(NAME##_c) * 4 * Nc * nrhs);                                                         \
      for (int dir = 0; dir < 8; ++dir) {                                                                           \
        const int mu = dir >> 1, fwd = !(dir & 1);                                                                  \
        /* neighbour (qcu_point.cuh:43-74) */                                                                       \
        int nx = x, ny = y, nz = z, nt = t;                                                                         \
        if (mu == 0) {                                                                                              \
          if (fwd)                                                                                                  \
            nx = x + (eo != parity) * (1 + (x == subLx - 1) * (-subLx));                                            \
          else                                                                                                      \
            nx = x + (eo == parity) * (-1 + (x == 0) * subLx);                                                      \
        } else if (mu == 1) {                                                                                       \
          ny = fwd ? y + 1 + (y == Ly - 1) * (-Ly) : y - 1 + (y == 0) * Ly;                                         \
        } else if (mu == 2) {                                                                                       \
          nz = fwd ? z + 1 + (z == Lz - 1) * (-Lz) : z - 1 + (z == 0) * Lz;                                         \
        } else {                                                                                                    \
          nt = fwd ? t + 1 + (t == Lt - 1) * (-Lt) : t - 1 + (t == 0) * Lt;                                         \
        }                                                                                                           \
        const long nidx = (((long)nt * Lz + nz) * Ly + ny) * subLx + nx;                                            \
        /* boundary skip only for a split dimension (wilson_dslash_4d.cuh:60,87,118,147,177,206,237,266) */         \
        int live = 1;                                                                                               \
        if (grid[mu] > 1) {                                                                                         \
          if (mu == 0) live = fwd ? !(x == subLx - 1 && parity != eo) : !(x == 0 && parity == eo);                  \
          if (mu == 1) live = fwd ? (y < Ly - 1) : (y >= 1);                                                        \
          if (mu == 2) live = fwd ? (z < Lz - 1) : (z >= 1);                                                        \
          if (mu == 3) live = fwd ? (t < Lt - 1) : (t >= 1);                                                        \
        }                                                                                                           \
        /* link: forward U_mu(x) at (mu, parity, x); backward U_mu(x-mu) at (mu, 1-parity, x-mu) */                 \
        const long lsite = fwd ? tid : nidx;                                                                        \
        const int lpar = fwd ? parity : 1 - parity;                                                                 \
        memcpy(U, gauge + (((long)mu * 2 + lpar) * S + lsite) * Nc * Nc, sizeof(NAME##_c) * Nc * Nc);               \
        for (int r = 0; r < nrhs; ++r) {                                                                            \
          NAME##_c *dst = dst_all + (long)r * 4 * Nc;                                                               \
          memcpy(src, in + ((long)r * S + nidx) * 4 * Nc, sizeof(NAME##_c) * 4 * Nc);                               \
          if (!live) continue;                                                                                      \
          for (int i = 0; i < Nc; ++i) {                                                                            \
            NAME##_c t1 = {0, 0}, t2 = {0, 0};                                                                      \
            for (int j = 0; j < Nc; ++j) {                                                                          \
              const NAME##_c u = fwd ? U[i * Nc + j] : NAME##_conj(U[j * Nc + i]);                                  \
              NAME##_c h1, h2;                                                                                      \
              switch (dir) {                                                                                        \
                case 0: /* X+ :74-76 */                                                                             \
                  h1 = NAME##_sub(src[0 * Nc + j], NAME##_scale(NAME##_mi(src[3 * Nc + j]), flag));                 \
                  h2 = NAME##_sub(src[1 * Nc + j], NAME##_scale(NAME##_mi(src[2 * Nc + j]), flag));                 \
                  break;                                                                                            \
                case 1: /* X- :102-107 */                                                                           \
                  h1 = NAME##_add(src[0 * Nc + j], NAME##_scale(NAME##_mi(src[3 * Nc + j]), flag));                 \
                  h2 = NAME##_add(src[1 * Nc + j], NAME##_scale(NAME##_mi(src[2 * Nc + j]), flag));                 \
                  break;                                                                                            \
                case 2: /* Y+ :134-136 */                                                                           \
                  h1 = NAME##_add(src[0 * Nc + j], NAME##_scale(src[3 * Nc + j], flag));                            \
                  h2 = NAME##_sub(src[1 * Nc + j], NAME##_scale(src[2 * Nc + j], flag));                            \
                  break;                                                                                            \
                case 3: /* Y- :162-166 */                                                                           \
                  h1 = NAME##_sub(src[0 * Nc + j], NAME##_scale(src[3 * Nc + j], flag));                            \
                  h2 = NAME##_add(src[1 * Nc + j], NAME##_scale(src[2 * Nc + j], flag));                            \
                  break;                                                                                            \
                case 4: /* Z+ :193-195 */                                                                           \
                  h1 = NAME##_sub(src[0 * Nc + j], NAME##_scale(NAME##_mi(src[2 * Nc + j]), flag));                 \
                  h2 = NAME##_add(src[1 * Nc + j], NAME##_scale(NAME##_mi(src[3 * Nc + j]), flag));                 \
                  break;                                                                                            \
                case 5: /* Z- :222-226 */                                                                           \
                  h1 = NAME##_add(src[0 * Nc + j], NAME##_scale(NAME##_mi(src[2 * Nc + j]), flag));                 \
                  h2 = NAME##_sub(src[1 * Nc + j], NAME##_scale(NAME##_mi(src[3 * Nc + j]), flag));                 \
                  break;                                                                                            \
                case 6: /* T+ :253-255 */                                                                           \
                  h1 = NAME##_sub(src[0 * Nc + j], NAME##_scale(src[2 * Nc + j], flag));                            \
                  h2 = NAME##_sub(src[1 * Nc + j], NAME##_scale(src[3 * Nc + j], flag));                            \
                  break;                                                                                            \
                default: /* T- :281-285 */                                                                          \
                  h1 = NAME##_add(src[0 * Nc + j], NAME##_scale(src[2 * Nc + j], flag));                            \
                  h2 = NAME##_add(src[1 * Nc + j], NAME##_scale(src[3 * Nc + j], flag));                            \
                  break;                                                                                            \
              }                                                                                                     \
              t1 = NAME##_add(t1, NAME##_mul(h1, u));                                                               \
              t2 = NAME##_add(t2, NAME##_mul(h2, u));                                                               \
            }                                                                                                       \
            dst[0 * Nc + i] = NAME##_add(dst[0 * Nc + i], t1);                                                      \
            dst[1 * Nc + i] = NAME##_add(dst[1 * Nc + i], t2);                                                      \
            switch (dir) {                                                                                          \
              case 0: /* :79,81 */                                                                                  \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(NAME##_mi(t1), flag));                   \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(NAME##_mi(t2), flag));                   \
                break;                                                                                              \
              case 1: /* :110,112 */                                                                                \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(NAME##_mmi(t1), flag));                  \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(NAME##_mmi(t2), flag));                  \
                break;                                                                                              \
              case 2: /* :139,141 */                                                                                \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(t1, flag));                              \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(t2, -flag));                             \
                break;                                                                                              \
              case 3: /* :169,171 */                                                                                \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(t1, -flag));                             \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(t2, flag));                              \
                break;                                                                                              \
              case 4: /* :198,200 */                                                                                \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(NAME##_mi(t1), flag));                   \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(NAME##_mmi(t2), flag));                  \
                break;                                                                                              \
              case 5: /* :229,231 */                                                                                \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(NAME##_mmi(t1), flag));                  \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(NAME##_mi(t2), flag));                   \
                break;                                                                                              \
              case 6: /* :258,260 */                                                                                \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(t1, -flag));                             \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(t2, -flag));                             \
                break;                                                                                              \
              default: /* :288,290 */                                                                               \
                dst[2 * Nc + i] = NAME##_add(dst[2 * Nc + i], NAME##_scale(t1, flag));                              \
                dst[3 * Nc + i] = NAME##_add(dst[3 * Nc + i], NAME##_scale(t2, flag));                              \
                break;                                                                                              \
            }                                                                                                       \
          }                                                                                                         \
        }                                                                                                           \
      }                                                                                                             \
      /* store (wilson_dslash_4d.cuh:295-300) */                                                                    \
      for (int r = 0; r < nrhs; ++r)                                                                                \
        memcpy(out + ((long)r * S + tid) * 4 * Nc, dst_all + (long)r * 4 * Nc, sizeof(NAME##_c) * 4 * Nc);          \
    }                                                                                                               \
  }

/* Limits of the fixed per-site scratch arrays above: N <= 8, nrhs <= 64. */
ORACLE_DEFINE(double, oracle_dslash_f64)
ORACLE_DEFINE(float, oracle_dslash_f32)

int oracle_max_nrhs(void) { return 64; }
int oracle_max_colors(void) { return 8; }

// Registration macros for the kernel instantiations (one translation unit per
// precision so the library compiles in parallel).
#pragma once
#include "dslash_dispatch.h"

#define QM_CAT2(a, b) a##b
#define QM_CAT(a, b) QM_CAT2(a, b)
#define QM_REG(...) static ::qmrhs::Reg<__VA_ARGS__> QM_CAT(qm_reg_, __COUNTER__);

// Native packed layout, full tile space (R = 1, 2; LDS staging on/off).
#define QM_PACKED_TILES(N, CS, CG, CO, A)                                          \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CG, CO, A) \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CG, CO, A)  \
  QM_REG(N, 2, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CG, CO, A) \
  QM_REG(N, 2, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CG, CO, A)

#define QM_PACKED_R4(N, CS, CG, CO, A)                                             \
  QM_REG(N, 4, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CG, CO, A) \
  QM_REG(N, 4, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CG, CO, A)

#define QM_PACKED_BASIC(N, CS, CG, CO, A)                                          \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CG, CO, A) \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CG, CO, A)

// Pointer layouts (one RHS per lane).
// User AoS pointers read and written by site groups (N <= 4, links in LDS;
// FP32 and FP16: in FP64 the held pieces cost the kernel a wave per SIMD).
#define QM_PTRS_COOP(N, CS, A) \
  QM_REG(N, 1, ::qmrhs::SL_AOS_COOP, ::qmrhs::GL_NATIVE, true, false, CS, CS, CS, A)

#define QM_PTRS(N, CS, A)                                                             \
  QM_REG(N, 1, ::qmrhs::SL_AOS_PTRS, ::qmrhs::GL_NATIVE, false, false, CS, CS, CS, A) \
  QM_REG(N, 1, ::qmrhs::SL_AOS_PTRS, ::qmrhs::GL_NATIVE, true, false, CS, CS, CS, A)  \
  QM_REG(N, 1, ::qmrhs::SL_SOA_PTRS, ::qmrhs::GL_REF_SOA, false, true, CS, CS, CS, A)

// Mixed storage for the solver: sloppy input, FP32 output.
#define QM_MIXED_TILES(N, CS)                                                                          \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CS, ::qmrhs::c32, float) \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CS, ::qmrhs::c32, float)  \
  QM_REG(N, 2, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CS, ::qmrhs::c32, float) \
  QM_REG(N, 2, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CS, ::qmrhs::c32, float)
#define QM_MIXED_R4(N, CS)                                                                             \
  QM_REG(N, 4, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CS, ::qmrhs::c32, float) \
  QM_REG(N, 4, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CS, ::qmrhs::c32, float)
#define QM_MIXED_BASIC(N, CS)                                                                          \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, false, false, CS, CS, ::qmrhs::c32, float) \
  QM_REG(N, 1, ::qmrhs::SL_PACKED, ::qmrhs::GL_NATIVE, true, false, CS, CS, ::qmrhs::c32, float)

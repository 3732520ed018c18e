// Device-side building blocks shared by every kernel of the library:
// storage element types and their conversions, a small complex type for
// accumulation, invariant-divisor integer division and the XCD-aware
// workgroup remap used by all bandwidth-bound launches on gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmrhs {

// ---------------------------------------------------------------------------
// Storage precisions (values of QCU_PRECISION plus bf16).
enum Prec : int { PREC_HALF = 0, PREC_SINGLE = 1, PREC_DOUBLE = 2, PREC_BF16 = 4, PREC_FP8 = 5 };

// Storage precision of the links that go with spinors of precision p: FP8
// spinors use binary16 links (three mantissa bits would be the whole link).
inline constexpr int gauge_prec_for(int p) { return p == PREC_FP8 ? int(PREC_HALF) : p; }

// One complex number as stored in HBM.
struct alignas(4) c16 {  // IEEE binary16 pair
  _Float16 x, y;
};
struct alignas(4) cb16 {  // bfloat16 pair, raw bits
  uint16_t x, y;
};
struct alignas(2) c8 {  // OCP FP8 E4M3 pair (4 exponent bits, bias 7, 3 mantissa bits, max 448), raw bits
  uint8_t x, y;
};
constexpr float kFp8Max = 448.0f;
using c32 = float2;
using c64 = double2;

// Real scalar type a stored complex type holds (for documentation / sizing).
template <class C> struct StoreInfo;
template <> struct StoreInfo<c64> { static constexpr int bytes = 16; static constexpr int prec = PREC_DOUBLE; };
template <> struct StoreInfo<c32> { static constexpr int bytes = 8; static constexpr int prec = PREC_SINGLE; };
template <> struct StoreInfo<c16> { static constexpr int bytes = 4; static constexpr int prec = PREC_HALF; };
template <> struct StoreInfo<cb16> { static constexpr int bytes = 4; static constexpr int prec = PREC_BF16; };
template <> struct StoreInfo<c8> { static constexpr int bytes = 2; static constexpr int prec = PREC_FP8; };

// ---------------------------------------------------------------------------
// Accumulation complex type.
template <class A>
struct cpx {
  A re, im;
};

template <class A>
__device__ __forceinline__ cpx<A> cz()
{
  return cpx<A>{A(0), A(0)};
}

// bf16 <-> f32 with round-to-nearest-even (NaN kept quiet).
__device__ __host__ __forceinline__ float bf16_to_f32(uint16_t h)
{
  union {
    uint32_t u;
    float f;
  } v;
  v.u = uint32_t(h) << 16;
  return v.f;
}
__device__ __host__ __forceinline__ uint16_t f32_to_bf16(float f)
{
  union {
    uint32_t u;
    float f;
  } v;
  v.f = f;
  if ((v.u & 0x7f800000u) == 0x7f800000u && (v.u & 0x007fffffu)) return uint16_t((v.u >> 16) | 0x40);
  uint32_t lsb = (v.u >> 16) & 1u;
  v.u += 0x7fffu + lsb;
  return uint16_t(v.u >> 16);
}

// Load a stored complex into the accumulation type.
template <class A> __device__ __forceinline__ cpx<A> ld_cvt(const c64 &v) { return cpx<A>{A(v.x), A(v.y)}; }
template <class A> __device__ __forceinline__ cpx<A> ld_cvt(const c32 &v) { return cpx<A>{A(v.x), A(v.y)}; }
template <class A> __device__ __forceinline__ cpx<A> ld_cvt(const c16 &v) { return cpx<A>{A(float(v.x)), A(float(v.y))}; }
template <class A> __device__ __forceinline__ cpx<A> ld_cvt(const cb16 &v)
{
  return cpx<A>{A(bf16_to_f32(v.x)), A(bf16_to_f32(v.y))};
}
// FP8 E4M3 pair -> two floats with one gfx950 conversion (exact).
template <class A> __device__ __forceinline__ cpx<A> ld_cvt(const c8 &v)
{
  const int w = int(uint32_t(v.x) | (uint32_t(v.y) << 8));
  const auto f = __builtin_amdgcn_cvt_pk_f32_fp8(w, false);
  return cpx<A>{A(f[0]), A(f[1])};
}

// Convert an accumulated complex to a storage complex.
template <class C> struct StCvt;
template <> struct StCvt<c64> {
  template <class A> __device__ __forceinline__ static c64 go(cpx<A> z) { return c64{double(z.re), double(z.im)}; }
};
template <> struct StCvt<c32> {
  template <class A> __device__ __forceinline__ static c32 go(cpx<A> z) { return c32{float(z.re), float(z.im)}; }
};
template <> struct StCvt<c16> {
  template <class A> __device__ __forceinline__ static c16 go(cpx<A> z)
  {
    c16 r;
    r.x = _Float16(float(z.re));
    r.y = _Float16(float(z.im));
    return r;
  }
};
template <> struct StCvt<c8> {
  // Round to nearest even, saturating at +-448: the format has no infinity, so
  // finite values beyond it and +-inf are clamped first (the conversion would
  // make them NaN). The clamp is written with comparisons so that a NaN stays
  // NaN (fminf / fmaxf would turn it into -448).
  __device__ __forceinline__ static float sat(float x) { return x > kFp8Max ? kFp8Max : (x < -kFp8Max ? -kFp8Max : x); }
  template <class A> __device__ __forceinline__ static c8 go(cpx<A> z)
  {
    const float re = sat(float(z.re));
    const float im = sat(float(z.im));
    const uint32_t w = uint32_t(__builtin_amdgcn_cvt_pk_fp8_f32(re, im, 0, false));
    c8 r;
    r.x = uint8_t(w & 0xffu);
    r.y = uint8_t((w >> 8) & 0xffu);
    return r;
  }
};
template <> struct StCvt<cb16> {
  template <class A> __device__ __forceinline__ static cb16 go(cpx<A> z)
  {
    cb16 r;
    r.x = f32_to_bf16(float(z.re));
    r.y = f32_to_bf16(float(z.im));
    return r;
  }
};

// Native vector types for the non-temporal load/store builtins.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// R consecutive stored complex numbers, loaded/stored as one vector access.
template <class C, int R>
struct alignas(sizeof(C) * R <= 16 ? sizeof(C) * R : 16) CVec {
  C v[R];
};

// ---------------------------------------------------------------------------
// Unsigned division by a run-time invariant d >= 1 (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, Fig. 4.1):
// with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (t + ((n - t) >> min(l,1))) >> max(l-1,0), t = umulhi(m, n),
// exact for every 32-bit n.
struct FastDiv {
  uint32_t d, m, s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d)
{
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  f.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 1 ? l - 1 : 0;
  return f;
}

__device__ __host__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f)
{
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t t = __umulhi(f.m, n);
#else
  uint32_t t = uint32_t((uint64_t(f.m) * n) >> 32);
#endif
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// ---------------------------------------------------------------------------
// gfx950 dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2. Remapping the hardware block id so that every XCD owns a
// contiguous range of logical blocks keeps neighbouring lattice sites (and the
// spinors they share) inside one L2.
constexpr int kNumXcd = 8;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t hw_block, uint32_t nblocks)
{
  const uint32_t q = nblocks / kNumXcd, rem = nblocks % kNumXcd;
  const uint32_t xcd = hw_block % kNumXcd, i = hw_block / kNumXcd;
  return xcd * q + (xcd < rem ? xcd : rem) + i;
}

// XCD-cyclic order: the logical workgroups are cut into chunks of `gran`,
// and the XCDs take consecutive chunks side by side (XCD x works on chunk
// 8 j + x while the others work on chunks 8 j .. 8 j + 7), each XCD in its
// chunk's own order. Needs nblocks to be a multiple of kNumXcd * gran.
__device__ __forceinline__ uint32_t xcd_cyclic(uint32_t hw_block, uint32_t gran)
{
  const uint32_t xcd = hw_block % kNumXcd, i = hw_block / kNumXcd;
  return ((i / gran) * kNumXcd + xcd) * gran + i % gran;
}

}  // namespace qmrhs

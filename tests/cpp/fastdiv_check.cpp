// Host check of the invariant-divisor division used by every kernel's index
// decoding: fdiv(n, d) == n / d for all divisors 1..4096 and a spread of
// numerators up to 2^31, plus every divisor/numerator edge case near powers of 2.
#include <cstdio>
#include <cstdint>
#include <random>

#include "qcu_device.h"

int main()
{
  std::mt19937_64 rng(1);
  long checked = 0;
  auto chk = [&](uint32_t n, uint32_t d, const qmrhs::FastDiv &f) {
    ++checked;
    if (qmrhs::fdiv(n, f) != n / d) {
      printf("FAIL n=%u d=%u got=%u\n", n, d, qmrhs::fdiv(n, f));
      return false;
    }
    return true;
  };
  for (uint32_t d = 1; d <= 4096; ++d) {
    qmrhs::FastDiv f = qmrhs::make_fastdiv(d);
    for (uint32_t n = 0; n < 3 * d + 5; ++n)
      if (!chk(n, d, f)) return 1;
    for (int i = 0; i < 2000; ++i)
      if (!chk(uint32_t(rng() & 0x7fffffffu), d, f)) return 1;
    for (uint32_t m : {0x7fffffffu, 0x7ffffffeu, 0x40000000u, 0x3fffffffu})
      if (!chk(m, d, f)) return 1;
  }
  for (int e = 12; e < 31; ++e)
    for (int off = -2; off <= 2; ++off) {
      const uint32_t d = uint32_t((1u << e) + off);
      qmrhs::FastDiv f = qmrhs::make_fastdiv(d);
      for (int i = 0; i < 20000; ++i)
        if (!chk(uint32_t(rng() & 0x7fffffffu), d, f)) return 1;
    }
  printf("OK %ld\n", checked);
  return 0;
}

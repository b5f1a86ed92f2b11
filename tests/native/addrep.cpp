// Host build of the engine's repeated addition (tn93g_add_repeat in swifttn93_amd/csrc/tn93g_math.h): m additions of
// v = 2^k (1/16 .. 1) to x, checked bit for bit against the m additions made one after the other, as tn93's site loop
// makes them.  Starting values: sums of tn93's 1/3 weight and dyadic weights, multiples of 1/16 plus 1/3, random
// doubles of 2^-4 .. 2^21, fractions of multiples of 1/3.
//
//   addrep N SEED MAXM    prints: cases checked, mismatches
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../swifttn93_amd/csrc/tn93g_math.h"

int main(int argc, char **argv)
{
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  std::mt19937_64 rng(argc > 2 ? strtoull(argv[2], nullptr, 10) : 1);
  const uint32_t maxm = argc > 3 ? (uint32_t)atol(argv[3]) : 20000;
  const double third = 1. / 3.f;  // resolutionsCount of a three-base code (refkernel.cpp:38-51)
  long bad = 0;
  for (long it = 0; it < n; it++) {
    const int k = -4 + (int)(rng() % 5);
    const double v = std::ldexp(1.0, k), inv = std::ldexp(1.0, -k);
    double x = 0;
    switch (rng() % 4) {
      case 0:
        for (int a = (int)(rng() % 6); a > 0; a--)
          x += rng() % 2 ? third : std::ldexp(1.0, -(int)(rng() % 5)) * (double)(1 + rng() % 8);
        break;
      case 1:
        x = (double)(rng() % (1u << 20)) * 0.0625 + third;
        break;
      case 2:
        x = tn93g_from_bits((tn93g_bits(1.0) & ~0xFFFFFFFFFFFFFull) | (rng() & 0xFFFFFFFFFFFFFull)) *
            std::ldexp(1.0, (int)(rng() % 25) - 4);
        break;
      default:
        x = (double)(rng() % 1000) * third * (1.0 / (double)(1 + rng() % 7));
    }
    const uint32_t m = (uint32_t)(rng() % (rng() % 2 ? 50 : maxm));
    double ref = x;
    for (uint32_t t = 0; t < m; t++) ref = ref + v;
    const double got = tn93g_add_repeat(x, v, inv, m);
    if (tn93g_bits(ref) != tn93g_bits(got)) {
      if (bad < 5) fprintf(stderr, "x=%a v=%a m=%u: %a, want %a\n", x, v, m, got, ref);
      bad++;
    }
  }
  printf("%ld %ld\n", n, bad);
  return bad ? 1 : 0;
}

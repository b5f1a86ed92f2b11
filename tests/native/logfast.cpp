// Host build of the engine's enclosed fast logarithm (tn93g_log_fast in swifttn93_amd/csrc/tn93g_math.h) checked
// against the running libm log() and against the double-double logarithm of the same header.
//   logfast N SEED  -> prints: samples outside_enclosure max_err_in_u max_halfwidth_in_u
// max_err_in_u: largest |v - log x| / (u |v|), log x from the double-double value (about 2^-70 accurate).
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../swifttn93_amd/csrc/tn93g_math.h"

static uint64_t st;
static uint64_t nxt()
{
  uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unif() { return (double)(nxt() >> 11) * 0x1p-53; }

int main(int argc, char **argv)
{
  long N = argc > 1 ? atol(argv[1]) : 1000000;
  st = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
  long outside = 0;
  double max_err = 0, max_hw = 0;
  const double u = 0x1p-53;
  for (long k = 0; k < N; k++) {
    double x;
    switch (k % 8) {
      case 0: x = 1.0 - unif() * 0.5; break;                                      // TN93 log arguments: (0.5, 1]
      case 1: x = 1.0 - unif() * 0x1p-6; break;                                   // near 1 from below
      case 2: x = 1.0 + unif() * 0x1p-5; break;                                   // near 1 from above
      case 3: x = 1.0 - ldexp(unif(), -(int)(nxt() % 50)); break;                 // every scale towards 1
      case 4: x = ldexp(0.5 + unif() * 0.5, (int)(nxt() % 2000) - 1000); break;   // any binade
      case 5: x = 1.0 + (unif() - 0.5) * 0x1p-7; break;                           // both sides of the 2^-9 switch
      case 6: x = ldexp(unif(), -1074 + (int)(nxt() % 60)); break;                // subnormal and tiny
      default: x = unif() * 2.0 + 1e-300; break;
    }
    if (!(x > 0)) continue;
    const tn93g_logenc r = tn93g_log_fast(x);
    const double ref = log(x);
    if (!(r.lo <= ref && ref <= r.hi)) {
      outside++;
      if (outside < 5) fprintf(stderr, "outside x=%a v=%a [%a, %a] libm=%a\n", x, r.v, r.lo, r.hi, ref);
    }
    const tn93g_logres dd = tn93g_log(x, 0.0);
    if (r.v != 0) {
      const double err = fabs(r.v - dd.v) / (u * fabs(r.v));
      if (err > max_err) max_err = err;
      const double hw = (r.hi - r.lo) / (2 * u * fabs(r.v));
      if (hw > max_hw) max_hw = hw;
    }
  }
  printf("%ld %ld %.3f %.3f\n", N, outside, max_err, max_hw);
  return outside ? 1 : 0;
}

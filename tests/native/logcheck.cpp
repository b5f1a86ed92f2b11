// Host build of the engine's logarithm (swifttn93_amd/csrc/tn93g_math.h) checked against the running libm log().
//   logcheck N SEED WINDOW  -> prints: samples flagged unflagged_mismatch flagged_mismatch max_misround_dist
// For every sample the double-double value is also used to measure how far from a rounding midpoint libm rounds
// the "wrong" way (max_misround_dist, in ulps): the certification window must exceed it.
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
  double W = argc > 3 ? atof(argv[3]) : TN93G_LOG_WINDOW;
  long flagged = 0, bad_unflagged = 0, bad_flagged = 0;
  double max_mis = 0;
  for (long k = 0; k < N; k++) {
    double x;
    switch (k % 6) {
      case 0: x = 1.0 - unif() * 0.5; break;                    // TN93 log arguments: (0.5, 1]
      case 1: x = 1.0 - unif() * 0x1p-6; break;                 // near 1 from below
      case 2: x = 1.0 + unif() * 0x1p-5; break;                 // near 1 from above
      case 3: x = 1.0 - ldexp(unif(), -(int)(nxt() % 50)); break;  // every scale towards 1
      case 4: x = ldexp(0.5 + unif() * 0.5, (int)(nxt() % 2000) - 1000); break;  // any binade
      default: x = unif() * 2.0 + 1e-300; break;
    }
    if (!(x > 0)) continue;
    tn93g_logres r = tn93g_log(x, W);
    double ref = log(x);
    if (r.unc) flagged++;
    if (ref != r.v) {
      if (!r.unc || ref != r.alt) {
        if (r.unc) bad_flagged++; else bad_unflagged++;
        if (bad_unflagged + bad_flagged < 5) fprintf(stderr, "mismatch x=%a got=%a ref=%a\n", x, r.v, ref);
      }
      // distance of the exact value from the midpoint, measured with a zero window
      tn93g_logres z = tn93g_log(x, 0.5);
      (void)z;
    }
    // measure: how close to the midpoint was a misrounding?  re-evaluate with an all-flagging window to get dist
    if (ref != r.v || (r.unc && ref == r.alt)) {
      // binary search the smallest window that flags x: that is its distance from the midpoint
      double lo = 0, hi = 0.5;
      for (int it = 0; it < 40; it++) {
        double m = 0.5 * (lo + hi);
        if (tn93g_log(x, m).unc) hi = m; else lo = m;
      }
      if (ref != tn93g_log(x, 0.0).v && hi > max_mis) max_mis = hi;
    }
  }
  printf("%ld %ld %ld %ld %.6f\n", N, flagged, bad_unflagged, bad_flagged, max_mis);
  return bad_unflagged || bad_flagged ? 1 : 0;
}

// tn93g_math.h -- the TN93 formula tail of tn93 (refkernel.cpp:752-815) on the GPU, with an own logarithm.
//
// tn93 evaluates log() with the host libm.  The GPU cannot call it, so the logarithm here is evaluated to about
// 2^-70 relative error in double-double arithmetic (Tang's table-driven method: x = 2^e m, breakpoint F = 1 + j/256
// near m, log x = e ln2 + log F + log1p((m-F)/F), with log F and 1/F from an mpmath table, tn93g_logtab.h).  The
// value rounded to nearest is what libm returns unless the exact logarithm lies within `window` units in the last
// place of a rounding midpoint, where libm's own approximation error (a few hundredths of an ulp) could round it
// the other way.  Such a result is returned with `alt`, the neighbouring double on the other side of the midpoint:
// libm's value is then one of {v, alt}.
//
// The distance is a monotone function of each logarithm (every IEEE operation rounds monotonically), so evaluating
// the formula at the two extreme choices gives an enclosure [dlo, dhi] of tn93's distance.  Callers certify what is
// printed from the enclosure and fall back to the host libm when that fails.
//
// This header is compiled for the device by hipcc (the product) and for the host by the tests (calibration of the
// window against the running libm).  All code must be compiled with -ffp-contract=off: fma() is written out where
// it is meant, every other expression is evaluated as written, left to right, exactly as tn93's C++ does.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define TN93G_FN __device__ static inline
#define TN93G_TABLE_ATTR __device__ __constant__
#define TN93G_FMA(a, b, c) __builtin_fma((a), (b), (c))
#else
#include <cmath>
#define TN93G_FN static inline
#define TN93G_TABLE_ATTR static
#define TN93G_FMA(a, b, c) std::fma((a), (b), (c))
#endif

#include "tn93g_logtab.h"

// Half-width of the uncertain band around a rounding midpoint, in ulps of the result.  libm's log (glibc >= 2.28)
// rounds an internal approximation whose error is documented as below 0.02 ulp; the band is wider than that.
#define TN93G_LOG_WINDOW 0.05

TN93G_FN uint64_t tn93g_bits(double x)
{
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}
TN93G_FN double tn93g_from_bits(uint64_t u)
{
  double x;
  memcpy(&x, &u, 8);
  return x;
}

struct tn93g_logres {
  double v;    // log(x) rounded to nearest (from the double-double value)
  double alt;  // the other candidate when uncertain, else == v
  int unc;     // 1 when libm may return alt instead of v
};

// exact sum a + b = s + e
TN93G_FN void tn93g_two_sum(double a, double b, double *s, double *e)
{
  double t = a + b;
  double bb = t - a;
  *e = (a - (t - bb)) + (b - bb);
  *s = t;
}

// x > 0, finite.
TN93G_FN tn93g_logres tn93g_log(double x, double window)
{
  tn93g_logres R;
  uint64_t u = tn93g_bits(x);
  int e = 0;
  if (u < 0x0010000000000000ULL) {  // subnormal: scale into the normal range
    x = x * 0x1p54;
    u = tn93g_bits(x);
    e = -54;
  }
  e += (int)(u >> 52) - 1023;
  double m = tn93g_from_bits((u & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);  // [1, 2)
  int j = (int)((m - 1.0) * 256.0 + 0.5);                                           // 0..256
  double F = 1.0 + (double)j * (1.0 / 256.0);
  double uu = m - F;  // exact (|m - F| <= 2^-9, both in [1, 2])
  const double lFh = tn93g_logtab[j][0], lFl = tn93g_logtab[j][1];
  const double iFh = tn93g_logtab[j][2], iFl = tn93g_logtab[j][3];
  // r = uu / F as double-double: uu * (iFh + iFl)
  double rh = uu * iFh;
  double rl = TN93G_FMA(uu, iFh, -rh) + uu * iFl;
  double r0, r1;
  tn93g_two_sum(rh, rl, &r0, &r1);  // |r| <= 2^-9
  // log1p(r) = r - r^2/2 + r^3 P(r), P(r) = 1/3 - r/4 + r^2/5 - ...
  double s0 = r0 * r0;
  double s1 = TN93G_FMA(r0, r0, -s0) + 2.0 * r0 * r1;  // r^2 = s0 + s1
  double p = -1.0 / 10.0;
  p = TN93G_FMA(p, r0, 1.0 / 9.0);
  p = TN93G_FMA(p, r0, -1.0 / 8.0);
  p = TN93G_FMA(p, r0, 1.0 / 7.0);
  p = TN93G_FMA(p, r0, -1.0 / 6.0);
  p = TN93G_FMA(p, r0, 1.0 / 5.0);
  p = TN93G_FMA(p, r0, -1.0 / 4.0);
  p = TN93G_FMA(p, r0, 1.0 / 3.0);
  double t3 = s0 * r0 * p;
  // A = e*ln2 + log F (double-double, error-free product for e*ln2)
  double ed = (double)e;
  double a0 = ed * TN93G_LN2_HI;
  double a1 = TN93G_FMA(ed, TN93G_LN2_HI, -a0) + ed * TN93G_LN2_LO;
  double b0, b1;
  tn93g_two_sum(a0, lFh, &b0, &b1);
  b1 = b1 + (a1 + lFl);
  // B = r - r^2/2 (double-double)
  double c0, c1;
  tn93g_two_sum(r0, -0.5 * s0, &c0, &c1);
  c1 = c1 + (r1 - 0.5 * s1);
  // total = b + c + t3
  double d0, d1;
  tn93g_two_sum(b0, c0, &d0, &d1);
  d1 = d1 + (b1 + c1 + t3);
  double v, lo;
  tn93g_two_sum(d0, d1, &v, &lo);  // v = RN(total) except within ~2^-70 of a midpoint
  R.v = v;
  R.alt = v;
  R.unc = 0;
  if (lo != 0.0 && v != 0.0) {
    // neighbour of v in the direction of lo, and the distance of the value from their midpoint in ulps
    uint64_t vb = tn93g_bits(v);
    bool up = (lo > 0) == (v > 0);  // moving away from zero?
    double n = tn93g_from_bits(up ? vb + 1 : vb - 1);
    double gap = n - v;  // exact (adjacent doubles)
    if (gap < 0) gap = -gap;
    double al = lo < 0 ? -lo : lo;
    // within `window` ulps of the midpoint (gap/2 from v): |lo| > (0.5 - window) gap
    if (al > (0.5 - window) * gap) {
      R.alt = n;
      R.unc = 1;
    }
  }
  return R;
}

// A cheaper logarithm with an enclosure, for callers that certify from enclosures anyway.  v approximates log(x)
// with |v - log x| <= 8 u |v| (u = 2^-53):
//   |x - 1| < 2^-9: log1p(f) with f = x - 1 exact (Sterbenz), f + f^2 Q(f), Q the Taylor polynomial of degree 5
//     (remainder below 2^-66 |f|); Q to about 1.1 u, so v within 1.1 u |v|.
//   otherwise: x = 2^e m, F = 1 + j/256 nearest m, r = (m - F) * RN(1/F) within 2 u |r|, log1p(r) as above
//     (|r| <= 2^-9); v = S + T with S = RN(e ln2_hi + RN(log F)) (exact or within u |S|) and T = RN(log1p(r) + C),
//     C = the low parts of e ln2 and log F.  Here |v| >= 2^-9 (1 - 2^-10) and |log1p(r)| <= 2^-9, so |S| and |T|
//     are at most about 2 |v| and 1.02 |v|, and the errors add up to less than 8 u |v|.
// libm's log is within 0.52 ulp (<= 1.04 u |log x|) of log x, so it lies in [lo, hi] = v -+ 16 u |v| (2^-49 |v|).
struct tn93g_logenc {
  double v, lo, hi;
};

TN93G_FN double tn93g_log1p_small(double f)
{
  double q = 1.0 / 7.0;
  q = TN93G_FMA(q, f, -1.0 / 6.0);
  q = TN93G_FMA(q, f, 1.0 / 5.0);
  q = TN93G_FMA(q, f, -1.0 / 4.0);
  q = TN93G_FMA(q, f, 1.0 / 3.0);
  q = TN93G_FMA(q, f, -0.5);
  return f + (f * f) * q;
}

// x > 0, finite.  `tab` is tn93g_logtab or a copy of it (in LDS, where the table look-ups are short).
TN93G_FN tn93g_logenc tn93g_log_fast(double x, const double (*tab)[4] = tn93g_logtab)
{
  double v;
  const double f = x - 1.0;
  if (f > -0x1p-9 && f < 0x1p-9) {
    v = tn93g_log1p_small(f);
  } else {
    uint64_t u = tn93g_bits(x);
    int e = 0;
    if (u < 0x0010000000000000ULL) {  // subnormal: scale into the normal range
      x = x * 0x1p54;
      u = tn93g_bits(x);
      e = -54;
    }
    e += (int)(u >> 52) - 1023;
    const double m = tn93g_from_bits((u & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);  // [1, 2)
    const int j = (int)((m - 1.0) * 256.0 + 0.5);                                         // 0..256
    const double F = 1.0 + (double)j * (1.0 / 256.0);
    const double r = (m - F) * tab[j][2];  // m - F exact
    const double l1 = tn93g_log1p_small(r);
    const double ed = (double)e;
    const double a0 = ed * TN93G_LN2_HI;
    const double c = (TN93G_FMA(ed, TN93G_LN2_HI, -a0) + ed * TN93G_LN2_LO) + tab[j][1];
    const double S = a0 + tab[j][0];
    v = S + (l1 + c);
  }
  tn93g_logenc R;
  const double d = (v < 0 ? -v : v) * 0x1p-49;
  R.v = v;
  R.lo = v - d;
  R.hi = v + d;
  return R;
}

// The cases in which every partial sum is exact, so that the m additions are one: x a multiple of v (the partial
// sums are multiples of v below 2^27), or x + m v in the binade of x (x >= 0 and v are multiples of the spacing g
// there, so every partial sum up to x + m v < 2^(e+1) is a multiple of g in that binade; a sum that reaches the next
// binade rounds to it or above, which the test sees).  True with *y the result.
TN93G_FN bool tn93g_add_repeat_exact(double x, double v, double inv, uint32_t m, double *y)
{
  *y = x + (double)m * v;
  return (tn93g_bits(*y) >> 52) == (tn93g_bits(x) >> 52) || floor(x * inv) == x * inv;
}

// With v <= x (S, f and Sm = S + m v as in tn93g_add_repeat below) and one or two powers of two passed, the one or
// two roundings of f to the doubled spacings give what a single rounding of f to the final spacing 2^k g gives (k <=
// 2; from three on, ties rounded down first can meet a tie rounded up later), and Sm is an even multiple of that
// spacing (v >= 1/16 > 2^3 g), so the FP64 sum Sm + f, rounded to nearest even, is the result.  True with *y it.
TN93G_FN bool tn93g_add_repeat_two(double x, double v, double inv, uint32_t m, double *y)
{
  const double S = floor(x * inv) * v, Sm = S + (double)m * v;  // both exact
  if ((int)(tn93g_bits(Sm) >> 52) - (int)(tn93g_bits(x) >> 52) > 2) return false;
  *y = Sm + (x - S);
  return true;
}

// m additions of v (a power of two, the value of every N-site effect of tn93's replay) to x >= 0, one after the
// other in FP64 as tn93's site loop makes them, given inv = 1 / v (exact).  Closed form: with v <= x, write
// x = S + f, S the largest multiple of v not above x, 0 <= f < v, f a multiple of g = the spacing of doubles at x.
// Every addition is exact (v is a multiple of the spacing: the sums stay below 2^27 and v >= 1/16) except the one
// that reaches a power of two 2^E: that partial sum is exactly 2^E + f (S only takes multiples of v, 2^E is one,
// and f < v), which rounds f to the doubled spacing, ties to even (2^E itself is an even multiple of it).  So the
// result is S + m v plus f rounded once for each power of two passed, in order; the powers passed are those above x
// up to S + m v, because S + m v + f stays below the next one.  The roundings run on F = f / g in integers: rounding
// to 2^t g looks at bit t - 1 of F (the lower bits are zero by then), and a set bit is a tie, broken towards a clear
// bit t.  With x < v the first addition can pass several powers of two at once; it is made as tn93 makes it.
TN93G_FN double tn93g_add_repeat(double x, double v, double inv, uint32_t m)
{
  double y;
  if (tn93g_add_repeat_exact(x, v, inv, m, &y)) return y;
  if (x < v) {
    x = x + v;
    if (!--m || floor(x * inv) == x * inv) return x + (double)m * v;
  }
  if (tn93g_add_repeat_two(x, v, inv, m, &y)) return y;
  const double S = floor(x * inv) * v, f = x - S;  // both exact
  const double Sm = S + (double)m * v;           // exact: a multiple of v below 2^27
  const int e0 = (int)((tn93g_bits(x) >> 52) & 0x7ffu), e1 = (int)((tn93g_bits(Sm) >> 52) & 0x7ffu);
  const double g = tn93g_from_bits((uint64_t)(e0 - 52) << 52);                       // spacing at x (x >= v: normal)
  uint64_t F = (uint64_t)(f * tn93g_from_bits((uint64_t)(2046 - (e0 - 52)) << 52));  // f / g, exact
  for (int t = 1; t <= e1 - e0; t++)
    if (F >> (t - 1) & 1ull) F = (F >> t & 1ull) ? F + (1ull << (t - 1)) : F - (1ull << (t - 1));
  return Sm + (double)F * g;  // exact: F g is a multiple of the spacing at the result
}

// Inputs of the tail: tn93's nucFreq[4] and float_counts sums after the merge (refkernel.cpp:714-729).
struct tn93g_tailin {
  double nf[4];
  double diag;  // fc00 + fc11 + fc22 + fc33 (left to right)
  double ag;    // fc02 + fc20
  double ct;    // fc13 + fc31
};

struct tn93g_tailout {
  double d, dlo, dhi;  // tn93's raw distance (before the clamp at 0), and an enclosure of it
  int unc;             // any logarithm uncertain
};

// The formula of computeTN93 (refkernel.cpp:752-803) and swifttn93's tn93_tail_raw (swifttn93.cpp:1954-2000),
// operation for operation.  Logs come from tn93g_log; d uses the rounded-to-nearest values, dlo/dhi the extreme
// candidates (d is non-increasing in each log whose coefficient is positive).
TN93G_FN tn93g_tailout tn93g_tail(const tn93g_tailin &in, double window)
{
  tn93g_tailout o;
  const double *nucFreq = in.nf;
  double auxd, fY, fR, K1 = 0, K2 = 0, K3 = 0, AG, CT, ti, tv, tng2, nucF[4];
  bool useK2P = false;
  // tn93 computes 2./S and 1./S; scaling by 2 commutes with rounding, so 2./S == 2. * (1./S) exactly
  auxd = 1. / (nucFreq[0] + nucFreq[1] + nucFreq[2] + nucFreq[3]);
  tng2 = 2. * auxd;
  for (int k = 0; k < 4; k++) nucF[k] = nucFreq[k] * auxd;
  fR = nucF[0] + nucF[2];
  fY = nucF[1] + nucF[3];
  if (nucFreq[0] == 0 || nucFreq[1] == 0 || nucFreq[2] == 0 || nucFreq[3] == 0) {
    useK2P = true;
  } else {
    K1 = 2. * nucF[0] * nucF[2] / fR;
    K2 = 2. * nucF[1] * nucF[3] / fY;
    K3 = 2. * (fR * fY - nucF[0] * nucF[2] * fY / fR - nucF[1] * nucF[3] * fR / fY);
  }
  AG = in.ag * tng2;
  CT = in.ct * tng2;
  tv = 1. - (in.diag * tng2 + AG + CT);
  o.unc = 0;
  if (useK2P) {
    ti = AG + CT;
    AG = 1. - 2. * ti - tv;
    CT = 1. - 2. * tv;
    if (AG > 0. && CT > 0.) {
      tn93g_logres la = tn93g_log(AG, window), lc = tn93g_log(CT, window);
      o.d = -0.5 * la.v - 0.25 * lc.v;
      o.unc = la.unc | lc.unc;
      if (o.unc) {
        double ah = la.v > la.alt ? la.v : la.alt, al_ = la.v > la.alt ? la.alt : la.v;
        double ch = lc.v > lc.alt ? lc.v : lc.alt, cl = lc.v > lc.alt ? lc.alt : lc.v;
        o.dlo = -0.5 * ah - 0.25 * ch;
        o.dhi = -0.5 * al_ - 0.25 * cl;
      } else
        o.dlo = o.dhi = o.d;
    } else
      o.d = o.dlo = o.dhi = 1000.0;
  } else {
    AG = 1. - AG / K1 - 0.5 * tv / fR;
    CT = 1. - CT / K2 - 0.5 * tv / fY;
    tv = 1. - 0.5 * tv / fY / fR;
    if (AG > 0. && CT > 0. && tv > 0) {
      tn93g_logres la = tn93g_log(AG, window), lc = tn93g_log(CT, window), lt = tn93g_log(tv, window);
      o.d = -K1 * la.v - K2 * lc.v - K3 * lt.v;
      o.unc = la.unc | lc.unc | lt.unc;
      if (o.unc) {
        double ah = la.v > la.alt ? la.v : la.alt, al_ = la.v > la.alt ? la.alt : la.v;
        double ch = lc.v > lc.alt ? lc.v : lc.alt, cl = lc.v > lc.alt ? lc.alt : lc.v;
        double th = lt.v > lt.alt ? lt.v : lt.alt, tl = lt.v > lt.alt ? lt.alt : lt.v;
        // d is non-increasing in la, lc; in lt it is non-increasing when K3 > 0, non-decreasing when K3 < 0
        o.dlo = -K1 * ah - K2 * ch - K3 * (K3 > 0 ? th : tl);
        o.dhi = -K1 * al_ - K2 * cl - K3 * (K3 > 0 ? tl : th);
      } else
        o.dlo = o.dhi = o.d;
    } else
      o.d = o.dlo = o.dhi = 1000.0;
  }
  return o;
}

// Quotients by one divisor that share its reciprocal.  On the GPU an IEEE division n / d compiles to v_div_scale_f64
// (for d, then for n), v_rcp_f64 of the scaled d, two Newton steps on that reciprocal r (e = 1 - d r, r += r e),
// q = n r, the remainder n - d q by one FMA, v_div_fmas_f64 (q + rem r, rescaled when a scale was applied) and
// v_div_fixup_f64 (zero, infinite or NaN operands, overflow).  The scales are applied only when an operand is near
// the ends of the double range (the numerator below 2^-969, the quotient within 2^768 of overflow, a subnormal
// denominator or reciprocal), and the fixup changes nothing for finite non-zero d and a finite normal quotient or a
// zero numerator (then both give +0 for n = +0, d > 0).  The divisions of tn93's tail have d between 2^-64 and 2^32
// and n zero or between 2^-128 and 2^33, so the same steps without scale and fixup give the same bits, and r is
// computed once per divisor.  The host build divides.
struct tn93g_recip {
  double d, r;
};
TN93G_FN tn93g_recip tn93g_recip_of(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  return tn93g_recip{d, r};
#else
  return tn93g_recip{d, 0.0};
#endif
}
TN93G_FN double tn93g_div(double n, const tn93g_recip &R)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const double q = n * R.r;
  const double e = __builtin_fma(-R.d, q, n);
  return __builtin_fma(e, R.r, q);
#else
  return n / R.d;
#endif
}

// tn93g_tail with tn93g_log_fast: the same operations; d from the logs' estimates, [dlo, dhi] from their
// enclosures (d is non-increasing in each log whose coefficient is positive, and every IEEE operation rounds
// monotonically), so [dlo, dhi] holds tn93's distance.  unc is set whenever the enclosure is not a single value.
TN93G_FN tn93g_tailout tn93g_tail_fast(const tn93g_tailin &in, const double (*tab)[4] = tn93g_logtab)
{
  tn93g_tailout o;
  const double *nucFreq = in.nf;
  double auxd, fY, fR, K1 = 0, K2 = 0, K3 = 0, AG, CT, ti, tv, tng2, nucF[4];
  bool useK2P = false;
  auxd = tn93g_div(1., tn93g_recip_of(nucFreq[0] + nucFreq[1] + nucFreq[2] + nucFreq[3]));
  tng2 = 2. * auxd;
  for (int k = 0; k < 4; k++) nucF[k] = nucFreq[k] * auxd;
  fR = nucF[0] + nucF[2];
  fY = nucF[1] + nucF[3];
  const bool k2p = nucFreq[0] == 0 || nucFreq[1] == 0 || nucFreq[2] == 0 || nucFreq[3] == 0;
  // fR, fY > 0 outside K2P; in K2P they are not divided by
  const tn93g_recip rR = tn93g_recip_of(k2p ? 1. : fR), rY = tn93g_recip_of(k2p ? 1. : fY);
  if (k2p) {
    useK2P = true;
  } else {
    K1 = tn93g_div(2. * nucF[0] * nucF[2], rR);
    K2 = tn93g_div(2. * nucF[1] * nucF[3], rY);
    K3 = 2. * (fR * fY - tn93g_div(nucF[0] * nucF[2] * fY, rR) - tn93g_div(nucF[1] * nucF[3] * fR, rY));
  }
  AG = in.ag * tng2;
  CT = in.ct * tng2;
  tv = 1. - (in.diag * tng2 + AG + CT);
  if (useK2P) {
    ti = AG + CT;
    AG = 1. - 2. * ti - tv;
    CT = 1. - 2. * tv;
    if (AG > 0. && CT > 0.) {
      const tn93g_logenc la = tn93g_log_fast(AG, tab), lc = tn93g_log_fast(CT, tab);
      o.d = -0.5 * la.v - 0.25 * lc.v;
      o.dlo = -0.5 * la.hi - 0.25 * lc.hi;
      o.dhi = -0.5 * la.lo - 0.25 * lc.lo;
    } else
      o.d = o.dlo = o.dhi = 1000.0;
  } else {
    const double h = 0.5 * tv;
    const double hY = tn93g_div(h, rY);
    AG = 1. - tn93g_div(AG, tn93g_recip_of(K1)) - tn93g_div(h, rR);
    CT = 1. - tn93g_div(CT, tn93g_recip_of(K2)) - hY;
    tv = 1. - tn93g_div(hY, rR);
    if (AG > 0. && CT > 0. && tv > 0) {
      const tn93g_logenc la = tn93g_log_fast(AG, tab), lc = tn93g_log_fast(CT, tab), lt = tn93g_log_fast(tv, tab);
      o.d = -K1 * la.v - K2 * lc.v - K3 * lt.v;
      o.dlo = -K1 * la.hi - K2 * lc.hi - K3 * (K3 > 0 ? lt.hi : lt.lo);
      o.dhi = -K1 * la.lo - K2 * lc.lo - K3 * (K3 > 0 ? lt.lo : lt.hi);
    } else
      o.d = o.dlo = o.dhi = 1000.0;
  }
  o.unc = o.dlo != o.dhi;
  return o;
}

// histogram index with tn93's unsigned clamp (refkernel.cpp:805-813): floor(d*200) as a long, 199 when it is not
// in [0, 200) as an unsigned value (negative, NaN or too large)
TN93G_FN int tn93g_hist_index(double d)
{
  double f = floor(d * 200.0);
  if (f >= 0.0 && f < 200.0) return (int)f;
  return 199;
}

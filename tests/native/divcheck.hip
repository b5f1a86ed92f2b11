// GPU check of the quotients the engine's fast tail computes with a shared reciprocal (tn93g_recip_of / tn93g_div
// in swifttn93_amd/csrc/tn93g_math.h) against the compiler's IEEE division, and of the whole fast tail against the
// same formula written with plain divisions (refkernel.cpp:752-803 operation for operation), bit for bit.
//
//   divcheck SEED BLOCKS    prints: quotients checked, quotient mismatches, tails checked, tail mismatches
//
// Quotients: numerators 0 or m 2^e with e in [-128, 33), denominators m 2^e with e in [-64, 32), m uniform in [1, 2)
// (the ranges of tn93's tail).  Tails: x16 count vectors of alignments of up to 2^26 sites, clean sites split over
// the four bases with random skew, zero base counts (the K2P branch) included, and the diagonal, A<->G and C<->T
// counts drawn inside what those base counts allow, as the pair kernel hands them to the tail.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../../swifttn93_amd/csrc/tn93g_math.h"

__device__ inline uint64_t mix(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double unit(uint64_t &s)  // [0, 1)
{
  s += 0x9E3779B97F4A7C15ull;
  return (double)(mix(s) >> 11) * 0x1p-53;
}

// the fast tail with the compiler's divisions (the engine's formula before the reciprocals were shared)
__device__ tn93g_tailout tail_plain(const tn93g_tailin &in)
{
  tn93g_tailout o;
  const double *nucFreq = in.nf;
  double auxd, fY, fR, K1 = 0, K2 = 0, K3 = 0, AG, CT, ti, tv, tng2, nucF[4];
  bool useK2P = false;
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
  if (useK2P) {
    ti = AG + CT;
    AG = 1. - 2. * ti - tv;
    CT = 1. - 2. * tv;
    if (AG > 0. && CT > 0.) {
      const tn93g_logenc la = tn93g_log_fast(AG), lc = tn93g_log_fast(CT);
      o.d = -0.5 * la.v - 0.25 * lc.v;
      o.dlo = -0.5 * la.hi - 0.25 * lc.hi;
      o.dhi = -0.5 * la.lo - 0.25 * lc.lo;
    } else
      o.d = o.dlo = o.dhi = 1000.0;
  } else {
    AG = 1. - AG / K1 - 0.5 * tv / fR;
    CT = 1. - CT / K2 - 0.5 * tv / fY;
    tv = 1. - 0.5 * tv / fY / fR;
    if (AG > 0. && CT > 0. && tv > 0) {
      const tn93g_logenc la = tn93g_log_fast(AG), lc = tn93g_log_fast(CT), lt = tn93g_log_fast(tv);
      o.d = -K1 * la.v - K2 * lc.v - K3 * lt.v;
      o.dlo = -K1 * la.hi - K2 * lc.hi - K3 * (K3 > 0 ? lt.hi : lt.lo);
      o.dhi = -K1 * la.lo - K2 * lc.lo - K3 * (K3 > 0 ? lt.lo : lt.hi);
    } else
      o.d = o.dlo = o.dhi = 1000.0;
  }
  o.unc = o.dlo != o.dhi;
  return o;
}

constexpr int PER_THREAD = 64;

__global__ void k_check(uint64_t seed, unsigned long long *out)
{
  uint64_t s = mix(seed ^ ((uint64_t)(blockIdx.x * blockDim.x + threadIdx.x) << 20));
  unsigned long long nq = 0, bq = 0, nt = 0, bt = 0;
  for (int it = 0; it < PER_THREAD; it++) {
    // quotients
    const double n = unit(s) < 0.01 ? 0.0 : ldexp(1.0 + unit(s), -128 + (int)(unit(s) * 161.0));
    const double d = ldexp(1.0 + unit(s), -64 + (int)(unit(s) * 96.0));
    const double q1 = n / d, q2 = tn93g_div(n, tn93g_recip_of(d));
    nq++;
    bq += tn93g_bits(q1) != tn93g_bits(q2) ? 1 : 0;
    // tails
    const double sites = floor(ldexp(1.0, 1 + (int)(unit(s) * 25.0)) * (1.0 + unit(s)));  // 2 .. 2^26 sites
    double w[4], tw = 0;
    for (int b = 0; b < 4; b++) {
      w[b] = unit(s) < 0.03 ? 0.0 : 0.05 + unit(s);
      tw += w[b];
    }
    if (tw == 0) w[0] = tw = 1;
    tn93g_tailin in;
    double used = 0;
    for (int b = 0; b < 4; b++) {
      in.nf[b] = b < 3 ? floor(32.0 * sites * w[b] / tw) : 32.0 * sites - used;
      if (in.nf[b] < 0) in.nf[b] = 0;
      used += in.nf[b];
    }
    const double dmax = 16.0 * sites;
    in.diag = floor(dmax * (1.0 - 0.3 * unit(s) * unit(s)));
    const double rest = dmax - in.diag;
    in.ag = floor(rest * unit(s));
    in.ct = floor((rest - in.ag) * unit(s));
    const tn93g_tailout a = tail_plain(in), b = tn93g_tail_fast(in);
    nt++;
    bt += (tn93g_bits(a.d) != tn93g_bits(b.d) || tn93g_bits(a.dlo) != tn93g_bits(b.dlo) ||
           tn93g_bits(a.dhi) != tn93g_bits(b.dhi))
              ? 1
              : 0;
  }
  atomicAdd(out + 0, nq);
  atomicAdd(out + 1, bq);
  atomicAdd(out + 2, nt);
  atomicAdd(out + 3, bt);
}

int main(int argc, char **argv)
{
  const uint64_t seed = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1;
  const unsigned blocks = argc > 2 ? (unsigned)strtoul(argv[2], nullptr, 10) : 4096;
  unsigned long long *d = nullptr, h[4] = {0, 0, 0, 0};
  if (hipMalloc(&d, sizeof h) != hipSuccess || hipMemset(d, 0, sizeof h) != hipSuccess) return 2;
  hipLaunchKernelGGL(k_check, dim3(blocks), dim3(256), 0, 0, seed, d);
  if (hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) return 2;
  printf("%llu %llu %llu %llu\n", h[0], h[1], h[2], h[3]);
  return h[1] || h[3] ? 1 : 0;
}

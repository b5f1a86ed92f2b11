// GPU check of the dense replay's repeated addition (tn93g_add_repeat and its inline cases in
// swifttn93_amd/csrc/tn93g_math.h, composed as the replay composes them) against the m additions made one after the
// other on the device, bit for bit.  Starting values as in addrep.cpp: sums of tn93's 1/3 weight and dyadic
// weights, multiples of 1/16 plus 1/3, random doubles of 2^-4 .. 2^21, fractions of multiples of 1/3.
//
//   addrep_gpu SEED BLOCKS MAXM    prints: cases checked, mismatches
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
__device__ inline uint64_t next(uint64_t &s)
{
  s += 0x9E3779B97F4A7C15ull;
  return mix(s);
}

constexpr int PER_THREAD = 8;

__global__ void k_check(uint64_t seed, uint32_t maxm, unsigned long long *out)
{
  uint64_t s = mix(seed ^ ((uint64_t)(blockIdx.x * blockDim.x + threadIdx.x) << 20));
  const double third = 1. / 3.f;  // resolutionsCount of a three-base code, as tn93 stores it
  unsigned long long n = 0, bad = 0;
  for (int it = 0; it < PER_THREAD; it++) {
    const int k = -4 + (int)(next(s) % 5);
    const double v = ldexp(1.0, k), inv = ldexp(1.0, -k);
    double x = 0;
    switch (next(s) % 4) {
      case 0:
        for (int a = (int)(next(s) % 6); a > 0; a--)
          x += next(s) % 2 ? third : ldexp(1.0, -(int)(next(s) % 5)) * (double)(1 + next(s) % 8);
        break;
      case 1:
        x = (double)(next(s) % (1u << 20)) * 0.0625 + third;
        break;
      case 2:
        x = tn93g_from_bits((tn93g_bits(1.0) & ~0xFFFFFFFFFFFFFull) | (next(s) & 0xFFFFFFFFFFFFFull)) *
            ldexp(1.0, (int)(next(s) % 25) - 4);
        break;
      default:
        x = (double)(next(s) % 1000) * third * (1.0 / (double)(1 + next(s) % 7));
    }
    const uint32_t m = (uint32_t)(next(s) % (next(s) % 2 ? 50 : maxm));
    double ref = x;
    for (uint32_t t = 0; t < m; t++) ref = ref + v;
    // the replay's composition: the exact cases and one or two powers of two passed inline, the rest in closed form
    double y;
    if (!(tn93g_add_repeat_exact(x, v, inv, m, &y) || (x >= v && tn93g_add_repeat_two(x, v, inv, m, &y))))
      y = tn93g_add_repeat(x, v, inv, m);
    n++;
    bad += tn93g_bits(ref) != tn93g_bits(y) ? 1 : 0;
  }
  atomicAdd(out + 0, n);
  atomicAdd(out + 1, bad);
}

int main(int argc, char **argv)
{
  const uint64_t seed = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1;
  const unsigned blocks = argc > 2 ? (unsigned)strtoul(argv[2], nullptr, 10) : 1024;
  const uint32_t maxm = argc > 3 ? (uint32_t)strtoul(argv[3], nullptr, 10) : 3000;
  unsigned long long *d = nullptr, h[2] = {0, 0};
  if (hipMalloc(&d, sizeof h) != hipSuccess || hipMemset(d, 0, sizeof h) != hipSuccess) return 2;
  hipLaunchKernelGGL(k_check, dim3(blocks), dim3(256), 0, 0, seed, maxm, d);
  if (hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) return 2;
  printf("%llu %llu\n", h[0], h[1]);
  return h[1] ? 1 : 0;
}

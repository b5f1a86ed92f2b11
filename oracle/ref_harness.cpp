// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.  A driver compiled together with the reference's own
// solver/refkernel.cpp (tn93's computeTN93 / ambiguity_handler / describe_sequence, built from the reference tree by
// oracle/Makefile into oracle/_ref/).  It dumps golden data that pins the oracle:
//
//   ref_harness table SUBSET            -> for every mode 0..4 and codes c1, c2 in 0..17, the contribution of
//                                          ref::ambiguity_handler (refkernel.cpp:819-937) from zero counts:
//                                          "mode c1 c2 amb ic[16] fc[16] (hex bits)"
//   ref_harness pairs < records.bin     -> for every record, ref::computeTN93 (refkernel.cpp:211-816) without and
//                                          with gap descriptors (describe_sequence, refkernel.cpp:104-209): the
//                                          distance bits and the histogram index it incremented (-1 if none).
//
// Record format (little endian): int32 mode, int32 subset_id, int64 L, int64 min_overlap, f64 resolve_fraction,
// then L code bytes of s1 and L code bytes of s2.  Subset ids index SUBSETS below.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <vector>

#include "refkernel.h"

static const char *SUBSETS[] = {"RY", "RYSWKM", "N", "BDHV", "RYN?", "ACGT"};

static void set_subset(const char *s)
{
  memset(ref::resolveTheseAmbigs, 0, 256);
  ref::initAlphabets(false, (char *)s, false);
}

static int dump_table(const char *subset)
{
  set_subset(subset);
  for (int mode = 0; mode < 5; mode++)
    for (unsigned c1 = 0; c1 < 18; c1++)
      for (unsigned c2 = 0; c2 < 18; c2++) {
        double fc[4][4] = {{0.}};
        long ic[4][4] = {{0}};
        unsigned long amb = 0;
        ref::ambiguity_handler(c1, c2, (char)mode, fc, ic, amb);
        printf("%d %u %u %lu", mode, c1, c2, amb);
        for (int k = 0; k < 16; k++) printf(" %ld", ic[k / 4][k % 4]);
        for (int k = 0; k < 16; k++) {
          uint64_t b;
          memcpy(&b, &fc[k / 4][k % 4], 8);
          printf(" %016" PRIx64, b);
        }
        printf("\n");
      }
  return 0;
}

static bool rd(void *p, size_t n) { return fread(p, 1, n, stdin) == n; }

static int dump_pairs()
{
  int32_t mode, sid;
  int64_t L, ov;
  double rf;
  while (rd(&mode, 4)) {
    if (!rd(&sid, 4) || !rd(&L, 8) || !rd(&ov, 8) || !rd(&rf, 8)) return 1;
    std::vector<char> s1((size_t)L), s2((size_t)L);
    if (!rd(s1.data(), (size_t)L) || !rd(s2.data(), (size_t)L)) return 1;
    set_subset(mode == 4 ? SUBSETS[sid] : nullptr);
    ref::resolve_fraction = rf;
    int64_t out_bits[2];
    long out_bin[2];
    for (int variant = 0; variant < 2; variant++) {
      unsigned long hist[200];
      memset(hist, 0, sizeof hist);
      std::vector<char> a = s1, b = s2;
      double d;
      if (variant == 0) {
        d = ref::computeTN93(a.data(), b.data(), (unsigned long)L, (char)mode, nullptr, (long)ov, hist, 200., 200, 1,
                             1, nullptr, nullptr, 0.0);
      } else {
        // descriptors with gap-run jump markers, as tn93 builds them (the markers rewrite the copies)
        ref::sequence_gap_structure d1 = ref::describe_sequence(a.data(), (unsigned long)L, 4, mode != 3);
        ref::sequence_gap_structure d2 = ref::describe_sequence(b.data(), (unsigned long)L, 4, mode != 3);
        d = ref::computeTN93(a.data(), b.data(), (unsigned long)L, (char)mode, nullptr, (long)ov, hist, 200., 200, 1,
                             1, &d1, &d2, 0.0);
      }
      memcpy(&out_bits[variant], &d, 8);
      out_bin[variant] = -1;
      for (int k = 0; k < 200; k++)
        if (hist[k]) out_bin[variant] = k;
    }
    printf("%016" PRIx64 " %ld %016" PRIx64 " %ld\n", (uint64_t)out_bits[0], out_bin[0], (uint64_t)out_bits[1],
           out_bin[1]);
  }
  return 0;
}

int main(int argc, char **argv)
{
  if (argc >= 2 && !strcmp(argv[1], "table")) return dump_table(argc >= 3 ? argv[2] : "RY");
  if (argc >= 2 && !strcmp(argv[1], "pairs")) return dump_pairs();
  fprintf(stderr, "usage: ref_harness table SUBSET | pairs < records.bin\n");
  return 2;
}

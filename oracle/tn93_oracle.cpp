// tn93_oracle.cpp -- CPU restatement of tn93 v1.1.0's standard all-pairs TN93 path, as swifttn93 runs it.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may build, load or run it; the product (swifttn93_amd/) never links or calls it.
//
// It is a plain, scalar, site-by-site restatement written for clarity, not speed: every pair is scanned site by
// site in increasing column order exactly as tn93's computeTN93 does, with libm log().  Every function cites the
// reference file:line it follows (paths relative to the reference repository).
//
// Pinning (see tests/test_oracle.py):
//   * whole-program bytes: the fuzz fixtures in tests/golden/ (outputs of the reference build, oracle/Makefile) and
//     the ten published tn93 sha256 pins of application/results/network.jsonl;
//   * handler(): the 18x18x(5 modes) contribution table dumped from the reference's own ambiguity_handler;
//   * pair_tn93(): per-pair golden vectors dumped from the reference's own computeTN93 (distance bits + bin).
//
// Build: oracle/Makefile -> oracle/_build/tn93_oracle (CLI, same command line as tn93) and
//        oracle/_build/liboracle.so (C entry points for the tests).
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <algorithm>
#include <string>
#include <utility>
#include <vector>

// Planted mistakes for the oracle's own tests (tests/test_oracle.py builds copies with -DORACLE_MUTANT=k and
// requires the golden vectors to reject each one).  0 in every real build.
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

namespace oracle {

enum { RESOLVE = 0, AVERAGE = 1, SKIP = 2, GAPMM = 3, SUBSET = 4 };  // refkernel.h:14-18
static const int GAPC = 17;                                          // refkernel.h:20, IS_GAP: code >= 17
static const int NCHAR = 15;                                          // refkernel.cpp:45 (N)
static const int HBINS = 200;                                         // swifttn93.cpp:44-45
static const double MAXD = 1000.0;                                    // refkernel.cpp:8

// Alphabet and resolution sets (refkernel.cpp:9, 19-35): bit b of RES[c] set iff base b (A,C,G,T) resolves c.
static const char ALPHA[] = "ACGTURYSWKMBDHVN?-";
static const unsigned RES[18] = {1, 2, 4, 8, 8, 5, 10, 6, 9, 12, 3, 14, 13, 11, 7, 15, 15, 0};
// resolutionsCount (refkernel.cpp:38-51): 1/|R| computed as 1./k.f in double; 0 for the gap.
static double RCNT[18];
static unsigned char RTA[256];  // resolveTheseAmbigs (refkernel.cpp:14-15, 90-101)
static signed char VALID[256];  // validFlags (refkernel.cpp:56-89), nucleotide alphabet, index map

// initAlphabets(false, subset, false) (refkernel.cpp:56-103)
static void init_alphabets(const char *subset)
{
  for (int c = 0; c < 18; c++) {
    int k = __builtin_popcount(RES[c]);
    RCNT[c] = k ? 1. / (double)(float)k : 0.;  // 1.f, 1./2.f, 1./3.f, 1./4.f, 0.f
  }
  memset(RTA, 0, sizeof RTA);
  for (int i = 0; i < 256; i++) VALID[i] = -1;
  for (int i = 0; ALPHA[i]; i++) VALID[(unsigned char)ALPHA[i]] = (signed char)i;
  if (subset)
    for (size_t k = 0; subset[k]; k++) {
      unsigned char u = (unsigned char)toupper((unsigned char)subset[k]);
      if (VALID[u] > 3) RTA[(unsigned char)VALID[u]] = 1;
    }
}

// One site's contribution: the ambiguityHandler lambda of computeTN93 (refkernel.cpp:267-380).  Called only for
// sites that are not both clean and carry no gap (the caller maps gaps to N under GAPMM).
static void handler(unsigned c1, unsigned c2, int mode, double fc[4][4], long ic[4][4], unsigned long &amb)
{
  if (mode == SKIP) return;
  if (c1 < 4) {  // c1 clean, c2 ambiguous (refkernel.cpp:268-293)
    if (RCNT[c2] > 0.) {
      if (mode == RESOLVE || (mode == SUBSET && RTA[c2])) {
        if (RES[c2] >> c1 & 1) {
          amb++;
          ic[c1][c1]++;
          return;
        }
      }
      for (int b = 0; b < 4; b++)
        if (RES[c2] >> b & 1) {
          if (ORACLE_MUTANT == 2)
            fc[b][c1] += RCNT[c2];
          else
            fc[c1][b] += RCNT[c2];
        }
    }
    return;
  }
  if (c2 < 4) {  // c2 clean, c1 ambiguous (refkernel.cpp:299-321)
    if (RCNT[c1] > 0.) {
      if (mode == RESOLVE || (mode == SUBSET && RTA[c1])) {
        if (RES[c1] >> c2 & 1) {
          amb++;
          ic[c2][c2]++;
          return;
        }
      }
      for (int a = 0; a < 4; a++)
        if (RES[c1] >> a & 1) fc[a][c2] += RCNT[c1];
    }
    return;
  }
  // both ambiguous (refkernel.cpp:323-376)
  double norm = RCNT[c1] * RCNT[c2];
  if (!(norm > 0.0)) return;
  if (mode == RESOLVE || (mode == SUBSET && RTA[c1] && RTA[c2])) {
    amb++;
    unsigned m = RES[c1] & RES[c2];
    if (m) {
      double norm2 = 1. / (long)__builtin_popcount(m);
      for (int i = 0; i < 4; i++)
        if (m >> i & 1) fc[i][i] += norm2;
      return;
    }
  }
  for (int i = 0; i < 4; i++)
    if (RES[c1] >> i & 1)
      for (int j = 0; j < 4; j++)
        if (RES[c2] >> j & 1) fc[i][j] += norm;
}

// The histogram index with tn93's unsigned clamp (refkernel.cpp:805-813; swifttn93.cpp:2003-2010).
static long hist_index(double dist)
{
  double f = floor(dist * (double)HBINS);
  long index;
  if (f != f || f >= 9.2233720368547758e18 || f < -9.2233720368547758e18)
    index = (long)0x8000000000000000ULL;  // x86 cvttsd2si result for NaN / out of range
  else
    index = (long)f;
  if (ORACLE_MUTANT == 4) {
    if (index < 0) index = 0;
    if (index >= HBINS) index = HBINS - 1;
  }
  if ((unsigned long)index >= (unsigned long)HBINS) index = HBINS - 1;
  return index;
}

// Per-pair TN93 distance in standard mode: computeTN93 (refkernel.cpp:211-816) with randomize == NULL and
// threshold <= 0.  Every site is visited in increasing order; tn93's gap jumps and skip windows only ever skip
// sites that carry a gap in one of the two sequences (refkernel.cpp:104-209, 384-619), so under every mode but
// GAPMM a site with a gap is simply not counted (refkernel.cpp:620-665 for GAPMM).
// Returns -1 when the overlap is too small (histogram untouched), else the distance clamped at 0; *bin gets the
// histogram index (or -1 when none was touched).
static double pair_tn93(const uint8_t *s1, const uint8_t *s2, long L, int mode, long min_overlap, double rf,
                        long *bin)
{
  double fc[4][4] = {{0.}};
  long ic[4][4] = {{0}};
  unsigned long amb = 0;
  for (long p = 0; p < L; p++) {
    unsigned c1 = s1[p], c2 = s2[p];
    if ((c1 | c2) < 4) {
      ic[c1][c2]++;
      continue;
    }
    bool g1 = c1 >= (unsigned)GAPC, g2 = c2 >= (unsigned)GAPC;
    if (g1 || g2) {
      if (mode != GAPMM || (g1 && g2)) continue;
      if (g1) c1 = NCHAR;
      else c2 = NCHAR;
    }
    handler(c1, c2, mode, fc, ic, amb);
  }
  // merge (refkernel.cpp:714-729)
  double tng = 0., nf[4] = {0., 0., 0., 0.};
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      double pc = (fc[a][b] += (double)ic[a][b]);
      tng += pc;
      nf[a] += pc;
      nf[ORACLE_MUTANT == 5 ? a : b] += pc;
    }
  if (tng <= (double)min_overlap) {  // refkernel.cpp:732-735
    *bin = -1;
    return -1.;
  }
  if ((mode == RESOLVE || mode == SUBSET) && rf < 1. && tng * rf <= (double)amb)  // refkernel.cpp:742-750
    return pair_tn93(s1, s2, L, AVERAGE, min_overlap, rf, bin);
  // formula (refkernel.cpp:752-803); this file is compiled with -ffp-contract=off
  bool k2p = false;
  double K1 = 0, K2 = 0, K3 = 0, F[4];
  double t2 = 2. / (nf[0] + nf[1] + nf[2] + nf[3]);
  double aux = 1. / (nf[0] + nf[1] + nf[2] + nf[3]);
  for (int k = 0; k < 4; k++) F[k] = nf[k] * aux;
  double fR = F[0] + F[2], fY = F[1] + F[3];
  if (nf[0] == 0 || nf[1] == 0 || nf[2] == 0 || nf[3] == 0)
    k2p = true;
  else {
    K1 = 2. * F[0] * F[2] / fR;
    K2 = 2. * F[1] * F[3] / fY;
    K3 = 2. * (fR * fY - F[0] * F[2] * fY / fR - F[1] * F[3] * fR / fY);
    if (ORACLE_MUTANT == 3) K3 = 2. * (fR * fY - F[0] * F[2] * fY / fR + F[1] * F[3] * fR / fY);
  }
  double AG = (fc[0][2] + fc[2][0]) * t2;
  double CT = (fc[1][3] + fc[3][1]) * t2;
  double tv = 1. - ((fc[0][0] + fc[1][1] + fc[2][2] + fc[3][3]) * t2 + AG + CT);
  if (ORACLE_MUTANT == 1) tv = 1. - ((fc[0][0] + fc[1][1] + fc[2][2] + fc[3][3]) * t2 + CT);
  double dist;
  if (k2p) {
    double ti = AG + CT;
    AG = 1. - 2. * ti - tv;
    CT = 1. - 2. * tv;
    dist = (AG > 0. && CT > 0.) ? -0.5 * log(AG) - 0.25 * log(CT) : MAXD;
  } else {
    AG = 1. - AG / K1 - 0.5 * tv / fR;
    CT = 1. - CT / K2 - 0.5 * tv / fY;
    tv = 1. - 0.5 * tv / fY / fR;
    dist = (AG > 0. && CT > 0. && tv > 0) ? -K1 * log(AG) - K2 * log(CT) - K3 * log(tv) : MAXD;
  }
  *bin = hist_index(dist);
  return dist <= 0. ? 0. : dist;  // refkernel.cpp:815
}

// ------------------------------------------------------------------------------------------------ command line
// tn93 v1.1.0 grammar (argparse.cpp:138-226, defaults argparse.hpp:8-16, parsers argparse.cpp:240-445).
enum { F_CSV = 0, F_CSVN = 1, F_HYPHY = 2 };
struct Args {
  FILE *output = stdout, *input1 = stdin, *input2 = nullptr;
  double distance = 0.015, min_distance = -1.e-10, include_prob = 1.0, resolve_fraction = 1.0;
  int ambig = RESOLVE, format = F_CSV;
  unsigned long overlap = 100;
  bool do_bootstrap = false, do_bootstrap_two_files = false, do_count = false, quiet = false, do_fst = false,
       skip_header = false, report_self = false, hamming_skip = false, subquadratic = false;
  char counts_in_name = ':', delimiter = ',';
  char *ambigs_to_resolve = nullptr;
};

static const char USAGE[] =
    "usage: tn93 [-h] [-v] [-o OUTPUT] [-t THRESHOLD] [-w MIN_THRESHOLD] [-a AMBIGS] [-g FRACTION] [-l OVERLAP] "
    "[-d COUNTS_IN_NAME] [-f FORMAT] [-s SECOND_FASTA] [-b] [-r] [-c] [-n] [-0] [-q] [-H] [-S] [-D DELIMITER]"
    "[FASTA]\n";

static void arg_error(const char *fmt, ...)  // argparse.cpp:118-128
{
  va_list ap;
  fprintf(stderr, "%stn93: error: ", USAGE);
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n");
  exit(1);
}

static const char *next_arg(int &i, int argc, const char **argv)  // argparse.cpp:130-137
{
  i++;
  if (i == argc) arg_error("ran out of command line arguments");
  return argv[i];
}

static void parse_args(Args &A, int argc, const char **argv)
{
  for (int i = 1; i < argc; ++i) {
    const char *arg = argv[i];
    if (arg[0] == '-' && arg[1] == '-') {
      if (!strcmp(arg + 2, "help")) {
        fprintf(stderr, "%s\n(help text)\n", USAGE);
        exit(1);
      } else if (!strcmp(arg + 2, "version")) {
        fprintf(stderr, "v1.1.0\n");
        exit(0);
      } else if (!strcmp(arg + 2, "hamming-skip"))
        A.hamming_skip = true;
      else if (!strcmp(arg + 2, "subquadratic"))
        A.subquadratic = true;
      else
        arg_error("unknown argument: %s", arg);
    } else if (arg[0] == '-') {
      const char o = arg[1];
      if (!strcmp(arg + 1, "h")) {
        fprintf(stderr, "%s\n(help text)\n", USAGE);
        exit(1);
      } else if (o == 'v') {
        fprintf(stderr, "v1.1.0\n");
        exit(0);
      } else if (o == 'o') {
        const char *s = next_arg(i, argc, argv);
        A.output = strcmp(s, "-") ? fopen(s, "wb") : stdout;
        if (!A.output) arg_error("failed to open the OUTPUT file %s", s);
      } else if (o == 't') {
        const char *s = next_arg(i, argc, argv);
        A.distance = atof(s);
        if (A.distance < 0.0) arg_error("genetic distance threshold must be >= 0, had: %s", s);
      } else if (o == 'w') {
        const char *s = next_arg(i, argc, argv);
        A.min_distance = atof(s);
        if (A.min_distance < 0.0 || A.min_distance > 1.0)
          arg_error("genetic minimum distance threshold must be >= 0, had: %s", s);
      } else if (o == 'l') {
        const char *s = next_arg(i, argc, argv);
        A.overlap = (unsigned long)(long)atoi(s);
        if (A.overlap == 0) arg_error("overlap must be positive, had: %s", s);
      } else if (o == 'f') {
        const char *s = next_arg(i, argc, argv);
        if (!strcmp(s, "csv")) A.format = F_CSV;
        else if (!strcmp(s, "csvn")) A.format = F_CSVN;
        else if (!strcmp(s, "hyphy")) A.format = F_HYPHY;
        else arg_error("invalid output format: %s", s);
      } else if (o == 'a') {
        const char *s = next_arg(i, argc, argv);
        if (!strcmp(s, "resolve")) A.ambig = RESOLVE;
        else if (!strcmp(s, "average")) A.ambig = AVERAGE;
        else if (!strcmp(s, "skip")) A.ambig = SKIP;
        else if (!strcmp(s, "gapmm")) A.ambig = GAPMM;
        else {
          A.ambig = SUBSET;
          A.ambigs_to_resolve = strdup(s);
        }
      } else if (o == 's') {
        const char *s = next_arg(i, argc, argv);
        A.input2 = strcmp(s, "-") ? fopen(s, "rb") : stdin;
        if (!A.input2) arg_error("failed to open the second INPUT file %s", s);
        if (A.input1 == A.input2) arg_error("input FASTA files must not both be stdin %s", s);
      } else if (o == 'd') {
        const char *s = next_arg(i, argc, argv);
        A.counts_in_name = s[0];
        if (!isprint((unsigned char)A.counts_in_name))
          arg_error("count separator must be a printable character, had: %s", s);
      } else if (o == 'u') {
        const char *s = next_arg(i, argc, argv);
        A.include_prob = atof(s);
        if (A.include_prob < 0.0 || A.include_prob > 1.0)
          arg_error("sequence inclusion probability must be >= 0, had: %s", s);
      } else if (o == 'D') {
        const char *s = next_arg(i, argc, argv);
        if (strlen(s) == 1) A.delimiter = s[0];
        else arg_error("invalid output format: %s", s);
      } else if (o == 'b') A.do_bootstrap = true;
      else if (o == 'r') A.do_bootstrap_two_files = true;
      else if (o == 'c') A.do_count = true;
      else if (o == 'n') A.skip_header = true;
      else if (o == 'q') A.quiet = true;
      else if (o == 'H') A.hamming_skip = true;
      else if (o == 'S') A.subquadratic = true;
      else if (o == 'm') A.do_fst = true;
      else if (o == '0') A.report_self = true;
      else if (o == 'g') {
        const char *s = next_arg(i, argc, argv);
        A.resolve_fraction = atof(s);
        if (A.resolve_fraction < 0.0 || A.resolve_fraction > 1.0)
          arg_error("resolve ambigous fraction must be >= 0, had: %s", s);
      } else
        arg_error("unknown argument: %s", arg);
    } else if (i == argc - 1) {
      A.input1 = strcmp(arg, "-") ? fopen(arg, "rb") : stdin;
      if (!A.input1) arg_error("failed to open the INPUT file %s", arg);
      if (A.input1 == A.input2) arg_error("input FASTA files must not both be stdin %s", arg);
    } else
      arg_error("unknown argument: %s", arg);
  }
}

// ------------------------------------------------------------------------------------------------ FASTA input
struct Seqs {
  std::vector<uint8_t> codes;  // N*L
  std::vector<std::string> names;
  std::vector<long> counts;
  long N = 0;
};

// add a finished record: the length rules of readFASTA (swifttn93.cpp:259-273); tn93 exits with status 1 after
// the message, here the reader returns 1 (LengthError unwinds to it)
struct LengthError {};
static void add_sequence(Seqs &S, size_t begin, long &firstLen)
{
  long len = (long)(S.codes.size() - begin);
  if (S.N == 0) {
    firstLen = len;
    if (firstLen <= 0) {
      std::cerr << "First sequence length must be positive." << std::endl;
      throw LengthError{};
    }
  } else if (len != firstLen) {
    std::cerr << "All sequences must have the same length (" << firstLen << "), but sequence '" << S.names.back()
              << "' had length " << len;
    throw LengthError{};
  }
  S.N++;
}

// readFASTA state machine (swifttn93.cpp:292-406): records start at '>' or '#'; a name runs to '\r' or '\n';
// sequence bytes are mapped through validFlags[toupper(c)], unknown bytes dropped; copy number = atoi of what
// follows the last separator (at least one character after it, not the first character), 1 if < 1 as unsigned.
static int read_fasta_body(FILE *F, Seqs &S, long &firstLen, char sep);
static int read_fasta(FILE *F, Seqs &S, long &firstLen, char sep)
{
  try {
    return read_fasta_body(F, S, firstLen, sep);
  } catch (const LengthError &) {
    return 1;
  }
}

static int read_fasta_body(FILE *F, Seqs &S, long &firstLen, char sep)
{
  std::vector<unsigned char> buf;
  {
    unsigned char tmp[1 << 16];
    size_t r;
    while ((r = fread(tmp, 1, sizeof tmp, F)) > 0) buf.insert(buf.end(), tmp, tmp + r);
  }
  signed char tab[256];
  for (int c = 0; c < 256; c++) tab[c] = VALID[(unsigned char)toupper(c)];
  const unsigned char *p = buf.data(), *end = p + buf.size();
  int state = 0;
  size_t seq_begin = S.codes.size();
  std::string name;
  while (p < end) {
    if (state == 0) {
      while (p < end && *p != '>' && *p != '#') p++;
      if (p == end) break;
      p++;
      state = 1;
      name.clear();
      continue;
    }
    if (state == 1) {
      const unsigned char *q = p;
      while (q < end && *q != '\n' && *q != '\r') q++;
      name.append((const char *)p, (size_t)(q - p));
      if (q == end) {
        p = q;
        break;
      }
      p = q + 1;
      long len = (long)name.size();
      if (len <= 0) {
        std::cerr << "Sequence names must be non-empty." << std::endl;
        return 1;
      }
      unsigned long count = 1UL;
      if (len >= 3) {
        long sep_loc;
        for (sep_loc = 2; sep_loc < len; sep_loc++)
          if (name[(size_t)(len - sep_loc)] == sep) break;  // names[ll - sep_loc - 1] with ll = len + 1
        if (sep_loc < len) count = (unsigned long)(long)atoi(name.c_str() + (len - sep_loc + 1));
        if (count < 1) count = 1;
      }
      S.counts.push_back((long)count);
      S.names.push_back(name);
      state = 2;
      seq_begin = S.codes.size();
      continue;
    }
    // state 2
    while (p < end) {
      signed char v = tab[*p];
      if (v < 0) break;
      S.codes.push_back((uint8_t)v);
      p++;
    }
    if (p == end) break;
    unsigned char c = *p++;
    if (c == '>' || c == '#') {
      add_sequence(S, seq_begin, firstLen);
      state = 1;
      name.clear();
    }
  }
  if (state == 2) {
    add_sequence(S, seq_begin, firstLen);
    return 0;
  }
  std::cerr << "Unexpected end of file: state " << state << std::endl;
  return 1;
}

// ------------------------------------------------------------------------------------------------ output helpers
static void dump_histogram(std::ostream &os, const char *tag, const unsigned long *h)  // swifttn93.cpp:60-74
{
  if (tag)
    os << "\t\"Histogram " << tag << "\" : [";
  else
    os << "\t\"Histogram\" : [";
  for (unsigned long k = 0; k < (unsigned long)HBINS; k++) {
    if (k) os << ',';
    os << '[' << (k + 1.) / HBINS << ',' << h[k] << ']';
  }
  os << "]" << std::endl;
}

static void put_g(std::string &o, double d)  // %g, precision 6 (fastfmt.h:1-5 semantics)
{
  char t[64];
  int k = snprintf(t, sizeof t, "%g", d);
  o.append(t, (size_t)k);
}

// ------------------------------------------------------------------------------------------------ main
// The main loop and summary of swifttn93 (swifttn93.cpp:2222-2966, transcribed from tn93's TN93.cpp), run
// single-threaded in row order.
static int run(int argc, const char **argv)
{
  Args A;
  parse_args(A, argc, argv);
  if (A.do_bootstrap || A.do_bootstrap_two_files || A.include_prob < 1.) {  // swifttn93.cpp:2226-2229
    std::cerr << "swifttn93: -b, -r and -u (random resampling) are not supported" << std::endl;
    return 1;
  }
  if (A.hamming_skip) {  // swifttn93.cpp:2230-2233
    std::cerr << "swifttn93: -H (Hamming early exit) is not supported" << std::endl;
    return 1;
  }
  init_alphabets(A.ambigs_to_resolve);
  Seqs S;
  long firstLen = 0;
  if (read_fasta(A.input1, S, firstLen, A.counts_in_name) == 1) return 1;
  unsigned long n1 = (unsigned long)S.N, n2 = 0;
  if (A.input2) {
    if (read_fasta(A.input2, S, firstLen, A.counts_in_name) == 1) return 1;
    n2 = (unsigned long)S.N - n1;
  }
  const bool do_fst = A.input2 && A.do_fst;
  const bool report_self = A.input2 == nullptr && A.report_self && !A.do_count;
  const unsigned long N = (unsigned long)S.N;
  const unsigned long pairwise = (A.input2 && !do_fst) ? n2 * n1 : (N - 1) * N / 2;  // swifttn93.cpp:2259-2261
  if (!A.quiet) {
    if (A.input2 == nullptr)
      std::cerr << "Read " << N << " sequences of length " << firstLen << std::endl
                << "Will perform " << pairwise << " pairwise distance calculations";
    else {
      std::cerr << "Read " << n1 << " sequences from file 1 and " << n2 << " sequences from file 2 of length "
                << firstLen << std::endl
                << "Will perform " << pairwise << " pairwise distance calculations";
      if (do_fst) std::cerr << ". Running in F_ST mode";
    }
  }
  const int mode = A.ambig;
  const double rf = A.resolve_fraction;
  const long min_overlap = (long)A.overlap;
  const long L = firstLen;
  if (A.format != F_HYPHY && !A.do_count && !A.skip_header) {  // swifttn93.cpp:2307-2316
    if (A.format == F_CSV)
      fprintf(A.output, "ID1%cID2%cDistance\n", A.delimiter, A.delimiter);
    else
      fprintf(A.output, "Seq1%cSeq2%cDistance\n", A.delimiter, A.delimiter);
  }
  const bool cross = A.input2 && !do_fst;
  const long upper = cross ? (long)n1 : (long)N;
  const bool hy = A.format == F_HYPHY, emit = !A.do_count;
  double max_[3] = {0., 0., 0.}, mean[3] = {0., 0., 0.}, wcounts[3] = {0., 0., 0.};
  long found = 0, skipped = 0;
  unsigned long hist[3][HBINS];
  memset(hist, 0, sizeof hist);
  std::vector<std::vector<std::pair<long, double>>> hylinks(hy && emit ? N : 0);
  // Rows are computed in parallel blocks and merged strictly in row order, so every floating-point reduction
  // follows tn93's single-thread order (per-row sums in column order, rows in row order).
  struct Row {
    std::string out;
    double lsum[3], lw[3], lmax[3];
    long llinks, lskip;
    std::vector<std::pair<long, long>> hinc;  // (bin*HBINS + index, weight)
    std::vector<std::pair<long, double>> links;
  };
  const long BLK = 512;
  std::vector<Row> rows((size_t)BLK);
  for (long i0 = 0; i0 < upper; i0 += BLK) {
    const long i1 = std::min(upper, i0 + BLK);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = i0; i < i1; i++) {
      Row &R = rows[(size_t)(i - i0)];
      const long lower = cross ? (long)n1 : i + 1;
      const long c1 = S.counts[(size_t)i];
      for (int k = 0; k < 3; k++) R.lsum[k] = R.lw[k] = R.lmax[k] = 0.;
      R.llinks = R.lskip = 0;
      R.out.clear();
      R.hinc.clear();
      R.links.clear();
      std::string &out = R.out;
      if (!cross) {  // swifttn93.cpp:2406-2434
        if (c1 > 1) {
          unsigned long svs = (unsigned long)((c1 * (c1 - 1L)) >> 1);
          int b = i < (long)n1 ? 0 : 1;
          R.lw[b] += svs;
          R.hinc.push_back(std::make_pair((long)b * HBINS, (long)svs));
        }
        if (report_self) {
          if (A.format == F_CSV) {
            out += S.names[(size_t)i];
            out += A.delimiter;
            out += S.names[(size_t)i];
            out += A.delimiter;
            out += "0\n";
          } else if (A.format == F_CSVN) {
            out += std::to_string(i);
            out += A.delimiter;
            out += std::to_string(i);
            out += A.delimiter;
            out += "0\n";
          }
        }
      }
      for (long j = lower; j < (long)N; j++) {  // swifttn93.cpp:2547-2603, 2680-2687
        long bin = 0;
        if (do_fst) {
          if (i < (long)n1) {
            if (j >= (long)n1) bin = 2;
          } else
            bin = 1;
        }
        const long w = c1 * S.counts[(size_t)j];
        long hidx;
        double d = pair_tn93(&S.codes[(size_t)i * L], &S.codes[(size_t)j * L], L, mode, min_overlap, rf, &hidx);
        if (hidx >= 0) R.hinc.push_back(std::make_pair(bin * HBINS + hidx, w));
        if (d >= A.min_distance && d <= A.distance) {
          R.llinks += w;
          if (emit) {
            if (A.format == F_CSV) {
              out += S.names[(size_t)i];
              out += A.delimiter;
              out += S.names[(size_t)j];
              out += A.delimiter;
              put_g(out, d);
              out += '\n';
            } else if (A.format == F_CSVN) {
              out += std::to_string(i);
              out += A.delimiter;
              out += std::to_string(j);
              out += A.delimiter;
              put_g(out, d);
              out += '\n';
            } else
              R.links.push_back(std::make_pair(j, d));
          }
        }
        if (d <= -0.5)
          R.lskip += 1;
        else {
          R.lsum[bin] += d * (double)w;
          R.lw[bin] += (double)w;
          if (d > R.lmax[bin]) R.lmax[bin] = d;
        }
      }
    }
    for (long i = i0; i < i1; i++) {  // row-order merge (swifttn93.cpp:2832-2848)
      Row &R = rows[(size_t)(i - i0)];
      fwrite(R.out.data(), 1, R.out.size(), A.output);
      found += R.llinks;
      skipped += R.lskip;
      for (auto &h : R.hinc) hist[h.first / HBINS][h.first % HBINS] += (unsigned long)h.second;
      for (auto &lk : R.links) {
        hylinks[(size_t)i].push_back(lk);
        hylinks[(size_t)lk.first].push_back(std::make_pair(i, lk.second));
      }
      for (int k = 0; k < 3; k++) {
        if (R.lmax[k] > max_[k]) max_[k] = R.lmax[k];
        mean[k] += R.lsum[k];
        wcounts[k] += R.lw[k];
      }
    }
  }
  if (emit && hy) {  // swifttn93.cpp:2869-2897
    std::string o = "{";
    for (long r = 0; r < (long)N; r++) {
      o += "\n{";
      auto &lk = hylinks[(size_t)r];
      size_t li = 0;
      for (long c = 0; c < (long)N; c++) {
        if (c) o += ',';
        if (c == r)
          o += '0';
        else if (li < lk.size() && lk[li].first == c) {
          put_g(o, lk[li].second);
          li++;
        } else
          o += "100";
      }
      o += '}';
    }
    o += "\n}";
    fwrite(o.data(), 1, o.size(), A.output);
  }
  if (!A.quiet)  // swifttn93.cpp:2900-2904
    std::cerr << std::endl
              << "Progress:" << std::setw(8) << 100. << "% (" << std::setw(8) << found << " links found, "
              << std::setw(12) << std::setprecision(3) << (double)pairwise << " evals)";
  std::cerr << std::endl;
  std::ostream &os = A.output == stdout ? std::cerr : std::cout;  // swifttn93.cpp:2908-2912
  os << "{" << std::endl;
  os << '\t' << "\"Actual comparisons performed\" :" << pairwise - skipped << ',' << std::endl;
  os << "\t\"Comparisons accounting for copy numbers \" :" << (wcounts[0] + wcounts[1] + wcounts[2]) << ','
     << std::endl;
  os << "\t\"Total comparisons possible\" : " << pairwise << ',' << std::endl;
  os << "\t\"Links found\" : " << found << ',' << std::endl;
  os << "\t\"Maximum distance\" : " << max_[0] << ',' << std::endl;
  if (A.input2 == nullptr)
    os << "\t\"Sequences\" : " << N << ',' << std::endl;
  else {
    os << "\t\"Sequences in first file\" : " << n1 << ',' << std::endl;
    os << "\t\"Sequences in second file\" : " << n2 << ',' << std::endl;
  }
  if (do_fst) {  // swifttn93.cpp:2929-2953
    const char *keys[4] = {"File 1", "File 2", "Between", "Combined"};
    for (int k = 0; k < 3; k++)
      os << "\t\"Mean distance " << keys[k] << "\" : " << mean[k] / wcounts[k] << ',' << std::endl;
    double meta = (mean[0] + mean[1] + mean[2]) / (wcounts[0] + wcounts[1] + wcounts[2]),
           intra = (mean[0] + mean[1]) / (wcounts[0] + wcounts[1]), pi_D = mean[2] / wcounts[2] - intra;
    os << "\t\"Mean distance [intra] " << "\" : " << intra << ',' << std::endl;
    os << "\t\"Mean distance [combined]" << "\" : " << meta << ',' << std::endl;
    os << "\t\"F_ST\" : " << pi_D / (intra + pi_D) << ',' << std::endl;
    for (int k = 0; k < 3; k++) {
      dump_histogram(os, keys[k], hist[k]);
      os << ',' << std::endl;
    }
    for (int k = 0; k < HBINS; k++) hist[0][k] += hist[1][k] + hist[2][k];
    dump_histogram(os, keys[3], hist[0]);
  } else {
    os << "\t\"Mean distance\" : " << mean[0] / wcounts[0] << ',' << std::endl;
    dump_histogram(os, nullptr, hist[0]);
  }
  os << '}' << std::endl;
  if (A.do_count)
    fprintf(A.output, "Found %ld links among %ld pairwise comparisons\n", found, (long)(pairwise - skipped));
  if (A.output && A.output != stdout) fclose(A.output);
  return 0;
}

}  // namespace oracle

// ------------------------------------------------------------------------------------------------ C entry points
extern "C" {
// Sets up the alphabet tables; subset may be NULL.
void oracle_init(const char *subset) { oracle::init_alphabets(subset); }

// One site's contribution for codes (c1, c2) under mode (gaps are mapped by the caller): fc[16], ic[16], amb.
void oracle_handler(unsigned c1, unsigned c2, int mode, double *fc16, long *ic16, unsigned long *amb)
{
  double fc[4][4] = {{0.}};
  long ic[4][4] = {{0}};
  unsigned long a = 0;
  oracle::handler(c1, c2, mode, fc, ic, a);
  memcpy(fc16, fc, sizeof fc);
  memcpy(ic16, ic, sizeof ic);
  *amb = a;
}

// Per-pair distance (clamped, -1 when skipped) and histogram index (-1 when untouched).
double oracle_pair(const uint8_t *s1, const uint8_t *s2, long L, int mode, long min_overlap, double rf, long *bin)
{
  return oracle::pair_tn93(s1, s2, L, mode, min_overlap, rf, bin);
}

// Row-major distances of the full upper triangle (i < j) of n sequences of length L: out[i*n + j].
void oracle_triangle(const uint8_t *codes, long n, long L, int mode, long min_overlap, double rf, double *out,
                     long *bins)
{
#pragma omp parallel for schedule(dynamic, 1)
  for (long i = 0; i < n; i++)
    for (long j = i + 1; j < n; j++) {
      long b;
      out[i * n + j] = oracle::pair_tn93(codes + i * L, codes + j * L, L, mode, min_overlap, rf, &b);
      if (bins) bins[i * n + j] = b;
    }
}

long oracle_hist_index(double d) { return oracle::hist_index(d); }

// Reads path1 (and path2 when not NULL, as tn93's second file) with the reader above; returns a malloc'd dump:
// "rc <rc>\nerr <bytes written to std::cerr>\nN <n> L <L>\n" then one line per record: name, tab, copy number, tab,
// codes as letters of ACGTURYSWKMBDHVN?-.  The caller frees it.
char *oracle_fasta_dump(const char *path1, const char *path2, char sep)
{
  oracle::init_alphabets(nullptr);
  oracle::Seqs S;
  long firstLen = 0;
  std::ostringstream err;
  std::streambuf *old = std::cerr.rdbuf(err.rdbuf());
  int rc = 0;
  for (const char *path : {path1, path2}) {
    if (!path || rc) continue;
    FILE *F = fopen(path, "rb");
    if (!F) {
      rc = 2;
      break;
    }
    rc = oracle::read_fasta(F, S, firstLen, sep);
    fclose(F);
  }
  std::cerr.rdbuf(old);
  std::string o = "rc " + std::to_string(rc) + "\nerr " + err.str() + "\n";
  if (rc == 0) {
    o += "N " + std::to_string(S.N) + " L " + std::to_string(firstLen) + "\n";
    for (long k = 0; k < S.N; k++) {
      o += S.names[(size_t)k] + "\t" + std::to_string(S.counts[(size_t)k]) + "\t";
      for (long c = 0; c < firstLen; c++) o += "ACGTURYSWKMBDHVN?-"[S.codes[(size_t)(k * firstLen + c)]];
      o += "\n";
    }
  }
  char *r = (char *)malloc(o.size() + 1);
  memcpy(r, o.c_str(), o.size() + 1);
  return r;
}
}

#ifdef ORACLE_MAIN
int main(int argc, const char **argv) { return oracle::run(argc, argv); }
#endif

// tn93g_cli.cpp -- tn93 v1.1.0-compatible command line over libtn93g (include/tn93g.h).
//
// Same command line (argparse.cpp:138-445), same reader (swifttn93.cpp:152-406), same distance file and JSON summary
// (swifttn93.cpp:2222-2966) as single-thread tn93; the pair computation runs on the GPU through tn93g_run, the
// reader and the distance-file formatting on all host threads (tn93g_fasta.h).
// Printed values that depend on a GPU logarithm are certified from the engine's enclosures: every printed string
// must be the same at both ends of the enclosure, otherwise the run is repeated with TN93G_FLAG_EXACT (logarithms
// resolved with the host libm).
#include <atomic>
#include <chrono>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <memory>
#include <fstream>
#include <unistd.h>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tn93g.h"
#include "tn93g_fasta.h"

namespace {

enum { F_CSV = 0, F_CSVN = 1, F_HYPHY = 2 };
constexpr int RESOLVE = 0, AVERAGE = 1, SKIP = 2, GAPMM = 3, SUBSET = 4;
const char ALPHA[] = "ACGTURYSWKMBDHVN?-";

// tn93 ends the process on every error; here errors unwind to main (or to the batch driver) as Exit
struct Exit {
  int code;
};

struct Args {
  FILE *output = stdout, *input1 = stdin, *input2 = nullptr;
  ~Args()
  {
    if (output && output != stdout) fclose(output);
    if (input1 && input1 != stdin) fclose(input1);
    if (input2 && input2 != stdin) fclose(input2);
  }
  double distance = 0.015, min_distance = -1.e-10, include_prob = 1.0, resolve_fraction = 1.0;
  int ambig = RESOLVE, format = F_CSV;
  unsigned long overlap = 100;
  bool do_bootstrap = false, do_bootstrap_two_files = false, do_count = false, quiet = false, do_fst = false,
       skip_header = false, report_self = false, hamming_skip = false, subquadratic = false;
  char counts_in_name = ':', delimiter = ',';
  const char *ambigs_to_resolve = nullptr;
};

const char USAGE[] =
    "usage: tn93 [-h] [-v] [-o OUTPUT] [-t THRESHOLD] [-w MIN_THRESHOLD] [-a AMBIGS] [-g FRACTION] [-l OVERLAP] "
    "[-d COUNTS_IN_NAME] [-f FORMAT] [-s SECOND_FASTA] [-b] [-r] [-c] [-n] [-0] [-q] [-H] [-S] [-D DELIMITER]"
    "[FASTA]\n";

[[noreturn]] void arg_error(const char *f, ...)  // argparse.cpp:118-128
{
  va_list ap;
  fprintf(stderr, "%stn93: error: ", USAGE);
  va_start(ap, f);
  vfprintf(stderr, f, ap);
  va_end(ap);
  fprintf(stderr, "\n");
  throw Exit{1};
}

const char *next_arg(int &i, int argc, const char **argv)
{
  if (++i == argc) arg_error("ran out of command line arguments");
  return argv[i];
}

// tn93's grammar (argparse.cpp:138-226): options are told apart by their first character; the FASTA file must be
// the last argument
void parse_args(Args &A, int argc, const char **argv)
{
  for (int i = 1; i < argc; ++i) {
    const char *arg = argv[i];
    if (arg[0] == '-' && arg[1] == '-') {
      if (!strcmp(arg + 2, "help")) {
        fprintf(stderr, "%s\ncompute Tamura Nei 93 distances between aligned sequences (tn93g, gfx950)\n", USAGE);
        throw Exit{1};
      } else if (!strcmp(arg + 2, "version")) {
        fprintf(stderr, "v1.1.0\n");
        throw Exit{0};
      } else if (!strcmp(arg + 2, "hamming-skip"))
        A.hamming_skip = true;
      else if (!strcmp(arg + 2, "subquadratic"))
        A.subquadratic = true;
      else
        arg_error("unknown argument: %s", arg);
      continue;
    }
    if (arg[0] != '-') {
      if (i != argc - 1) arg_error("unknown argument: %s", arg);
      A.input1 = strcmp(arg, "-") ? fopen(arg, "rb") : stdin;
      if (!A.input1) arg_error("failed to open the INPUT file %s", arg);
      if (A.input1 == A.input2) arg_error("input FASTA files must not both be stdin %s", arg);
      continue;
    }
    const char o = arg[1];
    if (!strcmp(arg + 1, "h")) {
      fprintf(stderr, "%s\ncompute Tamura Nei 93 distances between aligned sequences (tn93g, gfx950)\n", USAGE);
      throw Exit{1};
    } else if (o == 'v') {
      fprintf(stderr, "v1.1.0\n");
      throw Exit{0};
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
        A.ambigs_to_resolve = s;
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
      if (strlen(s) != 1) arg_error("invalid output format: %s", s);
      A.delimiter = s[0];
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
  }
}

using tn93g_io::Seqs;

signed char g_tab[256];

// readFASTA (swifttn93.cpp:292-406) on all host threads
int read_fasta(FILE *F, Seqs &S, long &firstLen, char sep)
{
  const tn93g_io::Input buf(F);
  return tn93g_io::read_fasta(buf.data(), buf.size(), S, firstLen, sep, g_tab, tn93g_io::io_threads());
}

int fmt_g(char *out, double d) { return snprintf(out, 64, "%g", d); }

void dump_histogram(std::ostream &os, const char *tag, const uint64_t *h)  // swifttn93.cpp:60-74
{
  if (tag)
    os << "\t\"Histogram " << tag << "\" : [";
  else
    os << "\t\"Histogram\" : [";
  for (unsigned long k = 0; k < 200; k++) {
    if (k) os << ',';
    os << '[' << (k + 1.) / 200 << ',' << (unsigned long)h[k] << ']';
  }
  os << "]" << std::endl;
}

struct Summary {
  unsigned long pairwise, n1, n2, N;
  long found, skipped;
  double weighted[3], max0, mean[3];
  const uint64_t *hist;  // [3][200]
};

// The JSON summary (swifttn93.cpp:2914-2958), written to `os` with the stream's current precision.
void write_json(std::ostream &os, const Summary &s, bool two_files, bool do_fst, uint64_t hist0_combined[200])
{
  os << "{" << std::endl;
  os << '\t' << "\"Actual comparisons performed\" :" << s.pairwise - (unsigned long)s.skipped << ',' << std::endl;
  os << "\t\"Comparisons accounting for copy numbers \" :" << (s.weighted[0] + s.weighted[1] + s.weighted[2]) << ','
     << std::endl;
  os << "\t\"Total comparisons possible\" : " << s.pairwise << ',' << std::endl;
  os << "\t\"Links found\" : " << s.found << ',' << std::endl;
  os << "\t\"Maximum distance\" : " << s.max0 << ',' << std::endl;
  if (!two_files)
    os << "\t\"Sequences\" : " << s.N << ',' << std::endl;
  else {
    os << "\t\"Sequences in first file\" : " << s.n1 << ',' << std::endl;
    os << "\t\"Sequences in second file\" : " << s.n2 << ',' << std::endl;
  }
  if (do_fst) {
    const char *keys[4] = {"File 1", "File 2", "Between", "Combined"};
    const double *mean = s.mean, *wc = s.weighted;
    for (int k = 0; k < 3; k++) os << "\t\"Mean distance " << keys[k] << "\" : " << mean[k] / wc[k] << ',' << std::endl;
    double meta = (mean[0] + mean[1] + mean[2]) / (wc[0] + wc[1] + wc[2]),
           intra = (mean[0] + mean[1]) / (wc[0] + wc[1]), pi_D = mean[2] / wc[2] - intra;
    os << "\t\"Mean distance [intra] " << "\" : " << intra << ',' << std::endl;
    os << "\t\"Mean distance [combined]" << "\" : " << meta << ',' << std::endl;
    os << "\t\"F_ST\" : " << pi_D / (intra + pi_D) << ',' << std::endl;
    for (int k = 0; k < 3; k++) {
      dump_histogram(os, keys[k], s.hist + 200 * k);
      os << ',' << std::endl;
    }
    for (int k = 0; k < 200; k++) hist0_combined[k] = s.hist[k] + s.hist[200 + k] + s.hist[400 + k];
    dump_histogram(os, keys[3], hist0_combined);
  } else {
    os << "\t\"Mean distance\" : " << s.mean[0] / s.weighted[0] << ',' << std::endl;
    dump_histogram(os, nullptr, s.hist);
  }
  os << '}' << std::endl;
}

tn93g_ctx *g_ctx = nullptr;  // opened on first use, shared by the cases of a batch

struct SetGuard {
  tn93g_set *s = nullptr;
  ~SetGuard() { tn93g_set_free(s); }
};
struct ResGuard {
  tn93g_result *r = nullptr;
  ~ResGuard() { tn93g_result_free(r); }
};

// One tn93 invocation in three phases: prepare (arguments, reader, upload: everything before the pair computation),
// the computation (tn93g_run, or one job of a tn93g_run_batch), then certify + output.
struct Case {
  Args A;
  Seqs S;
  long firstLen = 0;
  unsigned long n1 = 0, n2 = 0, N = 0, pairwise = 0;
  bool two = false, do_fst = false, report_self = false;  // two: -s (a second file)
  int prec = 6;
  tn93g_params P{};
  SetGuard sg;
  ResGuard rg;
};

void close_inputs(Args &A)
{
  if (A.input1 && A.input1 != stdin) fclose(A.input1);
  if (A.input2 && A.input2 != stdin) fclose(A.input2);
  A.input1 = A.input2 = nullptr;
}

// Arguments, reader and upload (swifttn93.cpp:2222-2306).  Returns -1 when the case is ready for its computation,
// otherwise the exit code the process ends with.
// TN93G_TIMES=1: the phases of a run (reader, device open, upload, computation, output) on stderr
double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
struct Times {
  double t0 = now_s(), read = 0, open = 0, upload = 0, run = 0, out = 0;
};
Times g_times;
void mark(double &slot)
{
  const double t = now_s();
  slot += t - g_times.t0;
  g_times.t0 = t;
}

int prepare(Case &C, int argc, const char **argv)
{
  Args &A = C.A;
  g_times.t0 = now_s();
  parse_args(A, argc, argv);
  if (A.do_bootstrap || A.do_bootstrap_two_files || A.include_prob < 1.) {
    std::cerr << "swifttn93: -b, -r and -u (random resampling) are not supported" << std::endl;
    return 1;
  }
  if (A.hamming_skip) {
    std::cerr << "swifttn93: -H (Hamming early exit) is not supported" << std::endl;
    return 1;
  }
  // validFlags[toupper(c)] and the subset of resolved codes (refkernel.cpp:56-103)
  signed char valid[256];
  for (int i = 0; i < 256; i++) valid[i] = -1;
  for (int i = 0; ALPHA[i]; i++) valid[(unsigned char)ALPHA[i]] = (signed char)i;
  for (int c = 0; c < 256; c++) g_tab[c] = valid[(unsigned char)toupper(c)];
  uint32_t subset_mask = 0;
  if (A.ambigs_to_resolve)
    for (const char *q = A.ambigs_to_resolve; *q; q++) {
      signed char v = valid[(unsigned char)toupper((unsigned char)*q)];
      if (v > 3) subset_mask |= 1u << v;
    }
  Seqs &S = C.S;
  if (read_fasta(A.input1, S, C.firstLen, A.counts_in_name) == 1) return 1;
  C.n1 = (unsigned long)S.N;
  if (A.input2) {
    if (read_fasta(A.input2, S, C.firstLen, A.counts_in_name) == 1) return 1;
    C.n2 = (unsigned long)S.N - C.n1;
  }
  const bool two = C.two = A.input2 != nullptr;
  close_inputs(A);
  mark(g_times.read);
  C.do_fst = two && A.do_fst;
  C.report_self = !two && A.report_self && !A.do_count;
  C.N = (unsigned long)S.N;
  C.pairwise = (two && !C.do_fst) ? C.n2 * C.n1 : (C.N - 1) * C.N / 2;
  if (!A.quiet) {
    if (!two)
      std::cerr << "Read " << C.N << " sequences of length " << C.firstLen << std::endl
                << "Will perform " << C.pairwise << " pairwise distance calculations";
    else {
      std::cerr << "Read " << C.n1 << " sequences from file 1 and " << C.n2 << " sequences from file 2 of length "
                << C.firstLen << std::endl
                << "Will perform " << C.pairwise << " pairwise distance calculations";
      if (C.do_fst) std::cerr << ". Running in F_ST mode";
    }
  }
  // ---- the GPU
  int rc;
  if (!g_ctx && (rc = tn93g_open(0, &g_ctx))) {
    std::cerr << "tn93g: no usable GPU (error " << rc << ")" << std::endl;
    return 1;
  }
  mark(g_times.open);
  rc = tn93g_set_upload(g_ctx, S.codes.data(), (int64_t)C.N, C.firstLen, S.counts.data(), &C.sg.s);
  mark(g_times.upload);
  if (rc) {
    std::cerr << "tn93g: " << tn93g_error(g_ctx) << std::endl;
    return 1;
  }
  tn93g_params &P = C.P;
  memset(&P, 0, sizeof P);
  P.mode = A.ambig;
  P.kind = (two && !C.do_fst) ? TN93G_CROSS : (C.do_fst ? TN93G_FST : TN93G_TRIANGLE);
  P.n1 = (int64_t)C.n1;
  P.subset_mask = subset_mask;
  P.backend = TN93G_BACKEND_AUTO;
  P.resolve_fraction = A.resolve_fraction;
  P.min_overlap = (long)A.overlap;
  P.threshold = A.distance;
  P.min_distance = A.min_distance;
  P.flags = C.do_fst ? TN93G_FLAG_EXACT : 0u;  // F_ST combines the means non-monotonically: resolve logs exactly
  if (getenv("TN93G_FLAGS")) P.flags |= (uint32_t)strtoul(getenv("TN93G_FLAGS"), nullptr, 0);
  if (getenv("TN93G_TILE_ROWS")) P.tile_rows = (int32_t)atoi(getenv("TN93G_TILE_ROWS"));
  if (const char *b = getenv("TN93G_BACKEND"))  // popc | mfma (diagnostics: the counts are exact either way)
    P.backend = !strcmp(b, "mfma") ? TN93G_BACKEND_MFMA : !strcmp(b, "popc") ? TN93G_BACKEND_POPC : P.backend;
  C.prec = (!A.quiet && A.output == stdout) ? 3 : 6;  // tn93's progress line leaves cerr at precision 3
  return -1;
}

// True when every printed link distance and the summary are the same at both ends of the result's enclosures.
bool certified(const Case &C, const tn93g_result *R)
{
  if (R->exact) return true;
  if (tn93g_result_certify(R, 6) != 1) return false;  // the link distances, printed with %g
  std::ostringstream s1, s2;
  s1 << std::setprecision(C.prec);
  s2 << std::setprecision(C.prec);
  uint64_t comb[200];
  Summary lo{C.pairwise, C.n1, C.n2, C.N, R->found_links, R->skipped, {R->weighted[0], R->weighted[1], R->weighted[2]},
             R->max_lo[0], {R->mean_lo[0], R->mean_lo[1], R->mean_lo[2]}, &R->hist[0][0]};
  Summary hi = lo;
  hi.max0 = R->max_hi[0];
  for (int k = 0; k < 3; k++) hi.mean[k] = R->mean_hi[k];
  write_json(s1, lo, C.two, C.do_fst, comb);
  write_json(s2, hi, C.two, C.do_fst, comb);
  return s1.str() == s2.str();
}

// The case's certified result in C.rg: the one it has when that certifies, else the run again with TN93G_FLAG_EXACT
// (logarithms resolved with the host libm).  Returns 0 or an exit code.
int settle(Case &C)
{
  for (;;) {
    if (!C.rg.r) {
      if (tn93g_run(g_ctx, C.sg.s, &C.P, &C.rg.r)) {
        std::cerr << "tn93g: " << tn93g_error(g_ctx) << std::endl;
        return 1;
      }
    }
    if (certified(C, C.rg.r) || (C.P.flags & TN93G_FLAG_EXACT)) return 0;
    tn93g_result_free(C.rg.r);
    C.rg.r = nullptr;
    C.P.flags |= TN93G_FLAG_EXACT;
  }
}

// The distance file and the summary (swifttn93.cpp:2307-2316, 2406-2434, 2547-2603, 2869-2963).
int output(Case &C)
{
  Args &A = C.A;
  const Seqs &S = C.S;
  const tn93g_result *R = C.rg.r;
  const unsigned long N = C.N, n1 = C.n1, pairwise = C.pairwise;
  const bool two = C.two, do_fst = C.do_fst, report_self = C.report_self;
  const int threads = tn93g_io::io_threads();
  if (getenv("TN93G_STATS"))
    fprintf(stderr,
            "tn93g: gpu %.3f ms (pair kernel %.3f ms) links %lld replayed %lld uncertain %lld host-exact %lld "
            "folded columns %lld\n",
            R->gpu_ms, R->pair_kernel_ms, (long long)R->n_links, (long long)R->n_replayed,
            (long long)R->n_uncertain, (long long)R->n_host_exact, (long long)R->folded_columns);
  // a full buffer for the distance file; stdout gets its own once, before anything is written to it
  static char obuf[1 << 16], sbuf[1 << 16];
  static bool stdout_buffered = false;
  if (A.output != stdout)
    setvbuf(A.output, obuf, _IOFBF, sizeof obuf);
  else if (!stdout_buffered) {
    setvbuf(stdout, sbuf, _IOFBF, sizeof sbuf);
    stdout_buffered = true;
  }
  if (A.format != F_HYPHY && !A.do_count && !A.skip_header) {
    if (A.format == F_CSV)
      fprintf(A.output, "ID1%cID2%cDistance\n", A.delimiter, A.delimiter);
    else
      fprintf(A.output, "Seq1%cSeq2%cDistance\n", A.delimiter, A.delimiter);
  }
  const bool emit = !A.do_count;
  const bool cross = two && !do_fst;
  const long upper = cross ? (long)n1 : (long)N;
  const int64_t nl = R->n_links;
  if (A.format != F_HYPHY && (emit || report_self)) {
    // rows in pieces of 1K-64K links (about two per thread), formatted on all threads and written in row order
    std::vector<long> rb;
    std::vector<int64_t> kb;
    const int64_t piece = std::max<int64_t>(1024, std::min<int64_t>(1 << 16, nl / (2 * (int64_t)threads)));
    tn93g_io::split_rows(R->link_row, nl, upper, piece, rb, kb);
    tn93g_io::emit_pieces(A.output, threads, rb.size() - 1, (size_t)threads * 4, [&](size_t p, std::string &out) {
      char t[64];
      int64_t k = kb[p];
      for (long i = rb[p]; i < rb[p + 1]; i++) {
        if (!cross && report_self) {
          if (A.format == F_CSV) {
            out += S.names[(size_t)i];
            out += A.delimiter;
            out += S.names[(size_t)i];
          } else {
            out += std::to_string(i);
            out += A.delimiter;
            out += std::to_string(i);
          }
          out += A.delimiter;
          out += "0\n";
        }
        for (; k < kb[p + 1] && R->link_row[k] == i; k++) {
          if (!emit) continue;
          const long j = (long)R->link_col[k];
          if (A.format == F_CSV) {
            out += S.names[(size_t)i];
            out += A.delimiter;
            out += S.names[(size_t)j];
          } else {
            out += std::to_string(i);
            out += A.delimiter;
            out += std::to_string(j);
          }
          out += A.delimiter;
          out.append(t, (size_t)fmt_g(t, R->link_dist[k]));
          out += '\n';
        }
      }
    });
  }
  if (emit && A.format == F_HYPHY) {
    // the full N x N matrix, 100 where no link (swifttn93.cpp:2869-2897); rows formatted on all threads
    std::vector<std::vector<std::pair<long, double>>> hylinks(N);
    for (int64_t k = 0; k < nl; k++) {
      const long i = (long)R->link_row[k], j = (long)R->link_col[k];
      hylinks[(size_t)i].push_back(std::make_pair(j, R->link_dist[k]));
      hylinks[(size_t)j].push_back(std::make_pair(i, R->link_dist[k]));
    }
    fputc('{', A.output);
    const size_t rows_per = std::max<size_t>(1, ((size_t)1 << 20) / std::max<size_t>(4 * N, 1));
    const size_t npieces = (N + rows_per - 1) / rows_per;
    tn93g_io::emit_pieces(A.output, threads, npieces, (size_t)threads * 4, [&](size_t p, std::string &o) {
      char t[64];
      for (long r = (long)(p * rows_per); r < (long)std::min<size_t>(N, (p + 1) * rows_per); r++) {
        o += "\n{";
        auto &lk = hylinks[(size_t)r];
        size_t li = 0;
        for (long c = 0; c < (long)N; c++) {
          if (c) o += ',';
          if (c == r)
            o += '0';
          else if (li < lk.size() && lk[li].first == c) {
            o.append(t, (size_t)fmt_g(t, lk[li].second));
            li++;
          } else
            o += "100";
        }
        o += '}';
      }
    });
    fputs("\n}", A.output);
  }
  if (!A.quiet)
    std::cerr << std::endl
              << "Progress:" << std::setw(8) << 100. << "% (" << std::setw(8) << R->found_links << " links found, "
              << std::setw(12) << std::setprecision(3) << (double)pairwise << " evals)";
  std::cerr << std::endl;
  std::ostream &os = A.output == stdout ? std::cerr : std::cout;
  uint64_t comb[200];
  Summary s{pairwise, n1, C.n2, N, R->found_links, R->skipped, {R->weighted[0], R->weighted[1], R->weighted[2]},
            R->max_lo[0], {R->mean_lo[0], R->mean_lo[1], R->mean_lo[2]}, &R->hist[0][0]};
  write_json(os, s, two, do_fst, comb);
  if (A.do_count)
    fprintf(A.output, "Found %ld links among %ld pairwise comparisons\n", (long)R->found_links,
            (long)(pairwise - (unsigned long)R->skipped));
  fflush(A.output);
  if (A.output != stdout) {
    fclose(A.output);
    A.output = stdout;
  }
  return 0;
}

int run(int argc, const char **argv)
{
  Case C;
  int rc = prepare(C, argc, argv);
  if (rc >= 0) return rc;
  g_times.t0 = now_s();
  if ((rc = settle(C))) return rc;
  mark(g_times.run);
  rc = output(C);
  mark(g_times.out);
  if (getenv("TN93G_TIMES"))
    fprintf(stderr, "tn93g: read %.1f ms, device open %.1f ms, upload %.1f ms, run %.1f ms, output %.1f ms\n",
            1e3 * g_times.read, 1e3 * g_times.open, 1e3 * g_times.upload, 1e3 * g_times.run, 1e3 * g_times.out);
  return rc;
}

template <class F>
int guarded(F f)
{
  try {
    return f();
  } catch (const Exit &e) {
    return e.code;
  }
}

int run_guarded(int argc, const char **argv)
{
  return guarded([&] { return run(argc, argv); });
}

// Test drivers: every line of FILE is one tn93 invocation, fields separated by 0x1f: the working directory, then the
// arguments.  Each case runs in this process on one GPU context; its standard output goes to DIR/stdout.bin and its
// exit code to DIR/rc.txt.
//   --batch FILE   the cases one after the other, each with tn93g_run
//   --jobs FILE    the cases in groups of up to 256: all of a group's cases are prepared, their computations run as
//                  one tn93g_run_batch (one graph launch), then each is certified (a case whose printed values differ
//                  at the ends of their enclosures runs again alone with TN93G_FLAG_EXACT) and written out
std::vector<std::vector<std::string>> read_cases(const char *file)
{
  std::ifstream in(file, std::ios::binary);
  std::string line;
  std::vector<std::vector<std::string>> out;
  while (std::getline(in, line)) {
    std::vector<std::string> f;
    size_t a = 0;
    for (size_t b; (b = line.find('\x1f', a)) != std::string::npos; a = b + 1) f.push_back(line.substr(a, b - a));
    f.push_back(line.substr(a));
    out.push_back(f);
  }
  return out;
}

std::vector<const char *> case_argv(const std::vector<std::string> &f)
{
  std::vector<const char *> av{"tn93g"};
  for (size_t k = 1; k < f.size(); k++) av.push_back(f[k].c_str());
  return av;
}

// Runs fn with std::cout captured into DIR/stdout.bin; writes the exit code to DIR/rc.txt.
template <class F>
void captured(const std::string &dir, F fn)
{
  std::ostringstream cap;
  std::streambuf *old = std::cout.rdbuf(cap.rdbuf());
  std::cerr.precision(6);
  const int rc = guarded(fn);
  std::cout.flush();
  std::cout.rdbuf(old);
  std::ofstream(dir + "/stdout.bin", std::ios::binary | std::ios::app) << cap.str();
  if (rc >= 0) std::ofstream(dir + "/rc.txt") << rc << "\n";
}

int batch(const char *file)
{
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return 2;
  for (auto &f : read_cases(file)) {
    if (f.empty() || chdir(f[0].c_str())) return 2;
    std::vector<const char *> av = case_argv(f);
    std::ofstream(f[0] + "/stdout.bin", std::ios::binary | std::ios::trunc);
    // the case's wall time (reader, upload and packing, the run, certification, formatting, the file written) in
    // DIR/wall.txt: the steady-state figure of a warm process
    const auto t0 = std::chrono::steady_clock::now();
    captured(f[0], [&] { return run((int)av.size(), av.data()); });
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::ofstream(f[0] + "/wall.txt") << std::setprecision(9) << wall << "\n";
    if (chdir(cwd)) return 2;
  }
  return 0;
}

int jobs(const char *file)
{
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return 2;
  const auto all = read_cases(file);
  constexpr size_t GROUP = 256;  // cases prepared at once (each holds its output file open)
  for (size_t g0 = 0; g0 < all.size(); g0 += GROUP) {
    const size_t g1 = std::min(all.size(), g0 + GROUP);
    std::vector<std::unique_ptr<Case>> cs(g1 - g0);
    std::vector<std::vector<const char *>> avs(g1 - g0);
    std::vector<size_t> ready;
    for (size_t k = g0; k < g1; k++) {
      const auto &f = all[k];
      if (f.empty() || chdir(f[0].c_str())) return 2;
      cs[k - g0].reset(new Case());
      avs[k - g0] = case_argv(f);
      std::ofstream(f[0] + "/stdout.bin", std::ios::binary | std::ios::trunc);
      Case &C = *cs[k - g0];
      // a case that ends here (bad arguments, reader errors) writes its exit code now; -1 means "prepared"
      captured(f[0], [&] {
        int rc = prepare(C, (int)avs[k - g0].size(), avs[k - g0].data());
        if (rc < 0) ready.push_back(k - g0);
        return rc;
      });
      if (chdir(cwd)) return 2;
    }
    std::vector<const tn93g_set *> sets;
    std::vector<tn93g_params> params;
    for (size_t q : ready) {
      sets.push_back(cs[q]->sg.s);
      params.push_back(cs[q]->P);
    }
    std::vector<tn93g_result *> res(ready.size(), nullptr);
    if (!ready.empty() && tn93g_run_batch(g_ctx, (int)ready.size(), sets.data(), params.data(), res.data())) {
      std::cerr << "tn93g: " << tn93g_error(g_ctx) << std::endl;
      return 1;
    }
    for (size_t r = 0; r < ready.size(); r++) cs[ready[r]]->rg.r = res[r];
    for (size_t q : ready) {
      const auto &f = all[g0 + q];
      if (chdir(f[0].c_str())) return 2;
      Case &C = *cs[q];
      captured(f[0], [&] {
        const int rc = settle(C);
        return rc ? rc : output(C);
      });
      cs[q].reset();
      if (chdir(cwd)) return 2;
    }
  }
  return 0;
}

}  // namespace

int main(int argc, const char **argv)
{
  int rc = (argc == 3 && !strcmp(argv[1], "--batch"))  ? batch(argv[2])
           : (argc == 3 && !strcmp(argv[1], "--jobs")) ? jobs(argv[2])
                                                        : run_guarded(argc, argv);
  if (g_ctx) tn93g_close(g_ctx);
  return rc;
}

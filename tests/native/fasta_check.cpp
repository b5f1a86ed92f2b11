// The multi-threaded FASTA reader of the command line (swifttn93_amd/csrc/tn93g_fasta.h) against the oracle's
// sequential restatement of readFASTA (oracle_fasta_dump in liboracle.so), case by case.
//
//   fasta_check MANIFEST THREADS MIN_CHUNK
// MANIFEST: one case per line, "path1<TAB>path2<TAB>sep" (path2 "-" for none).  Both readers' results are
// rendered in the same text form (exit status, the bytes written to std::cerr, N, L, then name, copy number and
// codes per record) and compared byte for byte.  Prints one line per mismatch and "cases N mismatches M".
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "../../swifttn93_amd/csrc/tn93g_fasta.h"

extern "C" char *oracle_fasta_dump(const char *path1, const char *path2, char sep);

static std::string product_dump(const char *path1, const char *path2, char sep, int threads, size_t min_chunk)
{
  static const char ALPHA[] = "ACGTURYSWKMBDHVN?-";
  signed char valid[256], tab[256];
  for (int i = 0; i < 256; i++) valid[i] = -1;
  for (int i = 0; ALPHA[i]; i++) valid[(unsigned char)ALPHA[i]] = (signed char)i;
  for (int c = 0; c < 256; c++) tab[c] = valid[(unsigned char)toupper(c)];
  tn93g_io::Seqs S;
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
    {
      const tn93g_io::Input buf(F);
      rc = tn93g_io::read_fasta(buf.data(), buf.size(), S, firstLen, sep, tab, threads, min_chunk);
    }
    fclose(F);
  }
  std::cerr.rdbuf(old);
  std::string o = "rc " + std::to_string(rc) + "\nerr " + err.str() + "\n";
  if (rc == 0) {
    o += "N " + std::to_string(S.N) + " L " + std::to_string(firstLen) + "\n";
    for (long k = 0; k < S.N; k++) {
      o += S.names[(size_t)k] + "\t" + std::to_string(S.counts[(size_t)k]) + "\t";
      for (long c = 0; c < firstLen; c++) o += ALPHA[S.codes[(size_t)(k * firstLen + c)]];
      o += "\n";
    }
  }
  return o;
}

int main(int argc, char **argv)
{
  if (argc != 4) {
    fprintf(stderr, "usage: fasta_check MANIFEST THREADS MIN_CHUNK\n");
    return 2;
  }
  const int threads = atoi(argv[2]);
  const size_t min_chunk = (size_t)atol(argv[3]);
  std::ifstream man(argv[1]);
  std::string line;
  long cases = 0, bad = 0;
  while (std::getline(man, line)) {
    const size_t t1 = line.find('\t'), t2 = line.find('\t', t1 + 1);
    if (t1 == std::string::npos || t2 == std::string::npos) return 2;
    const std::string p1 = line.substr(0, t1), p2 = line.substr(t1 + 1, t2 - t1 - 1);
    const char sep = line[t2 + 1];
    const char *q2 = p2 == "-" ? nullptr : p2.c_str();
    char *want = oracle_fasta_dump(p1.c_str(), q2, sep);
    const std::string got = product_dump(p1.c_str(), q2, sep, threads, min_chunk);
    cases++;
    if (got != want) {
      bad++;
      if (bad <= 10) {
        size_t d = 0;
        while (d < got.size() && want[d] && got[d] == want[d]) d++;
        printf("mismatch %s: first difference at byte %zu\n  oracle : %.120s\n  product: %.120s\n", p1.c_str(), d,
               want + (d > 40 ? d - 40 : 0), got.c_str() + (d > 40 ? d - 40 : 0));
      }
    }
    free(want);
  }
  printf("cases %ld mismatches %ld\n", cases, bad);
  return bad ? 1 : 0;
}

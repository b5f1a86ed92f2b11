// tn93g_fasta.h -- multi-threaded host I/O of the tn93g command line: the FASTA reader and the distance-file
// formatter.  Both give exactly the bytes, codes and error messages of tn93's sequential code (readFASTA,
// swifttn93.cpp:292-406 with the length rules of 259-273; the output loop of swifttn93.cpp:2406-2434 and 2547-2603),
// on any number of threads.
//
// Reader.  readFASTA is a three-state machine: state 0 skips to the first '>' or '#'; state 1 takes the name up to
// '\n' or '\r' (a '>' or '#' inside a name is part of it); state 2 maps bytes through validFlags[toupper(c)], drops
// unknown bytes and starts the next record at the first '>' or '#'.  So the record starts are exactly the '>'/'#'
// bytes that do not lie inside a name line, and a record's sequence is every valid byte between its name line and
// the next record start.  The reader (1) finds all '>'/'#' positions in parallel chunks, (2) walks them in order
// to pick the record starts and parse names and copy numbers (short work per record), (3) decodes every record's
// sequence in parallel straight into its row of the code matrix, counting its valid bytes, and (4) applies the
// length rules in record order, so the first error reported is the one the sequential reader reports.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <utility>
#include <vector>

namespace tn93g_io {

// Worker threads for host I/O: OMP_NUM_THREADS when set (the knob the reference's OpenMP build reads), else the
// hardware concurrency; at least 1, at most 64.
inline int io_threads()
{
  long t = 0;
  if (const char *e = getenv("OMP_NUM_THREADS")) t = strtol(e, nullptr, 10);
  if (t <= 0) t = (long)std::thread::hardware_concurrency();
  return (int)std::max<long>(1, std::min<long>(t, 64));
}

// Runs fn(task) for task in [0, ntasks) on up to `threads` threads (the calling thread included); tasks are taken
// in increasing order from a shared counter.
inline void parallel_for(int threads, int64_t ntasks, const std::function<void(int64_t)> &fn)
{
  if (ntasks <= 0) return;
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(threads, ntasks));
  if (nt == 1) {
    for (int64_t t = 0; t < ntasks; t++) fn(t);
    return;
  }
  std::atomic<int64_t> next{0};
  auto work = [&]() {
    for (int64_t t; (t = next.fetch_add(1)) < ntasks;) fn(t);
  };
  std::vector<std::thread> pool;
  pool.reserve((size_t)nt - 1);
  for (int k = 1; k < nt; k++) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

// vector storage whose new elements are left uninitialised (the code matrix is written once, by the decoder)
template <class T>
struct UninitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = UninitAlloc<U>;
  };
  UninitAlloc() = default;
  template <class U>
  UninitAlloc(const UninitAlloc<U> &) noexcept
  {
  }
  template <class U>
  void construct(U *p) noexcept
  {
    ::new ((void *)p) U;
  }
  template <class U, class... Args>
  void construct(U *p, Args &&...args)
  {
    ::new ((void *)p) U(std::forward<Args>(args)...);
  }
};

struct Seqs {
  std::vector<uint8_t, UninitAlloc<uint8_t>> codes;  // N x L codes 0..17
  std::vector<std::string> names;
  std::vector<int64_t> counts;
  long N = 0;
};

// The whole input of a stream: a read-only mapping of a regular file read from its start, else (pipes, stdin, a
// stream already read from) everything fread returns.
class Input {
 public:
  explicit Input(FILE *F)
  {
    struct stat st;
    if (fstat(fileno(F), &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0 && ftell(F) == 0) {
      void *m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fileno(F), 0);
      if (m != MAP_FAILED) {
        map_ = (unsigned char *)m;
        n_ = (size_t)st.st_size;
        madvise(m, n_, MADV_SEQUENTIAL);
        return;
      }
    }
    std::vector<unsigned char> tmp(1 << 20);
    size_t r;
    while ((r = fread(tmp.data(), 1, tmp.size(), F)) > 0) buf_.insert(buf_.end(), tmp.data(), tmp.data() + r);
    n_ = buf_.size();
  }
  ~Input()
  {
    if (map_) munmap(map_, n_);
  }
  Input(const Input &) = delete;
  Input &operator=(const Input &) = delete;
  const unsigned char *data() const { return map_ ? map_ : buf_.data(); }
  size_t size() const { return n_; }

 private:
  unsigned char *map_ = nullptr;
  size_t n_ = 0;
  std::vector<unsigned char> buf_;
};

// copy number of a record name (swifttn93.cpp:324-333): atoi of what follows the last separator, looked for from
// the third-last character back to the second; 1 when there is none or the value is below 1 as unsigned
inline int64_t name_count(const std::string &name, char sep)
{
  const long len = (long)name.size();
  unsigned long count = 1UL;
  if (len >= 3) {
    long k;
    for (k = 2; k < len; k++)
      if (name[(size_t)(len - k)] == sep) break;
    if (k < len) count = (unsigned long)(long)atoi(name.c_str() + (len - k + 1));
    if (count < 1) count = 1;
  }
  return (int64_t)(long)count;
}

// Appends the records of buf[0, n) to S.  tab = validFlags[toupper(c)] (-1: dropped byte).  firstLen carries the
// length of the first record over calls (a second file must match the first).  Returns 0, or 1 after printing
// tn93's message to std::cerr.  min_chunk: bytes per parallel scan task (tests set it small to force many tasks).
inline int read_fasta(const unsigned char *buf, size_t n, Seqs &S, long &firstLen, char sep,
                      const signed char *tab, int threads, size_t min_chunk = (size_t)1 << 20)
{
  // (1) every '>' and '#', in order
  const size_t nchunk = std::max<size_t>(1, std::min<size_t>((size_t)threads * 4, (n + min_chunk - 1) / min_chunk));
  std::vector<std::vector<size_t>> part(nchunk);
  parallel_for(threads, (int64_t)nchunk, [&](int64_t c) {
    const size_t a = n * (size_t)c / nchunk, b = n * (size_t)(c + 1) / nchunk;
    std::vector<size_t> &v = part[(size_t)c];
    size_t p = a;
    // 8 bytes at a time: a zero byte of x ^ ('>' x 8) or x ^ ('#' x 8) marks a candidate (exact test per word:
    // (y - 0x01..) & ~y & 0x80.. is non-zero iff y has a zero byte)
    constexpr uint64_t ONES = 0x0101010101010101ull, HIGH = 0x8080808080808080ull;
    for (; p + 8 <= b; p += 8) {
      uint64_t x;
      memcpy(&x, buf + p, 8);
      const uint64_t y1 = x ^ (ONES * (uint8_t)'>'), y2 = x ^ (ONES * (uint8_t)'#');
      if (((y1 - ONES) & ~y1 & HIGH) | ((y2 - ONES) & ~y2 & HIGH))
        for (size_t q = p; q < p + 8; q++)
          if (buf[q] == '>' || buf[q] == '#') v.push_back(q);
    }
    for (; p < b; p++)
      if (buf[p] == '>' || buf[p] == '#') v.push_back(p);
  });
  std::vector<size_t> cand;
  {
    size_t tot = 0;
    for (auto &v : part) tot += v.size();
    cand.reserve(tot);
    for (auto &v : part) cand.insert(cand.end(), v.begin(), v.end());
  }
  // (2) record starts, names and copy numbers in file order; stops at the first name error
  struct Rec {
    size_t seq_begin, seq_end;
  };
  std::vector<Rec> recs;
  std::vector<std::string> names;
  std::vector<int64_t> counts;
  enum { OK, EMPTY_NAME, EOF_STATE0, EOF_STATE1 } stop = OK;
  if (cand.empty()) stop = EOF_STATE0;
  size_t ci = 0;
  while (stop == OK && ci < cand.size()) {
    const size_t r = cand[ci];
    size_t q = r + 1;
    while (q < n && buf[q] != '\n' && buf[q] != '\r') q++;
    if (q == n) {  // the file ends inside a name line
      stop = EOF_STATE1;
      break;
    }
    if (q == r + 1) {
      stop = EMPTY_NAME;
      break;
    }
    names.emplace_back((const char *)buf + r + 1, q - r - 1);
    counts.push_back(name_count(names.back(), sep));
    const size_t sb = q + 1;
    ci = (size_t)(std::lower_bound(cand.begin() + (long)ci, cand.end(), sb) - cand.begin());
    recs.push_back(Rec{sb, ci < cand.size() ? cand[ci] : n});
  }
  // (3) decode: every record into its row, valid bytes counted (a row holds at most L codes; a record with more or
  // fewer fails the length rule below before its row is used)
  const size_t nrec = recs.size();
  std::vector<long> lens(nrec, 0);
  const size_t base = S.codes.size();
  long L = firstLen;
  size_t r0 = 0;  // records already placed
  if (S.N == 0 && nrec > 0) {  // the first record of the first file sets L
    long c = 0;
    for (size_t p = recs[0].seq_begin; p < recs[0].seq_end; p++) c += tab[buf[p]] >= 0;
    lens[0] = c;
    L = c;
    if (L > 0) {
      S.codes.resize(base + (size_t)L);
      uint8_t *o = S.codes.data() + base;
      for (size_t p = recs[0].seq_begin; p < recs[0].seq_end; p++)
        if (tab[buf[p]] >= 0) *o++ = (uint8_t)tab[buf[p]];
    }
    r0 = 1;
  }
  if (L > 0 && nrec > r0) {
    S.codes.resize(base + (size_t)L * nrec);
    uint8_t *out = S.codes.data() + base;
    // tasks of about equal bytes: whole records
    const size_t per = std::max<size_t>(min_chunk, (n / std::max<size_t>((size_t)threads * 8, 1)) + 1);
    std::vector<size_t> tb{r0};
    for (size_t k = r0, acc = 0; k < nrec; k++) {
      acc += recs[k].seq_end - recs[k].seq_begin;
      if (acc >= per) {
        tb.push_back(k + 1);
        acc = 0;
      }
    }
    if (tb.back() != nrec) tb.push_back(nrec);
    parallel_for(threads, (int64_t)tb.size() - 1, [&](int64_t t) {
      // branch-free: every byte is stored at the row position c (into a private row with one spare slot, so a
      // dropped byte's store is overwritten by the next valid one), c advances on valid bytes only
      std::vector<uint8_t> row((size_t)L + 1);
      for (size_t k = tb[(size_t)t]; k < tb[(size_t)t + 1]; k++) {
        long c = 0;
        for (size_t p = recs[k].seq_begin; p < recs[k].seq_end; p++) {
          const signed char v = tab[buf[p]];
          row[(size_t)std::min(c, L)] = (uint8_t)v;
          c += v >= 0;
        }
        lens[k] = c;
        if (c == L) memcpy(out + (size_t)L * k, row.data(), (size_t)L);
      }
    });
  }
  // (4) the length rules in record order (add_sequence, swifttn93.cpp:259-273)
  for (size_t k = 0; k < nrec; k++) {
    if (S.N == 0) {
      firstLen = lens[k];
      if (firstLen <= 0) {
        S.codes.resize(base);
        std::cerr << "First sequence length must be positive." << std::endl;
        return 1;
      }
    } else if (lens[k] != firstLen) {
      S.codes.resize(base);
      std::cerr << "All sequences must have the same length (" << firstLen << "), but sequence '" << names[k]
                << "' had length " << lens[k];
      return 1;
    }
    S.N++;
    S.names.push_back(std::move(names[k]));
    S.counts.push_back(counts[k]);
  }
  if (stop == EMPTY_NAME) {
    std::cerr << "Sequence names must be non-empty." << std::endl;
    return 1;
  }
  if (stop != OK) {
    std::cerr << "Unexpected end of file: state " << (stop == EOF_STATE0 ? 0 : 1) << std::endl;
    return 1;
  }
  return 0;
}

// ----------------------------------------------------------------------------------------------- distance file

// Pieces of the row range [0, upper) with about `target` links each: piece p holds rows [rb[p], rb[p + 1]) and
// links [kb[p], kb[p + 1]) (links are in row-major order, row = link_row[k]).
inline void split_rows(const int64_t *link_row, int64_t nl, long upper, int64_t target, std::vector<long> &rb,
                       std::vector<int64_t> &kb)
{
  rb.assign(1, 0);
  kb.assign(1, 0);
  int64_t k = 0;
  while (rb.back() < upper) {
    const long a = rb.back();
    long b;
    if (nl - k <= target) {
      b = upper;
    } else {
      b = (long)link_row[k + target];
      if (b <= a) b = a + 1;
    }
    int64_t k2 = k;
    while (k2 < nl && link_row[k2] < b) k2++;
    rb.push_back(b);
    kb.push_back(k2);
    k = k2;
  }
}

// Formats pieces [p0, p1) into outs[0..p1-p0) on `threads` threads with fmt(piece, string&), then writes them to
// F in order; repeats in rounds of `per_round` pieces so that only a round's text is held in memory.
inline void emit_pieces(FILE *F, int threads, size_t npieces, size_t per_round,
                        const std::function<void(size_t, std::string &)> &fmt)
{
  std::vector<std::string> outs;
  for (size_t p0 = 0; p0 < npieces; p0 += per_round) {
    const size_t p1 = std::min(npieces, p0 + per_round);
    outs.assign(p1 - p0, std::string());
    parallel_for(threads, (int64_t)(p1 - p0), [&](int64_t t) { fmt(p0 + (size_t)t, outs[(size_t)t]); });
    for (auto &o : outs) fwrite(o.data(), 1, o.size(), F);
  }
}

}  // namespace tn93g_io

// tn93g.hip -- MI355X (gfx950) all-pairs TN93 engine behind the C ABI of include/tn93g.h.
//
// Data layout in HBM (one set of n sequences, L sites, grown by appends up to its capacity; Ls = L rounded up to
// 16, npad = capacity rounded up to 64, plus 64; W = 32-site words):
//   codes   [n][Ls]    u8    ValidChars codes, row-major, padded with gaps        (replay kernel, j-side lookups)
//   codesT  [Ls][npad] u8    site-major copy                                     (row-side event codes, coalesced)
//   planes  [n][W]     uint4 per 32 sites: clean, C|T bit, G|T bit, T bit        (clean/clean counts, VALU popcount)
//   nmask   [n][W]     u32   N sites, for sets whose N occur in long runs         (pair kernel, dense-N sets only)
//   onehot  [npad/4][W64][4][16][16] u8 one-hot A,C,G,T per site in the operand order of v_mfma_i32_16x16x64_i8
//                                                                                (clean/clean counts, MFMA backend)
//   ev_off/ev   [n+1] i32, u32 (site << 5 | code): ambiguity events in site order, gaps excluded
//   evg_off/evg the same with gap sites (-a gapmm);   nrun_off/nruns  N runs [first, end) of dense-N sets;
//   grun_off/gruns gap runs;   pre [n][W+1] 32-byte records: uint4 plane counts before each word boundary and
//   the planes of that word (one cache line per prefix count), tb [n] the totals
// Per run: either the fast-path row accumulators and unordered links (distances certified in the kernel), or D
// [rows][cols] f64 + U [rows][cols] u16 (ulps that bound an uncertain logarithm) for the exact path; the histogram
// in per-stripe replicas, the flat replay list and the host-exact list; for long alignments the 16-cell count
// tables of the sequences with B/D/H/V codes against the other side of the run.
//
// Kernels of a run (on the context's stream, or one of its side streams in a captured batch):
//   k_spec16  16-cell count tables of the special sequences (long alignments with B/D/H/V codes)
//   k_pairs   clean/clean counts (popcount, popcount+nmask, or MFMA) + event tables from LDS + -g fallback pass;
//             dyadic pairs complete here, order-dependent ones are appended to the flat replay list
//   k_tail    FP64 TN93/K2P tail, certification, histogram, row accumulators and links (fast path) or D/U
//   k_replay  site-order FP64 replay for pairs with non-dyadic contributions (runs of N as repeated additions)
//   k_scan    per-row link offsets;  k_link_gather  links into one array;  k_link_order  links in tn93's row and
//             column order;  k_hist_sum  histogram replicas summed
//   k_patch   host-exact distances into D;  k_rows  per-row sums in column order (exact path only)
// A captured batch runs k_pairs, k_tail, k_scan, k_link_gather, k_link_order and k_hist_sum of the jobs that share
// a pair-kernel variant as one launch each (the *_multi kernels; each block finds its job from a table of offsets),
// and clears and reads back every job's buffers with one k_fill_ranges and one k_copy_ranges launch.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <iostream>
#include <map>
#include <mutex>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/tn93g.h"
#include "tn93g_math.h"

#define TN93G_VERSION "tn93g 0.1 (gfx950)"

namespace {

constexpr int NCODE = 18;
constexpr int GAPC = 17;
constexpr int NC = 15;  // N
constexpr int NCH = 15;
// resolution sets (refkernel.cpp:19-35): bit b set iff base b (A, C, G, T) resolves the code
constexpr unsigned RES[NCODE] = {1, 2, 4, 8, 8, 5, 10, 6, 9, 12, 3, 14, 13, 11, 7, 15, 15, 0};

// ------------------------------------------------------------------------------------------ host tables
// One site's contribution of tn93's ambiguityHandler (refkernel.cpp:267-380): every call adds ONE value to a set
// of float_counts cells, or resolves to one integer diagonal cell, and may count one ambiguity.
struct SiteEffect {
  double v = 0.0;   // value added to each cell of mask
  uint32_t mask = 0;  // float_counts cells (bit 4a+b)
  int icd = -1;       // integer_counts[icd][icd]++ (resolved ambiguity)
  int amb = 0;
};

static double rcnt(int c)  // resolutionsCount (refkernel.cpp:38-51): 1/|R| as 1./k.f
{
  int k = __builtin_popcount(RES[c]);
  return k ? 1. / (double)(float)k : 0.;
}

static SiteEffect handler_effect(int c1, int c2, int mode, uint32_t rta)
{
  SiteEffect E;
  if (mode == TN93G_SKIP) return E;
  const bool res_all = mode == TN93G_RESOLVE;
  if (c1 < 4) {
    if (rcnt(c2) > 0.) {
      if (res_all || (mode == TN93G_SUBSET && (rta >> c2 & 1))) {
        if (RES[c2] >> c1 & 1) {
          E.amb = 1;
          E.icd = c1;
          return E;
        }
      }
      E.v = rcnt(c2);
      for (int b = 0; b < 4; b++)
        if (RES[c2] >> b & 1) E.mask |= 1u << (4 * c1 + b);
    }
    return E;
  }
  if (c2 < 4) {
    if (rcnt(c1) > 0.) {
      if (res_all || (mode == TN93G_SUBSET && (rta >> c1 & 1))) {
        if (RES[c1] >> c2 & 1) {
          E.amb = 1;
          E.icd = c2;
          return E;
        }
      }
      E.v = rcnt(c1);
      for (int a = 0; a < 4; a++)
        if (RES[c1] >> a & 1) E.mask |= 1u << (4 * a + c2);
    }
    return E;
  }
  double norm = rcnt(c1) * rcnt(c2);
  if (!(norm > 0.0)) return E;
  if (res_all || (mode == TN93G_SUBSET && (rta >> c1 & 1) && (rta >> c2 & 1))) {
    E.amb = 1;
    unsigned m = RES[c1] & RES[c2];
    if (m) {
      E.v = 1. / (long)__builtin_popcount(m);
      for (int i = 0; i < 4; i++)
        if (m >> i & 1) E.mask |= 1u << (5 * i);
      return E;
    }
  }
  E.v = norm;
  for (int i = 0; i < 4; i++)
    if (RES[c1] >> i & 1)
      for (int j = 0; j < 4; j++)
        if (RES[c2] >> j & 1) E.mask |= 1u << (4 * i + j);
  return E;
}

// Effect of a site with codes (c1, c2) in the scan of computeTN93: gaps skip the site except under GAPMM, where a
// gap facing a non-gap is handled as N (refkernel.cpp:620-665).  Both clean: not an event (counted directly).
static SiteEffect site_effect(int c1, int c2, int mode, uint32_t rta)
{
  bool g1 = c1 >= GAPC, g2 = c2 >= GAPC;
  if (g1 || g2) {
    if (mode != TN93G_GAPMM || (g1 && g2)) return SiteEffect();
    return handler_effect(g1 ? NCH : c1, g2 ? NCH : c2, mode, rta);
  }
  if (c1 < 4 && c2 < 4) return SiteEffect();
  return handler_effect(c1, c2, mode, rta);
}

// Replay-kernel table entry: the effect in the form the FP64 site-order loop applies.
struct ReplayEnt {
  double v;
  uint32_t mask;
  int16_t icd;  // -1: none
  uint16_t amb;
};

// Event-table entry of the pair kernel: the effect scaled by 16 and aggregated into what the formula needs, packed
// as fields: x = nf0 | nf1 << 16, y = nf2 | nf3 << 16, z = diag | AG << 16 (16 bits each), w = CT | amb << 14 |
// nondyadic << 23 (14, 9 and 9 bits).  Sums of up to EVFLUSH = 256 entries fit every field (at most 32 per entry in
// the 16- and 14-bit fields, at most 1 in the 9-bit ones), so entries are added as plain 32-bit words.
// Exact whenever every value is a multiple of 1/16 (then all of tn93's float sums are exact and order-free).  An
// entry with other values (its pairs are replayed) carries only its site's total weight in the nf fields, 32 (2 x 16)
// for a counted site: on the field of the clean base when one code (a or b) is clean, where the marginal form
// subtracts that base, else on nf0; so the nf fields of any pair still sum to 32 x its counted sites.  Returns false
// when a weight is neither 0 nor 1 (then no pair's count is certified from the fields).
static bool event_entry(const SiteEffect &E, uint32_t out[4], int a = -1, int b = -1)
{
  double cell[16];
  bool dyadic = true;
  for (int k = 0; k < 16; k++) {
    double s = (E.mask >> k & 1) ? E.v * 16.0 : 0.0;
    if (s != std::floor(s)) dyadic = false;
    cell[k] = s;
  }
  if (E.icd >= 0) cell[5 * E.icd] += 16.0;
  uint32_t nf[4] = {0, 0, 0, 0}, dg = 0, ag = 0, ct = 0;
  if (dyadic) {
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        uint32_t c = (uint32_t)cell[4 * a + b];
        nf[a] += c;
        nf[b] += c;
        if (a == b) dg += c;
      }
    ag = (uint32_t)(cell[2] + cell[8]);
    ct = (uint32_t)(cell[7] + cell[13]);
  }
  bool unit = true;
  if (!dyadic) {
    // the site's weight: each cell of the mask gets v = 1/k of its resolutions, so k cells add up to 1 (to within
    // the rounding of v)
    const double w = (double)__builtin_popcount(E.mask) * E.v + (E.icd >= 0 ? 1.0 : 0.0);
    if (std::fabs(w - 1.0) < 1e-9) {
      const bool ca = a >= 0 && a < 4, cb = b >= 0 && b < 4;
      nf[ca != cb ? (ca ? a : b) : 0] = 32;
    } else {
      unit = w == 0.0;
    }
  }
  out[0] = nf[0] | nf[1] << 16;
  out[1] = nf[2] | nf[3] << 16;
  out[2] = dg | ag << 16;
  out[3] = ct | (uint32_t)E.amb << 14 | (dyadic ? 0u : 1u << 23);
  return unit;
}

// ------------------------------------------------------------------------------------------ device structures
// Plane counts of a sequence before word boundary w (clean, C|T, G|T, T over sites [0, 32 w)) beside the planes of
// word w itself, so that the counts before any site are one 32-byte record (one cache line).
struct PrefixRec {
  uint4 pre, word;
};

struct SetDev {
  int64_t n, L, Ls, npad, W, W64;
  uint8_t *codes, *codesT;
  uint4 *planes;
  uint8_t *onehot;
  int32_t *ev_off;
  uint32_t *ev;
  int32_t *evg_off;  // the lists with gap sites (GAPMM)
  uint32_t *evg;
  uint32_t *nmask;   // [n][W] N sites
  int32_t *nrun_off; // [n+1] N runs, for sets with dense N (ndense)
  uint2 *nruns;
  PrefixRec *pre;    // [n][W+1] plane counts before each word boundary, and the word (zero at W)
  uint4 *tb;         // [n] pre[s][W], the totals of sequence s
  uint32_t *npre;    // [n][W+1] N sites before each word boundary
  int32_t *grun_off; // [n+1] gap runs
  uint2 *gruns;
  int ndense;        // N sites are not in ev: handled through nmask (pair kernel) and nruns (replay)
  int64_t *weights;
  int64_t total_events;
  int64_t total_gruns;  // gap runs of all sequences
  // constant columns left out of every device array (the same clean base in every sequence): their number and
  // their count per base, added to every pair's clean/clean diagonal
  int64_t kt;
  uint32_t kc[4];
};

struct ExactRec {
  int64_t i, j;
  tn93g_tailin tin;
};
struct PatchRec {
  int64_t i, j;
  double thisD;
  int32_t idx;
  int32_t link;
};

// Per-row reductions of the fast path (no distance matrix): tn93 sums thisD*w of a row in column order, but only
// prints the final mean to 6 digits, so the sum is taken in any order together with sum |thisD*w| and turned into
// an enclosure of tn93's sequentially rounded value on the host (|fl(sum) - sum| <= gamma_n sum |t|).
struct RowAcc {
  double slo, shi, sabs;              // sum of thisD*w at the low / high end of the enclosure, sum of |thisD*w|
  unsigned long long mlo, mhi;        // bit patterns of max thisD (non-negative doubles order like their bits)
  long long wsum, skip, linkw;        // sum of w over counted pairs, skipped pairs, link weights
};
struct RowPart {
  double slo, shi, sabs, mlo, mhi;
  long long wsum, skip, linkw;
};
struct LinkRec {
  int64_t i, j;
  double d, lo, hi;
};

struct ChunkHdr;
struct RunArgs;
struct RunArgs {
  // set
  const uint4 *planes;
  const uint8_t *onehot;
  const uint8_t *codes, *codesT;
  const int32_t *ev_off;
  const uint32_t *ev;
  const uint32_t *nmask;   // [n][W] N sites (dense-N sets)
  const int32_t *nrun_off; // [n+1] and the N runs [start, end) of each sequence (dense-N sets)
  const uint2 *nruns;
  int nplane;              // N sites are counted from nmask / nruns, not in the event lists
  const PrefixRec *pre;    // [n][W+1] plane counts (clean, C|T, G|T, T) before each word boundary, and that word
  const uint32_t *npre;    // [n][W+1] the same for the N sites (nmask)
  const uint4 *tb;         // [n] the same over all words (the totals)
  const int32_t *grun_off; // [n+1] and the gap runs [start, end) of each sequence
  const uint2 *gruns;
  const int64_t *weights;
  int unitw;                 // every copy number of the set is 1: pair weights are 1 without loading them
  int64_t n, Ls, npad, W, W64;
  int64_t kt;                // constant columns folded out of the set (none of them is in W), kc[b] of base b:
  uint32_t kc[4];            // each adds one A/A, C/C, G/G or T/T site to every pair
  // pair space: rows [r0, r1) x cols [c0, c1), triangle => only j > i
  int64_t r0, r1, c0, c1, ncols, n1;
  int triangle, fst, gapmm, skip_events, avg_fallback, exact, all_slow;
  int gcert;                   // the event tables' nf fields count the sites of order-dependent pairs exactly
  double gE;                   // with gcert: twice the bound on |tng_fp - tng| below, (16 L^2 + 64 L) 2^-52
  double rf, thr, mind, window;
  int64_t min_overlap;
  const uint4 *evtab;          // [4][18*18] packed event entries: main mode, AVERAGE (the -g fallback), and the same
                               // two with the clean base facing an event subtracted (the marginal form of k_pairs)
  const ReplayEnt *rtab;       // [2][18*18]: main mode, AVERAGE
  // outputs
  double *D;
  uint16_t *U;
  unsigned long long *hist;    // [NSTRIPE][3*200]: histogram replicas, block b adds into replica b % NSTRIPE
  int32_t *row_links;          // [nrows]
  unsigned long long *slow_list;  // [slow_cap] pairs for the site-order replay: i << 32 | j (bit 63: AVERAGE first);
                                  // each pair tile appends its pairs as one contiguous range
  unsigned long long *slow_total; // entries in slow_list (64-bit counter in stripe 0)
  // Clean/clean 4x4 count tables of the "special" sequences (those with a B, D, H or V code; every order-dependent
  // pair of the fast path has one) against the other side of the run, for the replay.  Null when not built.
  const int32_t *spec_row;  // [nrows]: rank of row r0 + r among the special rows, or -1
  const int32_t *spec_col;  // [ncols]: rank of column c0 + c among the special columns, or -1
  const uint32_t *tab_row;  // [nsr][4][ncols][4]: [s][a][c][b] counts special row s's base a against column c's base b
  const uint32_t *tab_col;  // [nsc][4][nrows][4]: [s][a][r][b] counts special column s's base a against row r's base b
  const int64_t *spec_rows; // [nsr] the special rows (set indices), then [nsc] the special columns
  int nsr, nsc;
  ExactRec *exact_list;
  unsigned int *exact_cnt;
  unsigned int exact_cap;
  unsigned int *stripes;       // [NSTRIPE][32]: counters spread over cache lines (ST_* fields), summed on the host
  int fast;                    // 1: row reductions in RowAcc, links unordered in LinkRec; 0: D/U + k_rows
  RowAcc *racc;                // [nrows]
  LinkRec *links;              // [NSTRIPE][link_cap]: region s holds the links counted in stripe s; after the
                               // ordering, the run's links in tn93's order
  LinkRec *links_flat;         // the regions concatenated (k_link_gather)
  int64_t *link_off;           // [nrows + 1] exclusive scan of the per-row link counts
  uint32_t *lslot;             // [links] index into links_flat of each link, grouped by row
  unsigned int link_cap;       // per region
  // fast path: the pair kernel hands each wave-row (row i, 64 columns of a tile) to k_tail in a fixed slot
  // (i - r0) * gx + column tile: `need` lanes carry the tail inputs, `skip` lanes had too small an overlap
  struct ChunkHdr *chdr;
  uint32_t *cdata;             // [slot][7][64]: nf0..nf3, diag, AG, CT (x16) per lane; with c16, [slot][4][64]:
                               // nf0 | nf1 << 16, nf2 | nf3 << 16, diag | AG << 16, CT
  int64_t cstride;             // words from one slot to the next: 4 x 64 with c16, else 7 x 64
  int c16;                     // 32 * Ls < 2^16: every x16 count fits 16 bits
  int64_t gx;                  // column tiles
  // clean/clean counts from int8 GEMMs (the GEMM-count variant of the pair kernel): four [nrows][ncols] products
  // gcs apart, element (i - r0) ncols + (j - c0): [A|G].[A|G], [C|T].[C|T], R.R and Y.Y over the sites
  const uint32_t *gcnt;
  int64_t gcs;
  int ti;                      // rows of the pair kernel's tiles (8, 16 or 32)
};

struct ChunkHdr {
  unsigned long long need, skip;
};

constexpr int NSTRIPE = 64;  // one per lane of a wave (k_link_gather sums them across one)
enum { ST_UNCERTAIN = 0, ST_UNDECIDED = 1, ST_LINKS = 2, ST_EXACT = 3, ST_SLOW = 4 /* u64, stripe 0 only */ };

__device__ inline int fst_bin(const RunArgs &A, int64_t i, int64_t j)
{
  if (!A.fst) return 0;
  if (i < A.n1) return j >= A.n1 ? 2 : 0;
  return 1;
}

__device__ inline RowPart rowpart_zero() { return RowPart{0., 0., 0., -1., -1., 0, 0, 0}; }

__device__ inline unsigned stripe_of_block() { return (blockIdx.x + blockIdx.y * 7u) % NSTRIPE; }

// links go to the region of the block's stripe (one counter per region, so no single hot atomic)
__device__ inline void append_link(const RunArgs &A, int64_t i, int64_t j, double x, double lo, double hi)
{
  const unsigned st = stripe_of_block();
  const unsigned k = atomicAdd(&A.stripes[st * 32 + ST_LINKS], 1u);
  if (k < A.link_cap) A.links[(size_t)st * A.link_cap + k] = LinkRec{i, j, x, lo, hi};
  atomicAdd(&A.row_links[i - A.r0], 1);  // per-row counts, for the ordering of the links
}

// block-local counters flushed once per block into its stripe
struct BlockStats {
  unsigned unc, undecided;
};
__device__ inline void flush_block(const RunArgs &A, BlockStats *bs, unsigned long long *lhist, int tid, int nthr)
{
  __syncthreads();
  const unsigned st = stripe_of_block();
  if (tid == 0) {
    if (bs->unc) atomicAdd(&A.stripes[st * 32 + ST_UNCERTAIN], bs->unc);
    if (bs->undecided) atomicAdd(&A.stripes[st * 32 + ST_UNDECIDED], bs->undecided);
  }
  if (lhist)
    for (int k = tid; k < 600; k += nthr)
      if (lhist[k]) atomicAdd(&A.hist[st * 600 + k], lhist[k]);
}

// Everything after the counts of one pair: tests, tail, histogram, link (refkernel.cpp:732-815 and the phase C of
// swifttn93.cpp:2547-2603).  Fast path: the pair's share of its row's reductions in rp; otherwise D/U for k_rows.
// FASTLOG (fast path only): the tail with the enclosed fast logarithm, every output certified from [dlo, dhi];
// otherwise the double-double logarithm, exact unless a log lies near a rounding midpoint.
// Adds w to bin idx of the block's histogram.  With unit weights, the lanes of the wave that fall into the bin of the
// first active lane add their count with one LDS atomic: distances of closely related sequences all fall into the
// first bins, and 64 atomics on one LDS address serialise.
__device__ inline void hist_add(unsigned long long *lhist, int idx, int64_t w, bool unit)
{
  if (unit) {
    const int lead = __builtin_amdgcn_readfirstlane(idx);
    const unsigned long long m = __ballot(idx == lead);
    if (idx == lead) {
      if ((int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(&lhist[lead], (unsigned long long)__popcll(m));
      return;
    }
  }
  atomicAdd(&lhist[idx], (unsigned long long)w);
}

// ltab: the logarithm table, or a copy of it in the block's LDS.
template <bool FASTLOG>
__device__ void finish_pair(const RunArgs &A, int64_t i, int64_t j, const tn93g_tailin &tin,
                            unsigned long long *lhist, BlockStats *bs, RowPart &rp,
                            const double (*ltab)[4] = tn93g_logtab)
{
  tn93g_tailout t = FASTLOG ? tn93g_tail_fast(tin, ltab) : tn93g_tail(tin, A.window);
  if (A.exact && t.unc) {
    unsigned k = atomicAdd(A.exact_cnt, 1u);
    if (k < A.exact_cap) {
      A.exact_list[k].i = i;
      A.exact_list[k].j = j;
      A.exact_list[k].tin = tin;
    }
    return;
  }
  const int64_t w = A.unitw ? 1 : A.weights[i] * A.weights[j];
  const int bin = fst_bin(A, i, j);
  bool decided = true;
  // histogram bin: floor(200 d) must agree over the enclosure (or clamp to 199 on the same side)
  const double flo = floor(t.dlo * 200.0), fhi = floor(t.dhi * 200.0);
  if (flo != fhi && !(flo >= 200.0 && fhi >= 200.0) && !(flo < 0.0 && fhi < 0.0)) decided = false;
  hist_add(lhist, bin * 200 + tn93g_hist_index(t.d), w, A.unitw);
  const double x = t.d <= 0. ? 0. : t.d;
  const double xlo = t.dlo <= 0. ? 0. : t.dlo, xhi = t.dhi <= 0. ? 0. : t.dhi;
  const bool cert_in = xlo >= A.mind && xhi <= A.thr;
  const bool cert_out = xhi < A.mind || xlo > A.thr;
  if (!cert_in && !cert_out) decided = false;
  const bool link = x >= A.mind && x <= A.thr;
  if (!(x == x)) decided = false;
  if (A.fast) {
    if (xlo != xhi) atomicAdd(&bs->unc, 1u);
    if (!decided) atomicAdd(&bs->undecided, 1u);
    if (link) append_link(A, i, j, x, xlo, xhi);
    const double wd = (double)w;
    double tl = xlo * wd, th = xhi * wd;
    if (w < 0) {
      const double q = tl;
      tl = th;
      th = q;
    }
    rp.slo += tl;
    rp.shi += th;
    rp.sabs += fmax(fabs(tl), fabs(th));
    rp.mlo = fmax(rp.mlo, xlo);
    rp.mhi = fmax(rp.mhi, xhi);
    rp.wsum += w;
    if (link) rp.linkw += w;
    return;
  }
  const int64_t off = (i - A.r0) * A.ncols + (j - A.c0);
  if (link) atomicAdd(&A.row_links[i - A.r0], 1);
  uint16_t u = 0;
  if (t.unc) {
    uint64_t b = tn93g_bits(x), bl = tn93g_bits(xlo), bh = tn93g_bits(xhi);
    uint64_t dn = b - bl, up = bh - b;
    if (dn > 255 || up > 255) decided = false;
    u = (uint16_t)((dn > 255 ? 255 : dn) | (up > 255 ? 255 : up) << 8);
    if (dn || up) atomicAdd(&bs->unc, 1u);
  }
  if (!decided) atomicAdd(&bs->undecided, 1u);
  A.D[off] = x;
  A.U[off] = u;
}

// overlap too small: tn93's thisD is -1 (a link only if the caller's range reaches -1)
__device__ void skip_pair(const RunArgs &A, int64_t i, int64_t j, RowPart &rp)
{
  const bool link = -1.0 >= A.mind && -1.0 <= A.thr;
  if (A.fast) {
    rp.skip += 1;
    if (link) {
      append_link(A, i, j, -1.0, -1.0, -1.0);
      rp.linkw += A.unitw ? 1 : A.weights[i] * A.weights[j];
    }
    return;
  }
  const int64_t off = (i - A.r0) * A.ncols + (j - A.c0);
  A.D[off] = -1.0;
  A.U[off] = 0;
  if (link) atomicAdd(&A.row_links[i - A.r0], 1);
}

// adds one row's part to its accumulator
__device__ inline void rowacc_add(const RunArgs &A, int64_t i, const RowPart &rp)
{
  RowAcc *r = A.racc + (i - A.r0);
  if (rp.slo != 0.) atomicAdd(&r->slo, rp.slo);
  if (rp.shi != 0.) atomicAdd(&r->shi, rp.shi);
  if (rp.sabs != 0.) atomicAdd(&r->sabs, rp.sabs);
  if (rp.mlo > 0.) atomicMax(&r->mlo, (unsigned long long)tn93g_bits(rp.mlo));
  if (rp.mhi > 0.) atomicMax(&r->mhi, (unsigned long long)tn93g_bits(rp.mhi));
  if (rp.wsum) atomicAdd((unsigned long long *)&r->wsum, (unsigned long long)rp.wsum);
  if (rp.skip) atomicAdd((unsigned long long *)&r->skip, (unsigned long long)rp.skip);
  if (rp.linkw) atomicAdd((unsigned long long *)&r->linkw, (unsigned long long)rp.linkw);
}

__device__ inline double wave_sum_d(double x)
{
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ inline double wave_max_d(double x)
{
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ inline long long wave_sum_l(long long x)
{
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// ------------------------------------------------------------------------------------------ pack kernels
__device__ inline void add4(uint4 &a, const uint4 b)
{
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  a.w += b.w;
}
__device__ inline void sub4(uint4 &a, const uint4 b)
{
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
  a.w -= b.w;
}

__global__ void k_transpose(const uint8_t *__restrict__ codes, uint8_t *__restrict__ codesT, int64_t n, int64_t Ls,
                            int64_t npad, int64_t sbase)
{
  __shared__ uint8_t t[64][65];
  const int64_t s0 = sbase + (int64_t)blockIdx.y * 64, p0 = (int64_t)blockIdx.x * 64;
  for (int k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
    int r = k / 64, c = k % 64;
    int64_t s = s0 + r, p = p0 + c;
    t[r][c] = (s < n && p < Ls) ? codes[s * Ls + p] : (uint8_t)GAPC;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
    int r = k / 64, c = k % 64;  // r: site, c: sequence
    int64_t p = p0 + r, s = s0 + c;
    if (p < Ls && s < npad) codesT[p * npad + s] = t[c][r];
  }
}

// planes: per 32 sites, bit k of .x = site clean, .y = clean and bit0 of the code (C, T), .z = clean and bit1
// (G, T), .w = code T
__global__ void k_planes(const uint8_t *__restrict__ codes, uint4 *__restrict__ planes, uint32_t *__restrict__ nmask,
                         int64_t s0, int64_t s1, int64_t L, int64_t Ls, int64_t W)
{
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= (s1 - s0) * W) return;
  const int64_t s = s0 + t0 / W, w = t0 % W, t = s * W + w;
  uint32_t c = 0, h = 0, g = 0, tt = 0, nn = 0;
  for (int k = 0; k < 32; k++) {
    int64_t p = w * 32 + k;
    if (p >= L) break;
    unsigned code = codes[s * Ls + p];
    if (code < 4) {
      c |= 1u << k;
      h |= (code & 1u) << k;
      g |= (code >> 1 & 1u) << k;
      tt |= (code == 3 ? 1u : 0u) << k;
    } else if (code == (unsigned)NC) {
      nn |= 1u << k;
    }
  }
  planes[t] = make_uint4(c, h, g, tt);
  nmask[t] = nn;
}

// one-hot bytes for the MFMA backend, in operand order (see the layout above); sequences >= n are all zero
__global__ void k_onehot(const uint8_t *__restrict__ codes, uint8_t *__restrict__ onehot, int64_t n, int64_t L,
                         int64_t Ls, int64_t W64, int64_t sbase, int64_t send)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one thread per (sequence, site)
  if (t >= (send - sbase) * W64 * 64) return;
  const int64_t s = sbase + t / (W64 * 64), p = t % (W64 * 64);
  unsigned code = (s < n && p < L) ? codes[s * Ls + p] : (unsigned)GAPC;
  int64_t g = s >> 2, sl = s & 3, blk = p >> 6, kg = (p >> 4) & 3, m = p & 15;
  uint8_t *o = onehot + (((g * W64 + blk) * 4 + kg) * 16 + 4 * sl) * 16 + m;
  for (int b = 0; b < 4; b++) o[b * 16] = code == (unsigned)b ? 1 : 0;
}

// Event lists.  EV_AMB: the sites of codes 4..16 (ambiguous); EV_GAPS: also the gap sites (the lists of GAPMM,
// where a gap facing a non-gap is a site of its own); EV_NO_N: codes 4..16 but N, for sets whose N sites are long
// runs handled as a bit plane (nmask) and as run lists instead.
enum { EV_AMB = 0, EV_GAPS = 1, EV_NO_N = 2 };
__device__ inline bool is_event(unsigned code, int mode)
{
  if (code < 4) return false;
  if (code == (unsigned)GAPC) return mode == EV_GAPS;
  if (code == (unsigned)NC) return mode != EV_NO_N;
  return true;
}

// Runs of one code (N, or the gap) in each sequence: run k = [start, end) of consecutive sites holding it, in site
// order.
__global__ void k_runfill(const uint8_t *__restrict__ codes, const int32_t *__restrict__ off, uint2 *__restrict__ runs,
                          int64_t L, int64_t Ls, int64_t s0, int code)
{
  const int64_t s = s0 + blockIdx.x;
  const int lane = threadIdx.x;
  const uint8_t x = (uint8_t)code;
  int bs = off[s], be = off[s];
  for (int64_t p0 = 0; p0 < L; p0 += 64) {
    const int64_t p = p0 + lane;
    const bool in = p < L && codes[s * Ls + p] == x;
    const bool prev = p > 0 && p - 1 < L && codes[s * Ls + p - 1] == x;
    const bool st = in && !prev, en = !in && prev;  // a run ends where its first site of another code (or L) is
    const unsigned long long ms = __ballot(st), me = __ballot(en);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (st) runs[bs + __popcll(ms & below)].x = (uint32_t)p;
    if (en) runs[be + __popcll(me & below)].y = (uint32_t)p;
    bs += __popcll(ms);
    be += __popcll(me);
  }
  if (lane == 0 && be < bs) runs[be].y = (uint32_t)L;  // a run reaching the last site
}

// pre[s][w] = plane counts (clean, C|T, G|T, T) of sequence s over its words [0, w), w = 0..W: one wave per sequence,
// 64 word boundaries per step
__global__ void k_prefix(const uint4 *__restrict__ planes, const uint32_t *__restrict__ nmask,
                         PrefixRec *__restrict__ pre, uint4 *__restrict__ tb, uint32_t *__restrict__ npre, int64_t W,
                         int64_t s0)
{
  const int64_t s = s0 + blockIdx.x;
  const int lane = threadIdx.x;
  uint4 carry = make_uint4(0, 0, 0, 0);
  uint32_t ncarry = 0;
  for (int64_t w0 = 0; w0 <= W; w0 += 64) {
    const int64_t w = w0 + lane;
    uint4 v = make_uint4(0, 0, 0, 0);
    uint32_t nv = 0;
    if (w > 0 && w <= W) {
      const uint4 q = planes[s * W + w - 1];
      v = make_uint4(__popc(q.x), __popc(q.y), __popc(q.z), __popc(q.w));
      nv = __popc(nmask[s * W + w - 1]);
    }
    for (int o = 1; o < 64; o <<= 1) {  // inclusive scan over the wave
      const uint4 t = make_uint4(__shfl_up(v.x, o, 64), __shfl_up(v.y, o, 64), __shfl_up(v.z, o, 64),
                                 __shfl_up(v.w, o, 64));
      const uint32_t tn = __shfl_up(nv, o, 64);
      if (lane >= o) {
        add4(v, t);
        nv += tn;
      }
    }
    add4(v, carry);
    nv += ncarry;
    if (w <= W) {
      pre[s * (W + 1) + w] = PrefixRec{v, w < W ? planes[s * W + w] : make_uint4(0, 0, 0, 0)};
      npre[s * (W + 1) + w] = nv;
    }
    if (w == W) tb[s] = v;
    carry = make_uint4(__shfl(v.x, 63, 64), __shfl(v.y, 63, 64), __shfl(v.z, 63, 64), __shfl(v.w, 63, 64));
    ncarry = __shfl(nv, 63, 64);
  }
}

// Every list's length for sequences s0 + blockIdx.x (one wave each) in one pass over the codes: cnt[k * n + s - s0]
// for k = 0: EV_AMB events, 1: EV_NO_N events, 2: EV_GAPS events, 3: N runs, 4: gap runs.
__global__ void k_listcount(const uint8_t *__restrict__ codes, int32_t *__restrict__ cnt, int64_t L, int64_t Ls,
                            int64_t s0, int64_t n)
{
  const int64_t s = s0 + blockIdx.x;
  const uint8_t *q = codes + s * Ls;
  int c[5] = {0, 0, 0, 0, 0};
  for (int64_t p = threadIdx.x; p < L; p += blockDim.x) {
    const unsigned v = q[p], u = p ? q[p - 1] : 255u;
    c[0] += is_event(v, EV_AMB) ? 1 : 0;
    c[1] += is_event(v, EV_NO_N) ? 1 : 0;
    c[2] += is_event(v, EV_GAPS) ? 1 : 0;
    c[3] += v == (unsigned)NC && u != (unsigned)NC ? 1 : 0;
    c[4] += v == (unsigned)GAPC && u != (unsigned)GAPC ? 1 : 0;
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    for (int o = 32; o > 0; o >>= 1) c[k] += __shfl_down(c[k], o, 64);
    if (threadIdx.x == 0) cnt[k * n + blockIdx.x] = c[k];
  }
}

// events in site order: one wave per sequence
__global__ void k_evfill(const uint8_t *__restrict__ codes, const int32_t *__restrict__ ev_off,
                         uint32_t *__restrict__ ev, int64_t L, int64_t Ls, int gaps, int64_t s0)
{
  const int64_t s = s0 + blockIdx.x;
  const int lane = threadIdx.x;
  int base = ev_off[s];
  for (int64_t p0 = 0; p0 < L; p0 += 64) {
    int64_t p = p0 + lane;
    const unsigned code = p < L ? codes[s * Ls + p] : 0u;
    const bool f = is_event(code, gaps);
    unsigned long long m = __ballot(f);
    if (f) {
      int k = __popcll(m & ((1ull << lane) - 1ull));
      ev[base + k] = (uint32_t)(p << 5) | code;
    }
    base += __popcll(m);
  }
}

// ------------------------------------------------------------------------------------------ pair kernel
// Tile: TI rows x 64 columns, 256 threads = 4 waves; lane = column j, each wave owns RP rows (uniform i), TI = 4 RP.
// The pair kernel is instantiated for RP = 2, 4 and 8 (tiles of 8, 16 and 32 rows: the autotuner's shapes); RPW and
// TI are the default shape, which the other kernels built on the same tiling (k_spec16) use.
constexpr int RPW = 4, TI = 4 * RPW, TJ = 64, KW = 16;
constexpr int EVCAP = 2048;  // events of a tile's rows and columns held in LDS

struct Cnt {
  uint32_t T, DG, TS, CT, B0, B1, BT;
};

// v_bitop3_b32 (gfx950): any bitwise function of three operands in one instruction; the truth table is the function
// evaluated on 0xf0, 0xcc, 0xaa (first, second, third operand).
constexpr unsigned BOP_A_NB_NC = 0x10;  // a & ~b & ~c
constexpr unsigned BOP_A_NB_C = 0x20;   // a & ~b & c

// Per word pair: M = both clean; x, y = the C|T and G|T bits differ.  Same base: M & ~x & ~y; a transition (A<->G
// or C<->T): M & ~x & y, a C<->T one when the C|T bit is set.
__device__ inline void count_word(const uint4 pi, const uint4 pj, Cnt &c)
{
  const uint32_t M = pi.x & pj.x;
  const uint32_t x = pi.y ^ pj.y, y = pi.z ^ pj.z;
  c.T += __popc(M);
  c.DG += __popc(__builtin_amdgcn_bitop3_b32(M, x, y, BOP_A_NB_NC));
  const uint32_t tr = __builtin_amdgcn_bitop3_b32(M, x, y, BOP_A_NB_C);
  c.TS += __popc(tr);
  c.CT += __popc(tr & pi.y);
  c.B0 += __popc(pj.x & pi.y) + __popc(pi.x & pj.y);
  c.B1 += __popc(pj.x & pi.z) + __popc(pi.x & pj.z);
  c.BT += __popc(pj.x & pi.w) + __popc(pi.x & pj.w);
}

// The same without the base sums.  With per-sequence totals the base sums over the sites where both sequences are
// clean are the totals of both sequences less their clean sites facing a non-clean site of the other sequence, and
// those sites are all visited anyway: ambiguity events (their table entries subtract the clean base), N runs of
// dense-N sets (the N-plane counts, through the same entries) and gap runs (prefix counts).  11 operations per word
// pair instead of 24 (9 with v_bitop3_b32).
struct CntM {
  uint32_t DG, TS, CT;
};

__device__ inline void count_word(const uint4 pi, const uint4 pj, CntM &c)
{
  const uint32_t M = pi.x & pj.x;
  const uint32_t x = pi.y ^ pj.y, y = pi.z ^ pj.z;
  c.DG += __popc(__builtin_amdgcn_bitop3_b32(M, x, y, BOP_A_NB_NC));
  const uint32_t tr = __builtin_amdgcn_bitop3_b32(M, x, y, BOP_A_NB_C);
  c.TS += __popc(tr);
  c.CT += __popc(tr & pi.y);
}

// Plane counts (clean, C|T, G|T, T) of sequence s over the sites [0, p): the prefix at the word boundary plus the
// bits of the partial word.
__device__ inline uint4 plane_prefix(const PrefixRec *__restrict__ pre, int64_t W, int64_t s, uint32_t p)
{
  const PrefixRec e = pre[s * (W + 1) + (p >> 5)];  // both halves: one cache line
  uint4 r = e.pre;
  const uint32_t b = p & 31u;
  if (b) {
    const uint4 v = e.word;
    const uint32_t m = (1u << b) - 1u;
    r.x += __popc(v.x & m);
    r.y += __popc(v.y & m);
    r.z += __popc(v.z & m);
    r.w += __popc(v.w & m);
  }
  return r;
}

// N sites of sequence s over [0, p) (dense-N sets: nmask and its prefix counts).
__device__ inline uint32_t n_prefix(const uint32_t *__restrict__ npre, const uint32_t *__restrict__ nmask, int64_t W,
                                    int64_t s, uint32_t p)
{
  const int64_t w = p >> 5;
  uint32_t r = npre[s * (W + 1) + w];
  const uint32_t b = p & 31u;
  if (b) r += __popc(nmask[s * W + w] & ((1u << b) - 1u));
  return r;
}

// event accumulators: packed 16-bit fields plus 32-bit totals (flushed before a 16-bit field can overflow)
constexpr int EVFLUSH = 256;  // entries summed in the packed words between flushes
// The four packed words are added as two 64-bit words (one v_lshl_add_u64 each): no field overflows between flushes,
// so no carry crosses from one field (or word) into the next.
struct EvAcc {
  unsigned long long pxy, pzw;  // px | py << 32, pz | pw << 32
  uint32_t fl;                  // non-zero once an order-dependent entry was added
};
struct Tot {
  uint32_t nf[4], dg, ag, ct, amb;
};

__device__ inline void ev_add(EvAcc &e, const uint4 t)
{
  e.pxy += (unsigned long long)t.x | (unsigned long long)t.y << 32;
  e.pzw += (unsigned long long)t.z | (unsigned long long)t.w << 32;
}
__device__ inline void ev_flush(EvAcc &e, Tot &o)
{
  const uint32_t px = (uint32_t)e.pxy, py = (uint32_t)(e.pxy >> 32), pz = (uint32_t)e.pzw,
                 pw = (uint32_t)(e.pzw >> 32);
  o.nf[0] += px & 0xffffu;
  o.nf[1] += px >> 16;
  o.nf[2] += py & 0xffffu;
  o.nf[3] += py >> 16;
  o.dg += pz & 0xffffu;
  o.ag += pz >> 16;
  o.ct += pw & 0x3fffu;
  o.amb += (pw >> 14) & 0x1ffu;
  e.fl |= pw >> 23;
  e.pxy = e.pzw = 0;
}

typedef int v4i __attribute__((ext_vector_type(4)));

// sum over the 4 lanes of a quad
__device__ inline uint32_t quad_sum(uint32_t x)
{
  x += (uint32_t)__shfl_xor((int)x, 1, 64);
  x += (uint32_t)__shfl_xor((int)x, 2, 64);
  return x;
}

// LDS of the pair kernel.  Popcount backend: per chunk of KW words, the bit planes of the tile's rows and columns,
// padded by one element per word so that lanes storing different words of one sequence use different LDS banks.
// MFMA backend: the operands of one 64-site block, and at the end the per-pair counts transposed into the
// lane = column layout of the epilogue.
template <int KWORDS, bool MFMA, int TI_>
struct PairLds {
  union {
    uint4 ops[(TI_ + TJ) / 4][64];  // one 64-site block: TI/4 row groups + 16 column groups, 1 KB each
    uint32_t agg[7][TI_][TJ];       // T, DG, TS, CT, B0, B1, BT per pair
  } u;
};
template <int KWORDS, int TI_>
struct PairLds<KWORDS, false, TI_> {
  uint4 sI[KWORDS][TI_ + 1];  // the rows (unless they are read with scalar loads)
  uint4 sJ[KWORDS][TJ + 1];
};

// Staging of the planes of words [w0, w0 + kw) of the tile's sequences, in two steps so that the loads of the next
// chunk are in flight while the current one is counted: SPT values per thread, N-plane words too for dense-N sets.
// Every sequence has KW slots whatever kw is: thread t takes word t % KW of sequences t / KW + 16 q, so the slot of
// a thread is fixed for the whole tile (no division by kw) and consecutive threads read consecutive words of one
// sequence; slots at or past kw load nothing and are never read.  SR: only the 64 columns are staged; the rows' words
// are the same in every lane of a wave and the counting loop reads them with scalar loads, which keeps them off the
// LDS: the LDS read bandwidth bounds the loop when every lane reads a row word (and its N word) per word and row.
static_assert(KW == 16, "staging slots: 16 words per sequence, 16 sequences per 256 threads");
template <int TI_, bool SR>
constexpr int spt() { return (KW * ((SR ? 0 : TI_) + TJ) + 255) / 256; }
template <bool NP, int TI_, bool SR>
struct StageRegs {
  uint4 v[spt<TI_, SR>()];
  uint32_t nm[NP ? spt<TI_, SR>() : 1];
};
template <bool NP, int TI_, bool SR>
__device__ __attribute__((always_inline)) inline void stage_load(const RunArgs &A, int64_t i0, int64_t j0, int64_t w0,
                                                                 int kw, int tid, StageRegs<NP, TI_, SR> &r)
{
  constexpr int R = SR ? 0 : TI_;
  const int w = tid & (KW - 1);
#pragma unroll
  for (int q = 0; q < spt<TI_, SR>(); q++) {
    const int s = (tid >> 4) + 16 * q;
    const int64_t seq = s < R ? i0 + s : j0 + (s - R);
    const bool ok = s < R + TJ && w < kw && seq < A.n;
    r.v[q] = ok ? A.planes[seq * A.W + w0 + w] : make_uint4(0, 0, 0, 0);
    if constexpr (NP) r.nm[q] = ok ? A.nmask[seq * A.W + w0 + w] : 0u;
  }
}
template <bool NP, int TI_, bool SR, class LDS>
__device__ __attribute__((always_inline)) inline void stage_store(LDS &lds, uint32_t (*sNI)[TI_ + 1],
                                                                  uint32_t (*sNJ)[TJ + 1], int tid,
                                                                  const StageRegs<NP, TI_, SR> &r)
{
  constexpr int R = SR ? 0 : TI_;
  const int w = tid & (KW - 1);
#pragma unroll
  for (int q = 0; q < spt<TI_, SR>(); q++) {
    const int s = (tid >> 4) + 16 * q;
    if (s >= R + TJ) break;
    if (s < R) {
      lds.sI[w][s] = r.v[q];
      if constexpr (NP) sNI[w][s] = r.nm[q];
    } else {
      lds.sJ[w][s - R] = r.v[q];
      if constexpr (NP) sNJ[w][s - R] = r.nm[q];
    }
  }
}

// Scalar (SMEM) views of the planes and N masks: loads through them with an address that is the same in every lane
// go to SGPRs, which the VALU reads as operands.  The planes do not change during a run.
typedef __attribute__((address_space(4))) const uint32_t suint32;
__device__ inline uint4 sload4(suint32 *p) { return make_uint4(p[0], p[1], p[2], p[3]); }

// Events of a tile.  The rows' and the columns' event lists are each one contiguous range of the set's CSR array:
// rows [er0, er0 + nr), columns [ec0, ec0 + nc).  They are taken in chunks that fit the LDS: chunk rows [r0, r1) and
// columns [c0, c1) of those ranges (relative), row events first.  A chunk holds its event words in sEv (rows, then
// columns) and, in the LDS of the counting buffers, the codes its events face: row event e the codes of the tile's
// 64 columns at its site (g[64 e ...]), column event c the codes of the tile's TI rows at its site
// (g[64 R + TI c ...], R = r1 - r0; cb = TI bytes per column event).  Almost every tile is one chunk.
constexpr int GCAP = 30 * 1024;
template <bool MF, int TI_>
union TileLds {
  PairLds<KW, MF, TI_> p;
  uint8_t g[GCAP];
};

struct TileEv {
  int er0, ec0, nr, nc;
  int r0, r1, c0, c1;
  int cb;  // bytes of row codes per column event (the tile's rows)
  __device__ inline void next()
  {
    r0 = r1;
    c0 = c1;
    const int R = min(nr - r0, min(GCAP / 64, EVCAP));
    r1 = r0 + R;
    c1 = c0 + min(nc - c0, min((GCAP - 64 * R) / cb, EVCAP - R));
  }
  __device__ inline void first()
  {
    r1 = c1 = 0;
    next();
  }
  __device__ inline bool last() const { return r1 == nr && c1 == nc; }
  __device__ inline bool single() const { return r0 == 0 && c0 == 0 && last(); }
};

// The chunk's event words into sEv.
__device__ inline void chunk_events(const RunArgs &A, const TileEv &te, uint32_t *sEv, int tid)
{
  const int R = te.r1 - te.r0, C = te.c1 - te.c0;
  for (int k = tid; k < R; k += 256) sEv[k] = A.ev[te.er0 + te.r0 + k];
  for (int k = tid; k < C; k += 256) sEv[R + k] = A.ev[te.ec0 + te.c0 + k];
}

// The codes the chunk's events face, from the site-major codes into g (16 bytes per load, 8 for the columns of a
// tile of 8 rows; all loads of a thread in flight together).  i0 is a multiple of TI and j0 of 64, so every load is
// aligned to its size.
template <int TI_>
__device__ __attribute__((always_inline)) inline void chunk_codes(const RunArgs &A, const TileEv &te,
                                                                   const uint32_t *sEv, uint8_t *g, int64_t i0,
                                                                   int64_t j0, int tid)
{
  const size_t np = (size_t)A.npad;
  const int R = te.r1 - te.r0, C = te.c1 - te.c0;
  const int nrow = 4 * R;
  if constexpr (TI_ >= 16) {
    constexpr int PC = TI_ / 16;  // 16-byte loads per column event
    const int tot = nrow + PC * C;
#pragma unroll 4
    for (int k = tid; k < tot; k += 256) {
      const uint8_t *src;
      if (k < nrow) {
        src = A.codesT + (size_t)(sEv[k >> 2] >> 5) * np + j0 + (k & 3) * 16;
      } else {
        const int q = k - nrow, ce = q / PC, part = q % PC;
        src = A.codesT + (size_t)(sEv[R + ce] >> 5) * np + i0 + part * 16;
      }
      *reinterpret_cast<uint4 *>(g + k * 16) = *reinterpret_cast<const uint4 *>(src);
    }
  } else {
    static_assert(TI_ == 8, "tiles of 8, 16 or 32 rows");
#pragma unroll 4
    for (int k = tid; k < nrow; k += 256)
      *reinterpret_cast<uint4 *>(g + k * 16) =
          *reinterpret_cast<const uint4 *>(A.codesT + (size_t)(sEv[k >> 2] >> 5) * np + j0 + (k & 3) * 16);
#pragma unroll 4
    for (int c = tid; c < C; c += 256)
      *reinterpret_cast<uint2 *>(g + nrow * 16 + c * 8) =
          *reinterpret_cast<const uint2 *>(A.codesT + (size_t)(sEv[R + c] >> 5) * np + i0);
  }
}

// The codes of a wave's RP rows at a column event's site (one byte per row).
template <int RP>
using RowCodes = typename std::conditional<RP == 8, unsigned long long, uint32_t>::type;
template <int RP>
__device__ inline RowCodes<RP> row_codes(const uint8_t *p)
{
  if constexpr (RP == 2)
    return *reinterpret_cast<const uint16_t *>(p);
  else if constexpr (RP == 4)
    return *reinterpret_cast<const uint32_t *>(p);
  else
    return *reinterpret_cast<const unsigned long long *>(p);
}

// Event tables in LDS, TABW entries per event code; code EVNONE is an all-zero row that pads partial batches.
//   sTabR[c][x]  a row event of code c facing the column's code x (any code: a site where both sequences are
//                ambiguous is counted here);
//   sTabC[c][x]  a column event of code c facing the row's code x, zero where x is ambiguous (that site is counted
//                as the row's event).
constexpr int TABW = NCODE + 1, EVNONE = NCODE, EVB = 4;
constexpr int BACKEND_GEMMC = 3;  // the pair kernel's clean/clean counts from int8 GEMMs (rocBLAS)

// Fills the LDS event tables from one [NCODE][NCODE] table of packed entries (entry [row code][column code]).
// With `gaps` (GAPMM), gap sites are events too, so a column event facing a row gap is the row's event.  With
// `dense_n`, N sites are not events, so a column event facing a row N counts the site itself.
__device__ inline void load_event_tables(const uint4 *src, uint4 *sTabR, uint4 *sTabC, int tid, bool gaps,
                                         bool dense_n)
{
  for (int k = tid; k < (NCODE + 1) * TABW; k += 256) {
    const int c = k / TABW, x = k % TABW;
    const bool in = c < NCODE && x < NCODE;
    const bool xamb = x >= 4 && (x < GAPC || gaps) && !(dense_n && x == NC);
    sTabR[k] = in ? src[c * NCODE + x] : make_uint4(0, 0, 0, 0);
    sTabC[k] = in && !xamb ? src[x * NCODE + c] : make_uint4(0, 0, 0, 0);
  }
}

__device__ inline uint4 keep_if(bool k, uint4 t) { return k ? t : make_uint4(0, 0, 0, 0); }

// The events of one chunk for the wave's RPW rows against the lane's column, every operand in LDS.  Row events
// are uniform over the wave; column events are per lane, the codes of the wave's rows at the site being one 32-bit
// word of the event's 16-byte record.  sRo[TI + 1]: the rows' event offsets relative to the tile's row list.  The
// packed words are flushed every EVFLUSH events.  MASKED: only the (row, lane) pairs with use[a] set take the entries
// (the others add zero), and rows or lanes with no such pair skip the loops.
template <bool MASKED, int RP>
__device__ __attribute__((always_inline)) inline void chunk_apply(const RunArgs &A, const uint4 *sTabR,
                                                                  const uint4 *sTabC, const TileEv &te,
                                                                  const uint32_t *sEv, const uint8_t *g,
                                                                  const int *sRo, int64_t ib, int wave, int lane,
                                                                  int pj, int pj_end, const bool (&use)[RP],
                                                                  EvAcc (&acc)[RP], Tot (&tot)[RP])
{
  constexpr int TI_ = 4 * RP;
  const int R = te.r1 - te.r0;
  // Row events are the same in every lane: their bounds and words go to SGPRs (the words by scalar loads from the
  // set's event list, which the chunk's sEv copies), so the loop and the table row are scalar work.
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  suint32 *evs = (suint32 *)(A.ev + te.er0 + te.r0);
#pragma unroll
  for (int a = 0; a < RP; a++) {
    const int64_t i = ib + a;
    if (i >= A.r1) continue;  // uniform over the wave
    if (MASKED && !__ballot(use[a])) continue;
    const int e0 = __builtin_amdgcn_readfirstlane(max(sRo[wv * RP + a], te.r0) - te.r0);
    const int e1 = __builtin_amdgcn_readfirstlane(min(sRo[wv * RP + a + 1], te.r1) - te.r0);
    for (int h0 = e0; h0 < e1; h0 += EVFLUSH) {
    const int h1 = min(e1, h0 + EVFLUSH);
    int e = h0;
    for (; e + EVB <= h1; e += EVB) {
      uint32_t c[EVB], x[EVB];
      const uint8_t *gx = g + e * 64 + lane;
#pragma unroll
      for (int k = 0; k < EVB; k++) {
        c[k] = (evs[e + k] & 31u) * TABW;
        x[k] = gx[k * 64];
      }
#pragma unroll
      for (int k = 0; k < EVB; k++) {
        const uint4 t = sTabR[c[k] + x[k]];
        ev_add(acc[a], MASKED ? keep_if(use[a], t) : t);
      }
    }
    for (; e < h1; e++) {
      const uint4 t = sTabR[(evs[e] & 31u) * TABW + g[e * 64 + lane]];
      ev_add(acc[a], MASKED ? keep_if(use[a], t) : t);
    }
    ev_flush(acc[a], tot[a]);
    }
  }
  if constexpr (MASKED) {
    bool any = false;
#pragma unroll
    for (int a = 0; a < RP; a++) any |= use[a];
    if (!any) return;
  }
  const int k0 = max(pj - te.ec0, te.c0) - te.c0, k1 = min(pj_end - te.ec0, te.c1) - te.c0;
  const uint32_t *evc = sEv + R;
  const uint8_t *gc = g + R * 64 + wave * RP;
  for (int f0 = k0; f0 < k1; f0 += EVFLUSH) {
    const int f1 = min(k1, f0 + EVFLUSH);
    for (int k = f0; k < f1; k++) {
      const uint4 *t = sTabC + (evc[k] & 31u) * TABW;
      const RowCodes<RP> rc = row_codes<RP>(gc + k * TI_);
#pragma unroll
      for (int a = 0; a < RP; a++) {
        const uint4 v = t[(uint32_t)(rc >> (8 * a)) & 255u];
        ev_add(acc[a], MASKED ? keep_if(use[a], v) : v);
      }
    }
#pragma unroll
    for (int a = 0; a < RP; a++) ev_flush(acc[a], tot[a]);
  }
}

// All events of the tile, chunk by chunk.  The first chunk's event words and codes are in LDS already (ready); every
// further chunk is loaded between barriers, so all threads of the block call this together.
template <bool MASKED, int RP>
__device__ __attribute__((always_inline)) inline void tile_events(const RunArgs &A, const uint4 *sTabR,
                                                                  const uint4 *sTabC, TileEv &te, uint32_t *sEv,
                                                                  uint8_t *g, const int *sRo, int64_t i0, int64_t j0,
                                                                  int64_t ib, int tid, int pj, int pj_end,
                                                                  const bool (&use)[RP], EvAcc (&acc)[RP],
                                                                  Tot (&tot)[RP])
{
  const int wave = tid >> 6, lane = tid & 63;
  for (;;) {
    chunk_apply<MASKED, RP>(A, sTabR, sTabC, te, sEv, g, sRo, ib, wave, lane, pj, pj_end, use, acc, tot);
    if (te.last()) return;
    te.next();
    __syncthreads();
    chunk_events(A, te, sEv, tid);
    __syncthreads();
    chunk_codes<4 * RP>(A, te, sEv, g, i0, j0, tid);
    __syncthreads();
  }
}

struct NCnt {
  uint32_t r[4], c[4], nn;
};

__device__ inline void tot_add_mul(Tot &t, EvAcc &acc, const uint4 e, uint32_t k)
{
  if (!k) return;
  t.nf[0] += (e.x & 0xffffu) * k;
  t.nf[1] += (e.x >> 16) * k;
  t.nf[2] += (e.y & 0xffffu) * k;
  t.nf[3] += (e.y >> 16) * k;
  t.dg += (e.z & 0xffffu) * k;
  t.ag += (e.z >> 16) * k;
  t.ct += (e.w & 0x3fffu) * k;
  t.amb += ((e.w >> 14) & 0x1ffu) * k;
  acc.fl |= e.w >> 23;
}

__device__ __attribute__((always_inline)) inline void nplane_add(const NCnt &q, const uint4 *sTabR, Tot &t, EvAcc &acc)
{
  // base counts A, C, G, T from (clean, C|T, G|T, T)
  const uint32_t rb[4] = {q.r[0] - q.r[1] - q.r[2] + q.r[3], q.r[1] - q.r[3], q.r[2] - q.r[3], q.r[3]};
  const uint32_t cb[4] = {q.c[0] - q.c[1] - q.c[2] + q.c[3], q.c[1] - q.c[3], q.c[2] - q.c[3], q.c[3]};
#pragma unroll
  for (int b = 0; b < 4; b++) {
    tot_add_mul(t, acc, sTabR[NC * TABW + b], rb[b]);  // row N, column base b
    tot_add_mul(t, acc, sTabR[b * TABW + NC], cb[b]);  // row base b, column N
  }
  tot_add_mul(t, acc, sTabR[NC * TABW + NC], q.nn);
}

// Gap runs of a tile: the rows' runs then the columns' runs (each a contiguous range of the set's run list), in LDS
// when they fit.
constexpr int GRCAP = 128;
struct TileRuns {
  const uint2 *s;
  int gr0, nrr, gc0;
  bool lds;
  __device__ inline uint2 row(const RunArgs &A, int g) const { return lds ? s[g] : A.gruns[gr0 + g]; }
  __device__ inline uint2 col(const RunArgs &A, int g) const { return lds ? s[nrr + g - gc0] : A.gruns[g]; }
};

// Base sums of the marginal form: S[a] = plane totals of row ib + a and of the lane's column (sTB: the tile's rows,
// then its columns), less the clean sites of each facing a gap run of the other (gaps are not events outside GAPMM,
// so no table entry sees them).  sRg[TI + 1]: the rows' run offsets relative to the tile's row runs; [h0, h1): the
// column's runs in the set's list.
template <int RP>
__device__ __attribute__((always_inline)) inline void base_sums(const RunArgs &A, const uint4 *sTB,
                                                                const TileRuns &tr, const int *sRg, int64_t ib,
                                                                int wave, int lane, int64_t j, int h0, int h1,
                                                                uint4 (&S)[RP])
{
  const int64_t W = A.W;
  const bool jin = j < A.n;
#pragma unroll
  for (int a = 0; a < RP; a++) {
    S[a] = sTB[4 * RP + lane];
    add4(S[a], sTB[wave * RP + a]);
  }
  if (A.gapmm) return;
#pragma unroll
  for (int a = 0; a < RP; a++) {
    const int64_t i = ib + a;
    if (i >= A.r1) continue;  // uniform over the wave
    const int g0 = sRg[wave * RP + a], g1 = sRg[wave * RP + a + 1];
    for (int g = g0; g < g1; g++) {
      const uint2 r = tr.row(A, g);
      if (jin) {
        sub4(S[a], plane_prefix(A.pre, W, j, r.y));
        add4(S[a], plane_prefix(A.pre, W, j, r.x));
      }
    }
  }
  for (int g = h0; g < h1; g++) {
    const uint2 r = tr.col(A, g);
#pragma unroll
    for (int a = 0; a < RP; a++) {
      const int64_t i = ib + a;
      if (i >= A.r1) continue;
      sub4(S[a], plane_prefix(A.pre, W, i, r.y));
      add4(S[a], plane_prefix(A.pre, W, i, r.x));
    }
  }
}

// MARG: the base sums come from base_sums and tables whose entries facing a clean base subtract it (popcount
// backend, every mode but SKIP, whose ambiguity sites are not visited).
// One pair tile: rows A.r0 + by * TI .. + TI, columns (c0 rounded down to 64) + bx * TJ .. + TJ; TI = 4 RP.
template <int BACKEND, bool FAST, bool NP, bool MARG, int RP>
__device__ __attribute__((always_inline)) inline void pairs_tile(const RunArgs &A, const int64_t bx, const int64_t by)
{
  constexpr int TI_ = 4 * RP;
  const int64_t c0a = A.c0 & ~(int64_t)63;  // column tiles start on a 64 boundary (whole 4-sequence groups)
  const int64_t i0 = A.r0 + by * TI_, j0 = c0a + bx * TJ;
  if (A.triangle && j0 + TJ - 1 <= i0) return;  // no pair j > i in this tile
  constexpr bool MF = BACKEND == TN93G_BACKEND_MFMA;
  constexpr bool GC = BACKEND == BACKEND_GEMMC;  // clean/clean counts read from the GEMM products
  static_assert(!GC || (MARG && NP && FAST && RP == 4), "the GEMM-count variant: fast dense-N marginal form");
  constexpr bool SR = NP;  // dense-N sets (long N-masked alignments): the rows' words through scalar loads
  __shared__ unsigned lslow;
  __shared__ unsigned long long sbase;
  __shared__ BlockStats bs;
  if (threadIdx.x == 0) {
    lslow = 0;
    bs.unc = bs.undecided = 0;
  }
  __shared__ TileLds<MF, TI_> L;
  PairLds<KW, MF, TI_> &lds = L.p;
  __shared__ uint32_t sNI[NP && !SR ? KW : 1][TI_ + 1], sNJ[NP ? KW : 1][TJ + 1];
  __shared__ uint4 sTabR[(NCODE + 1) * TABW], sTabC[(NCODE + 1) * TABW];
  __shared__ unsigned long long lhist[FAST ? 1 : 3 * 200];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  StageRegs<NP, TI_, SR> stg;  // the first chunk's planes, loaded before everything else
  if constexpr (!MF && !GC)
    if (!A.all_slow) stage_load<NP, TI_, SR>(A, i0, j0, 0, (int)min((int64_t)KW, A.W), tid, stg);
  const uint4 *tabs = A.evtab + (MARG ? 2 * NCODE * NCODE : 0);  // [main, AVERAGE], plain or folded
  load_event_tables(tabs, sTabR, sTabC, tid, A.gapmm, A.nplane);
  if constexpr (!FAST)
    for (int k = tid; k < 600; k += 256) lhist[k] = 0;
  const int64_t j = j0 + lane;
  const int64_t ib = i0 + wave * RP;
  using CntT = typename std::conditional<MARG, CntM, Cnt>::type;
  CntT cnt[RP];
  uint4 S[MARG ? RP : 1];
  Tot tot[RP];
  EvAcc acc[RP];
  NCnt ncn[NP ? RP : 1];
#pragma unroll
  for (int a = 0; a < RP; a++) {
    cnt[a] = CntT{};
    tot[a] = Tot{{0, 0, 0, 0}, 0, 0, 0, 0};
    acc[a] = EvAcc{0, 0, 0};
  }
  if constexpr (NP)
#pragma unroll
    for (int a = 0; a < RP; a++) ncn[a] = NCnt{{0, 0, 0, 0}, {0, 0, 0, 0}, 0};
  const bool do_fast = !A.all_slow;
  const bool do_events = do_fast && !A.skip_events;
  // column event cursors (columns past the set have none)
  const int pj = j < A.n ? A.ev_off[j] : 0, pj_end = j < A.n ? A.ev_off[j + 1] : 0;
  // the tile's event lists, the rows' event offsets, and (marginal form) the tile's plane totals and gap runs into
  // LDS when they fit (synchronised by the first barrier of the counts below)
  __shared__ uint32_t sEv[EVCAP];
  __shared__ int sRo[TI_ + 1], sRg[MARG ? TI_ + 1 : 1];
  __shared__ uint4 sTB[MARG ? TI_ + TJ : 1];
  __shared__ uint2 sGR[MARG ? GRCAP : 1];
  TileEv te;
  te.cb = TI_;
  TileRuns tr{sGR, 0, 0, 0, false};
  int h0 = 0, h1 = 0;
  {
    const int64_t rl = i0 < A.r1 ? min(i0 + (int64_t)TI_, A.r1) : i0;
    const int64_t cl = min(j0 + (int64_t)TJ, A.n);
    te.er0 = i0 < A.r1 ? A.ev_off[i0] : 0;
    te.nr = i0 < A.r1 ? A.ev_off[rl] - te.er0 : 0;
    te.ec0 = j0 < A.n ? A.ev_off[j0] : 0;
    te.nc = j0 < A.n ? A.ev_off[cl] - te.ec0 : 0;
    te.first();
    if (do_events) {
      chunk_events(A, te, sEv, tid);
      if (tid <= TI_) sRo[tid] = (i0 < A.r1 ? A.ev_off[min(i0 + tid, rl)] : 0) - te.er0;
    }
    if constexpr (MARG) {
      if (do_fast) {
        if (tid < TI_ + TJ) {
          const int64_t q = tid < TI_ ? i0 + tid : j0 + (tid - TI_);
          sTB[tid] = q < A.n ? A.tb[q] : make_uint4(0, 0, 0, 0);
        }
        if (j < A.n) {
          h0 = A.grun_off[j];
          h1 = A.grun_off[j + 1];
        }
        tr.gr0 = i0 < A.r1 ? A.grun_off[i0] : 0;
        tr.nrr = i0 < A.r1 ? A.grun_off[rl] - tr.gr0 : 0;
        tr.gc0 = j0 < A.n ? A.grun_off[j0] : 0;
        const int ncr = j0 < A.n ? A.grun_off[cl] - tr.gc0 : 0;
        tr.lds = tr.nrr + ncr <= GRCAP;
        if (tr.lds) {
          for (int k = tid; k < tr.nrr; k += 256) sGR[k] = A.gruns[tr.gr0 + k];
          for (int k = tid; k < ncr; k += 256) sGR[tr.nrr + k] = A.gruns[tr.gc0 + k];
        }
        if (tid <= TI_) sRg[tid] = (i0 < A.r1 ? A.grun_off[min(i0 + tid, rl)] : 0) - tr.gr0;
      }
    }
  }
  if constexpr (GC) {
    // DG = AA + CC + GG + TT, TS = (RR + YY) - DG, CT = YY - (CC + TT): the marginal form's three counts
    // (a triangle's products hold the pairs j > i only)
    if (do_fast && j >= A.c0 && j < A.c1)
#pragma unroll
      for (int a = 0; a < RP; a++) {
        const int64_t i = ib + a;
        if (i >= A.r1 || (A.triangle && j <= i)) continue;
        const size_t o = (size_t)(i - A.r0) * (size_t)A.ncols + (size_t)(j - A.c0);
        const uint32_t p1 = A.gcnt[o], p2 = A.gcnt[o + A.gcs], p3 = A.gcnt[o + 2 * A.gcs], p4 = A.gcnt[o + 3 * A.gcs];
        cnt[a].DG = p1 + p2;
        cnt[a].TS = p3 + p4 - (p1 + p2);
        cnt[a].CT = p4 - p2;
      }
  }
  if constexpr (!MF && !GC) {
    if (do_fast) {
      // SR: the wave's rows through scalar loads, RG words at a time (words past the row's end read the next row or
      // the arrays' padding and are not used; row indices past the set read its last sequence, whose counts are
      // never used)
      constexpr int RG = 2;
      const int64_t ibs = i0 + (int64_t)__builtin_amdgcn_readfirstlane(wave) * RP;
      suint32 *rowp[SR ? RP : 1];
      suint32 *rown[SR && NP ? RP : 1];
      if constexpr (SR)
#pragma unroll
        for (int a = 0; a < RP; a++) {
          const int64_t ia = min(ibs + a, A.n - 1);
          rowp[a] = (suint32 *)(A.planes + ia * A.W);
          if constexpr (NP) rown[a] = (suint32 *)(A.nmask + ia * A.W);
        }
      for (int64_t w0 = 0; w0 < A.W; w0 += KW) {
        const int kw = (int)min((int64_t)KW, A.W - w0);
        __syncthreads();
        stage_store<NP, TI_, SR>(lds, sNI, sNJ, tid, stg);
        __syncthreads();
        if (w0 + KW < A.W)
          stage_load<NP, TI_, SR>(A, i0, j0, w0 + KW, (int)min((int64_t)KW, A.W - w0 - KW), tid, stg);
        for (int k0 = 0; k0 < kw; k0 += RG) {
          uint4 rw[RP][RG];
          uint32_t rn[RP][RG];
#pragma unroll
          for (int a = 0; a < RP; a++)
#pragma unroll
            for (int q = 0; q < RG; q++) {
              if constexpr (SR) {
                rw[a][q] = sload4(rowp[a] + 4 * (w0 + k0 + q));
                if constexpr (NP) rn[a][q] = rown[a][w0 + k0 + q];
              }
            }
#pragma unroll
          for (int q = 0; q < RG; q++) {
            const int k = k0 + q;
            if (k >= kw) break;
            const uint4 pjv = lds.sJ[k][lane];
#pragma unroll
            for (int a = 0; a < RP; a++) count_word(SR ? rw[a][q] : lds.sI[k][wave * RP + a], pjv, cnt[a]);
            if constexpr (NP) {
              // the rows' N sites against the column (a uniform branch); the column's N sites against the rows are
              // counted per N run after the loop
              const uint32_t nj = sNJ[k][lane];
#pragma unroll
              for (int a = 0; a < RP; a++) {
                const uint32_t ni = SR ? rn[a][q] : sNI[k][wave * RP + a];  // the same in every lane
                if (ni) {
                  ncn[a].r[0] += __popc(ni & pjv.x);
                  ncn[a].r[1] += __popc(ni & pjv.y);
                  ncn[a].r[2] += __popc(ni & pjv.z);
                  ncn[a].r[3] += __popc(ni & pjv.w);
                  ncn[a].nn += __popc(ni & nj);
                }
              }
            }
          }
        }
      }
    }
  }
  if constexpr (MF) {
    static_assert(RP == 4, "the MFMA backend tiles 16 rows");
    if (do_fast) {
      // wave w: all TI_/4 row groups x column groups 4w..4w+3; one v_mfma_i32_16x16x64_i8 per (row group, column
      // group) and 64-site block gives the 4x4 base table of 16 pairs in int32
      v4i macc[TI_ / 4][4];
#pragma unroll
      for (int ig = 0; ig < TI_ / 4; ig++)
#pragma unroll
        for (int jg = 0; jg < 4; jg++) macc[ig][jg] = v4i{0, 0, 0, 0};
      // Operands of one 64-site block: the 4 row groups (1 KB each, shared by the waves) pass through LDS, double
      // buffered (groups 4 buf .. 4 buf + 3 of the operand area), one group per wave's 64 lanes; each wave's own 4
      // column groups go straight to registers.  The next block's operands are loaded while this one's MFMAs run, so
      // one barrier per block remains.  The groups lie inside the one-hot array (npad rounded past every tile).
      const uint4 *gI = reinterpret_cast<const uint4 *>(A.onehot) + (i0 >> 2) * A.W64 * 64;
      const uint4 *gJ = reinterpret_cast<const uint4 *>(A.onehot) + (j0 >> 2) * A.W64 * 64;
      uint4 ra = gI[(int64_t)wave * A.W64 * 64 + lane];
      uint4 cb[4];
#pragma unroll
      for (int jg = 0; jg < 4; jg++) cb[jg] = gJ[(int64_t)(4 * wave + jg) * A.W64 * 64 + lane];
      for (int64_t blk = 0; blk < A.W64; blk++) {
        const int buf = (int)(blk & 1);
        lds.u.ops[4 * buf + wave][lane] = ra;
        v4i bop[4];
#pragma unroll
        for (int jg = 0; jg < 4; jg++) bop[jg] = v4i{(int)cb[jg].x, (int)cb[jg].y, (int)cb[jg].z, (int)cb[jg].w};
        __syncthreads();
        if (blk + 1 < A.W64) {
          ra = gI[((int64_t)wave * A.W64 + blk + 1) * 64 + lane];
#pragma unroll
          for (int jg = 0; jg < 4; jg++) cb[jg] = gJ[((int64_t)(4 * wave + jg) * A.W64 + blk + 1) * 64 + lane];
        }
#pragma unroll
        for (int ig = 0; ig < TI_ / 4; ig++) {
          const uint4 v = lds.u.ops[4 * buf + ig][lane];
          const v4i aop = v4i{(int)v.x, (int)v.y, (int)v.z, (int)v.w};
#pragma unroll
          for (int jg = 0; jg < 4; jg++)
            macc[ig][jg] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aop, bop[jg], macc[ig][jg], 0, 0, 0);
        }
      }
      __syncthreads();
      // lane l of tile (ig, jg) holds C[a][b] for a = 0..3 of the pair (row 4ig + l/16, column 4(4w+jg) + (l%16)/4)
      // with b = l%4; reduce over the quad into the seven counts of the popcount backend
      const int b = lane & 3;
#pragma unroll
      for (int ig = 0; ig < TI_ / 4; ig++)
#pragma unroll
        for (int jg = 0; jg < 4; jg++) {
          const v4i v = macc[ig][jg];
          const uint32_t v0 = (uint32_t)v.x, v1 = (uint32_t)v.y, v2 = (uint32_t)v.z, v3 = (uint32_t)v.w;
          const uint32_t col = v0 + v1 + v2 + v3;
          const uint32_t dg = b == 0 ? v0 : b == 1 ? v1 : b == 2 ? v2 : v3;
          const uint32_t ct = b == 1 ? v3 : b == 3 ? v1 : 0u;
          const uint32_t ts = ct + (b == 0 ? v2 : b == 2 ? v0 : 0u);
          const uint32_t b0 = v1 + v3 + ((b & 1) ? col : 0u);
          const uint32_t b1 = v2 + v3 + ((b >> 1) ? col : 0u);
          const uint32_t bt = v3 + (b == 3 ? col : 0u);
          const uint32_t q0 = quad_sum(col), q1 = quad_sum(dg), q2 = quad_sum(ts), q3 = quad_sum(ct);
          const uint32_t q4 = quad_sum(b0), q5 = quad_sum(b1), q6 = quad_sum(bt);
          const int il = 4 * ig + (lane >> 4), jloc = 4 * (4 * wave + jg) + ((lane & 15) >> 2);
          // lane b of the quad stores fields b and b + 4
          lds.u.agg[b][il][jloc] = b == 0 ? q0 : b == 1 ? q1 : b == 2 ? q2 : q3;
          if (b < 3) lds.u.agg[b + 4][il][jloc] = b == 0 ? q4 : b == 1 ? q5 : q6;
        }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < RP; a++) {
        const int il = wave * RP + a;
        cnt[a] = Cnt{lds.u.agg[0][il][lane], lds.u.agg[1][il][lane], lds.u.agg[2][il][lane], lds.u.agg[3][il][lane],
                     lds.u.agg[4][il][lane], lds.u.agg[5][il][lane], lds.u.agg[6][il][lane]};
      }
    }
  }
  __syncthreads();  // the event tables and lists are in LDS (also when no counts ran); the counting buffers are free
  if (do_events) chunk_codes<TI_>(A, te, sEv, L.g, i0, j0, tid);
  if constexpr (MARG)
    if (do_fast) base_sums<RP>(A, sTB, tr, sRg, ib, wave, lane, j, h0, h1, S);
  if constexpr (NP && (MF || GC))
    if (do_fast && j < A.n) {
      // MFMA backend: the one-hot operands carry clean sites only, so the rows' N sites against the column (its
      // planes and its N sites) come from the rows' N runs, each a difference of the column's prefix counts
#pragma unroll
      for (int a = 0; a < RP; a++) {
        const int64_t i = ib + a;
        if (i >= A.r1) continue;  // uniform over the wave
        const int q1 = A.nrun_off[i + 1];
        for (int q = A.nrun_off[i]; q < q1; q++) {
          const uint2 r = A.nruns[q];
          const uint4 hi = plane_prefix(A.pre, A.W, j, r.y);
          const uint4 lo = plane_prefix(A.pre, A.W, j, r.x);
          ncn[a].r[0] += hi.x - lo.x;
          ncn[a].r[1] += hi.y - lo.y;
          ncn[a].r[2] += hi.z - lo.z;
          ncn[a].r[3] += hi.w - lo.w;
          ncn[a].nn += n_prefix(A.npre, A.nmask, A.W, j, r.y) - n_prefix(A.npre, A.nmask, A.W, j, r.x);
        }
      }
    }
  if constexpr (NP)
    if (do_fast && j < A.n) {
      // the column's N sites against the rows' planes: N comes in runs, so each run is a difference of the rows'
      // plane prefix counts (clean, C|T, G|T, T) instead of 8 operations per word and row
      const int q1 = A.nrun_off[j + 1];
      for (int q = A.nrun_off[j]; q < q1; q++) {
        const uint2 r = A.nruns[q];
#pragma unroll
        for (int a = 0; a < RP; a++) {
          if (ib + a >= A.r1) continue;  // uniform over the wave
          const uint4 hi = plane_prefix(A.pre, A.W, ib + a, r.y);
          const uint4 lo = plane_prefix(A.pre, A.W, ib + a, r.x);
          ncn[a].c[0] += hi.x - lo.x;
          ncn[a].c[1] += hi.y - lo.y;
          ncn[a].c[2] += hi.z - lo.z;
          ncn[a].c[3] += hi.w - lo.w;
        }
      }
    }
  __syncthreads();
  bool all_rows[RP];
#pragma unroll
  for (int a = 0; a < RP; a++) all_rows[a] = true;
  if (do_events) tile_events<false, RP>(A, sTabR, sTabC, te, sEv, L.g, sRo, i0, j0, ib, tid, pj, pj_end, all_rows, acc, tot);
  if constexpr (NP)
    if (do_events)
#pragma unroll
      for (int a = 0; a < RP; a++) nplane_add(ncn[a], sTabR, tot[a], acc[a]);
  // -g: pairs whose resolved ambiguities reach the fraction are computed again in AVERAGE (refkernel.cpp:742-750):
  // same clean/clean counts, the events through the AVERAGE table.  Order-dependent ones go to the replay instead.
  bool avg[RP], avgr[RP];  // avgr: order-dependent pairs whose -g test is decided for AVERAGE
#pragma unroll
  for (int a = 0; a < RP; a++) avg[a] = avgr[a] = false;
  bool any_avg = false;
  if (A.avg_fallback && do_events) {
#pragma unroll
    for (int a = 0; a < RP; a++) {
      const int64_t i = ib + a;
      const bool valid = i < A.r1 && j >= A.c0 && j < A.c1 && (!A.triangle || j > i);
      uint32_t nfs = tot[a].nf[0] + tot[a].nf[1] + tot[a].nf[2] + tot[a].nf[3];
      if constexpr (MARG)
        nfs += 16u * S[a].x;
      else
        nfs += 32u * cnt[a].T;
      nfs += 32u * (uint32_t)A.kt;
      const double tng = (double)nfs * 0.03125;
      avg[a] = valid && !acc[a].fl && tng > (double)A.min_overlap && tng * A.rf <= (double)tot[a].amb;
      any_avg |= avg[a];
      // an order-dependent pair: tng is the exact count of the sites tn93 sums in FP64, whose rounded sum stays within
      // (16 L^2 + 64 L) 2^-53 of it (16 cells of at most L additions of values below L, v rounded, 32 merge
      // additions); when tng is farther than that above the overlap and tng rf farther than that from amb, tn93's
      // -g test is decided: true => the replay starts in AVERAGE (margins doubled, the product's rounding included)
      if (valid && acc[a].fl && A.gcert) {
        const double E = A.gE;  // twice the bound on |tng_fp - tng|
        const double M = E * A.rf + 2.0 * tng * A.rf * 0x1p-52;
        avgr[a] = tng - E > (double)A.min_overlap && tng * A.rf + M < (double)tot[a].amb;
      }
    }
  }
  if (__syncthreads_or(any_avg)) {
    load_event_tables(tabs + NCODE * NCODE, sTabR, sTabC, tid, A.gapmm, A.nplane);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < RP; a++)
      if (avg[a]) {
        tot[a] = Tot{{0, 0, 0, 0}, 0, 0, 0, 0};
        acc[a] = EvAcc{0, 0, 0};
      }
    if (!te.single()) {  // the tile's first chunk again
      te.first();
      chunk_events(A, te, sEv, tid);
      __syncthreads();
      chunk_codes<TI_>(A, te, sEv, L.g, i0, j0, tid);
      __syncthreads();
    }
    tile_events<true, RP>(A, sTabR, sTabC, te, sEv, L.g, sRo, i0, j0, ib, tid, pj, pj_end, avg, acc, tot);
    if constexpr (NP)
#pragma unroll
      for (int a = 0; a < RP; a++)
        if (avg[a]) nplane_add(ncn[a], sTabR, tot[a], acc[a]);
  }
  // totals x16 from the clean/clean counts, plus the events
#pragma unroll
  for (int a = 0; a < RP; a++) {
    const CntT &c = cnt[a];
    uint32_t nf0, nf1, nf2, nf3;
    if constexpr (MARG) {
      const uint4 &q = S[a];  // clean, C|T, G|T, T summed over both sequences
      nf3 = q.w;
      nf1 = q.y - q.w;
      nf2 = q.z - q.w;
      nf0 = q.x - q.y - q.z + q.w;
    } else {
      nf3 = c.BT;
      nf1 = c.B0 - c.BT;
      nf2 = c.B1 - c.BT;
      nf0 = 2 * c.T - c.B0 - c.B1 + c.BT;
    }
    tot[a].nf[0] += 16 * nf0 + 32 * A.kc[0];
    tot[a].nf[1] += 16 * nf1 + 32 * A.kc[1];
    tot[a].nf[2] += 16 * nf2 + 32 * A.kc[2];
    tot[a].nf[3] += 16 * nf3 + 32 * A.kc[3];
    tot[a].dg += 16 * (c.DG + (uint32_t)A.kt);
    tot[a].ag += 16 * (c.TS - c.CT);
    tot[a].ct += 16 * c.CT;
  }
  // per pair: tests and routing.  Fast path: the tail inputs go to k_tail in one chunk per wave-row; exact path:
  // the tail right here, into D/U.  Pairs for the replay are marked in smask (bit a: row ib + a; bit 4 + a: the -g
  // test already chose AVERAGE) and appended after the loop.
  unsigned smask = 0;
#pragma unroll
  for (int a = 0; a < RP; a++) {
    const int64_t i = ib + a;
    const bool valid = i < A.r1 && j >= A.c0 && j < A.c1 && (!A.triangle || j > i);
    int state = 0;  // 1: tail, 2: overlap too small
    const Tot &o = tot[a];
    if (valid) {
      bool slow = !do_fast || acc[a].fl;  // order-dependent: exact site-order replay
      if (!slow) {
        const double tng = (double)(o.nf[0] + o.nf[1] + o.nf[2] + o.nf[3]) * 0.03125;
        if (tng <= (double)A.min_overlap)  // refkernel.cpp:732-735
          state = 2;
        else if (A.avg_fallback && !do_events && tng * A.rf <= (double)o.amb)  // refkernel.cpp:742-750
          slow = true;
        else
          state = 1;
      }
      if (slow) smask |= (1u << a) | (avg[a] || avgr[a] ? 1u << (RP + a) : 0u);
    }
    if constexpr (FAST) {
      const unsigned long long need = __ballot(state == 1), skp = __ballot(state == 2);
      if (i < A.r1) {
        const size_t slot = (size_t)(i - A.r0) * (size_t)A.gx + (size_t)bx;
        if (lane == 0) A.chdr[slot] = ChunkHdr{need, skp};
        if (state == 1) {
          uint32_t *d = A.cdata + slot * A.cstride + lane;
          if (A.c16) {  // every value below 2^16
            d[0] = o.nf[0] | o.nf[1] << 16;
            d[64] = o.nf[2] | o.nf[3] << 16;
            d[128] = o.dg | o.ag << 16;
            d[192] = o.ct;
          } else {
            d[0] = o.nf[0];
            d[64] = o.nf[1];
            d[128] = o.nf[2];
            d[192] = o.nf[3];
            d[256] = o.dg;
            d[320] = o.ag;
            d[384] = o.ct;
          }
        }
      }
    } else if (state) {
      RowPart rp = rowpart_zero();  // unused: the exact path writes D/U
      if (state == 2) {
        skip_pair(A, i, j, rp);
      } else {
        tn93g_tailin tin;
        for (int k = 0; k < 4; k++) tin.nf[k] = (double)o.nf[k] * 0.0625;
        tin.diag = (double)o.dg * 0.0625;
        tin.ag = (double)o.ag * 0.0625;
        tin.ct = (double)o.ct * 0.0625;
        finish_pair<false>(A, i, j, tin, lhist, &bs, rp);
      }
    }
  }
  // the tile's replay pairs: a slot range per thread in LDS, one range of the flat list per tile
  const unsigned nmine = __popc(smask & ((1u << RP) - 1));
  const unsigned k0 = nmine ? atomicAdd(&lslow, nmine) : 0u;  // lslow <= TI * TJ
  flush_block(A, &bs, FAST ? nullptr : lhist, tid, 256);  // begins with a barrier: lslow is final
  if (tid == 0 && lslow) sbase = atomicAdd(A.slow_total, (unsigned long long)lslow);
  __syncthreads();
  if (nmine) {
    unsigned long long *dst = A.slow_list + sbase + k0;
    unsigned q = 0;
#pragma unroll
    for (int a = 0; a < RP; a++)
      if (smask >> a & 1u)  // bit 63: the replay starts in AVERAGE
        dst[q++] = (unsigned long long)(ib + a) << 32 | (unsigned long long)j |
                   (smask >> (RP + a) & 1u ? 1ull << 63 : 0ull);
  }
}

template <int BACKEND, bool FAST, bool NP, bool MARG, int RP>
__global__ __launch_bounds__(256) void k_pairs(RunArgs A)
{
  pairs_tile<BACKEND, FAST, NP, MARG, RP>(A, blockIdx.x, blockIdx.y);
}

// The pair tiles of several jobs of one kernel variant in one launch (a batch): block b is tile b - toff[k] of job k,
// toff[k] <= b < toff[k + 1], so small alignments fill the device together instead of one after another.
template <int BACKEND, bool FAST, bool NP, bool MARG, int RP>
__global__ __launch_bounds__(256) void k_pairs_multi(const RunArgs *__restrict__ jobs, const int64_t *__restrict__ toff,
                                                     int njobs)
{
  const int64_t b = blockIdx.x;
  int lo = 0, hi = njobs;  // uniform over the block
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (toff[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  const RunArgs &A = jobs[lo];
  const int64_t t = b - toff[lo];
  pairs_tile<BACKEND, FAST, NP, MARG, RP>(A, t % A.gx, t / A.gx);
}

// The FP64 tail of the fast path.  A block takes whole rows; its 4 waves take the row's chunks (64 columns each) in
// turn, at high occupancy so the long dependent chains of divisions and logarithms of many pairs overlap.  Each lane
// keeps its share of the row's reductions over the chunks; the block combines them into one set of atomics per row.
// C16: the chunks hold 16-bit packed inputs (4 words per lane), else 7 words.  A chunk's header and inputs are loaded
// one chunk ahead and no other load sits between, so the wait for them never covers the next chunk's loads.
struct TailLds {
  unsigned long long lhist[3 * 200];
  BlockStats bs;
  RowPart wpart[4];
  double log[TN93G_LOGTAB_N + 1][4];  // the logarithm's table: LDS look-ups instead of memory ones
};
template <bool C16>
__device__ __attribute__((always_inline)) inline void tail_rows(const RunArgs &A, int64_t rbeg, int64_t rstep,
                                                                TailLds &S)
{
  constexpr int NQ = C16 ? 4 : 7;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int k = tid; k < 600; k += 256) S.lhist[k] = 0;
  for (int k = tid; k < (TN93G_LOGTAB_N + 1) * 4; k += 256) S.log[k >> 2][k & 3] = tn93g_logtab[k >> 2][k & 3];
  if (tid == 0) S.bs.unc = S.bs.undecided = 0;
  __syncthreads();
  const int64_t nrows = A.r1 - A.r0;
  const int64_t c0a = A.c0 & ~(int64_t)63;
  for (int64_t r = rbeg; r < nrows; r += rstep) {
    const int64_t i = A.r0 + r;
    RowPart rp = rowpart_zero();
    // the wave's chunks of the row, bx = wave, wave + 4, ...
    const int64_t tri = A.r0 + (r / A.ti) * A.ti;  // first row of the pair kernel's tile row
    auto live = [&](int64_t bx) { return bx < A.gx && !(A.triangle && c0a + bx * TJ + TJ - 1 <= tri); };
    auto load = [&](int64_t bx, ChunkHdr &h, uint32_t (&v)[NQ]) {
      const int64_t k = r * A.gx + bx;
      h = A.chdr[k];
      const uint32_t *d = A.cdata + (size_t)k * A.cstride + lane;
#pragma unroll
      for (int f = 0; f < NQ; f++) v[f] = d[64 * f];
    };
    int64_t bx = wave;
    while (bx < A.gx && !live(bx)) bx += 4;
    ChunkHdr h{0, 0};
    uint32_t p[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) p[f] = 0;
    if (bx < A.gx) load(bx, h, p);
    while (bx < A.gx) {
      int64_t nb = bx + 4;
      while (nb < A.gx && !live(nb)) nb += 4;
      ChunkHdr hn{0, 0};
      uint32_t n[NQ];
#pragma unroll
      for (int f = 0; f < NQ; f++) n[f] = 0;
      if (nb < A.gx) load(nb, hn, n);
      const int64_t j = c0a + bx * TJ + lane;
      if (h.need >> lane & 1ull) {
        uint32_t q[7];
        if constexpr (C16) {
          q[0] = p[0] & 0xffffu, q[1] = p[0] >> 16, q[2] = p[1] & 0xffffu, q[3] = p[1] >> 16;
          q[4] = p[2] & 0xffffu, q[5] = p[2] >> 16, q[6] = p[3];
        } else {
#pragma unroll
          for (int f = 0; f < 7; f++) q[f] = p[f];
        }
        // the x16 counts as they are: the tail only uses ratios of its inputs (1 / sum, then products with it), and
        // scaling every input by 16 scales those intermediates by exact powers of two, so every rounded value after
        // them, and the distance, are the same bits as with the counts divided by 16
        tn93g_tailin tin;
        tin.nf[0] = (double)q[0];
        tin.nf[1] = (double)q[1];
        tin.nf[2] = (double)q[2];
        tin.nf[3] = (double)q[3];
        tin.diag = (double)q[4];
        tin.ag = (double)q[5];
        tin.ct = (double)q[6];
        finish_pair<true>(A, i, j, tin, S.lhist, &S.bs, rp, S.log);
      } else if (h.skip >> lane & 1ull) {
        skip_pair(A, i, j, rp);
      }
      bx = nb;
      h = hn;
#pragma unroll
      for (int f = 0; f < NQ; f++) p[f] = n[f];
    }
    RowPart q;
    q.slo = wave_sum_d(rp.slo);
    q.shi = wave_sum_d(rp.shi);
    q.sabs = wave_sum_d(rp.sabs);
    q.mlo = wave_max_d(rp.mlo);
    q.mhi = wave_max_d(rp.mhi);
    q.wsum = wave_sum_l(rp.wsum);
    q.skip = wave_sum_l(rp.skip);
    q.linkw = wave_sum_l(rp.linkw);
    if (lane == 0) S.wpart[wave] = q;
    __syncthreads();
    if (tid == 0) {
      RowPart s = S.wpart[0];
      for (int w = 1; w < 4; w++) {
        s.slo += S.wpart[w].slo;
        s.shi += S.wpart[w].shi;
        s.sabs += S.wpart[w].sabs;
        s.mlo = fmax(s.mlo, S.wpart[w].mlo);
        s.mhi = fmax(s.mhi, S.wpart[w].mhi);
        s.wsum += S.wpart[w].wsum;
        s.skip += S.wpart[w].skip;
        s.linkw += S.wpart[w].linkw;
      }
      rowacc_add(A, i, s);
    }
    __syncthreads();
  }
  flush_block(A, &S.bs, S.lhist, tid, 256);
}

template <bool C16>
__global__ __launch_bounds__(256, 4) void k_tail(RunArgs A)
{
  __shared__ TailLds S;
  tail_rows<C16>(A, blockIdx.x, gridDim.x, S);
}

// The block of flat index b in a list of jobs: the largest k with off[k] <= b (uniform over the block).
__device__ inline int job_of(const int64_t *__restrict__ off, int nj, int64_t b)
{
  int lo = 0, hi = nj;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// The tails of several jobs (a batch) in one launch: job k owns blocks [boff[k], boff[k + 1]), which stride over
// its rows as k_tail's grid does.
__global__ __launch_bounds__(256, 4) void k_tail_multi(const RunArgs *__restrict__ jobs, const int64_t *__restrict__ boff,
                                                    int nj)
{
  __shared__ TailLds S;
  const int k = job_of(boff, nj, blockIdx.x);
  if (jobs[k].c16)
    tail_rows<true>(jobs[k], (int64_t)blockIdx.x - boff[k], boff[k + 1] - boff[k], S);
  else
    tail_rows<false>(jobs[k], (int64_t)blockIdx.x - boff[k], boff[k + 1] - boff[k], S);
}

// ------------------------------------------------------------------------------------------ replay kernel
// computeTN93's scan (refkernel.cpp:384-729) for one pair per thread, in site order: clean/clean sites counted in
// integers, every other site applied through the handler table in FP64 in increasing site order, then tn93's merge.
// The pair's 16 float cells, 16 clean/clean counts and 4 resolved-diagonal counts live in registers; a site's effect
// is added to every cell, 0.0 where its mask is clear (the cells hold non-negative sums, so x + 0.0 == x).
constexpr int RB = 256;

struct ReplayState {
  double fc[16];
  uint32_t dgi[4];  // integer_counts[d][d] increments of resolved ambiguities
  uint32_t amb;
};

__device__ inline void replay_reset(ReplayState &s)
{
#pragma unroll
  for (int k = 0; k < 16; k++) s.fc[k] = 0.0;
#pragma unroll
  for (int d = 0; d < 4; d++) s.dgi[d] = 0;
  s.amb = 0;
}

__device__ inline void replay_apply(ReplayState &s, const ReplayEnt &E)
{
  const double v = E.v;
#pragma unroll
  for (int k = 0; k < 16; k++) s.fc[k] += (E.mask >> k & 1u) ? v : 0.0;
#pragma unroll
  for (int d = 0; d < 4; d++) s.dgi[d] += E.icd == d ? 1u : 0u;
  s.amb += E.amb;
}

// tn93's merge (refkernel.cpp:714-729): pc = float + integer counts, summed in row-major cell order
__device__ inline void replay_merge(const ReplayState &s, const uint32_t (&icc)[16], tn93g_tailin &tin, double &tng)
{
  double pc[16];
#pragma unroll
  for (int k = 0; k < 16; k++) pc[k] = s.fc[k] + (double)(icc[k] + (k % 5 == 0 ? s.dgi[k / 5] : 0u));
  double nf[4] = {0., 0., 0., 0.};
  tng = 0.;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      tng += pc[4 * a + b];
      nf[a] += pc[4 * a + b];
      nf[b] += pc[4 * a + b];
    }
#pragma unroll
  for (int k = 0; k < 4; k++) tin.nf[k] = nf[k];
  tin.diag = pc[0] + pc[5] + pc[10] + pc[15];
  tin.ag = pc[2] + pc[8];
  tin.ct = pc[7] + pc[13];
}

// The clean/clean 4x4 integer table of a pair from the bit planes (integer_counts before any resolved ambiguity).
__device__ inline void cc_table(const RunArgs &A, int64_t i, int64_t j, uint32_t (&c)[16])
{
#pragma unroll
  for (int k = 0; k < 16; k++) c[k] = 0;
  const uint4 *Pi = A.planes + i * A.W, *Pj = A.planes + j * A.W;
#pragma unroll 4
  for (int64_t w = 0; w < A.W; w++) {
    const uint4 pi = Pi[w], pj = Pj[w];
    const uint32_t M = pi.x & pj.x;
    const uint32_t ai[4] = {M & ~(pi.y | pi.z), M & pi.y & ~pi.z, M & pi.z & ~pi.y, M & pi.w};
    const uint32_t bj[4] = {~(pj.y | pj.z), pj.y & ~pj.z, pj.z & ~pj.y, pj.w};
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) c[4 * x + y] += __popc(ai[x] & bj[y]);
  }
}

// The same table for the 64 pairs of a wave at once (long alignments): the wave takes the pairs one after the other,
// its lanes split the words (coalesced loads of both rows) and the 16 counts are summed over the wave into
// out[pair][16].  Lanes with valid == false bring no pair.  Every lane of the wave must call it.
__device__ void cc_table_wave(const RunArgs &A, int64_t i, int64_t j, bool valid, uint32_t (*out)[16])
{
  const int lane = threadIdx.x & 63;
  unsigned long long m = __ballot(valid);
  while (m) {
    const int src = __ffsll((long long)m) - 1;
    m &= m - 1;
    const int64_t pi = __shfl(i, src, 64), pj = __shfl(j, src, 64);
    uint32_t c[16];
#pragma unroll
    for (int k = 0; k < 16; k++) c[k] = 0;
    const uint4 *Pi = A.planes + pi * A.W, *Pj = A.planes + pj * A.W;
    for (int64_t w = lane; w < A.W; w += 64) {
      const uint4 a = Pi[w], b = Pj[w];
      const uint32_t M = a.x & b.x;
      const uint32_t ai[4] = {M & ~(a.y | a.z), M & a.y & ~a.z, M & a.z & ~a.y, M & a.w};
      const uint32_t bj[4] = {~(b.y | b.z), b.y & ~b.z, b.z & ~b.y, b.w};
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) c[4 * x + y] += __popc(ai[x] & bj[y]);
    }
    // butterfly reduce-scatter of the 16 counters over the wave: at the step over lane bit b each lane keeps half of
    // its counters, adds its partner's copy of that half and sends the other half (8 + 4 + 2 + 1 exchanges), so
    // that lane l holds counter l >> 2 summed over the lanes that differ in bits 2..5; two more exchanges add the
    // lanes that differ in bits 0..1.  17 exchanges instead of 16 x 6.
    const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4;
    uint32_t c8[8], c4[4], c2[2];
#pragma unroll
    for (int q = 0; q < 8; q++)
      c8[q] = (h5 ? c[q + 8] : c[q]) + (uint32_t)__shfl_xor((int)(h5 ? c[q] : c[q + 8]), 32, 64);
#pragma unroll
    for (int q = 0; q < 4; q++)
      c4[q] = (h4 ? c8[q + 4] : c8[q]) + (uint32_t)__shfl_xor((int)(h4 ? c8[q] : c8[q + 4]), 16, 64);
#pragma unroll
    for (int q = 0; q < 2; q++)
      c2[q] = (h3 ? c4[q + 2] : c4[q]) + (uint32_t)__shfl_xor((int)(h3 ? c4[q] : c4[q + 2]), 8, 64);
    uint32_t v = (h2 ? c2[1] : c2[0]) + (uint32_t)__shfl_xor((int)(h2 ? c2[0] : c2[1]), 4, 64);
    v += (uint32_t)__shfl_xor((int)v, 2, 64);
    v += (uint32_t)__shfl_xor((int)v, 1, 64);
    if ((lane & 3) == 0) out[src][lane >> 2] = v;
  }
}

// Clean/clean 4x4 count tables of `ns` special sequences ids[0..ns) against the sequences [x0, x0 + nx):
// T[(s * 16 + 4a + b) * nx + x - x0] = sites where sequence ids[s] has clean base a and sequence x clean base b.  A
// block takes 16 special sequences x 64 others (the pair kernel's tile); a wave's lanes are the 64 others, each with
// the 16 counters of the wave's 4 special sequences.  The others' planes are staged through LDS KW words at a time;
// the special sequences' words are the same in every lane of a wave and are read with scalar loads, two words at a
// time (words past a sequence's end read the next one or the planes' padding and are not used).
__global__ __launch_bounds__(256) void k_spec16(RunArgs A, const int64_t *__restrict__ ids, int ns, int64_t x0,
                                                int64_t nx, uint32_t *__restrict__ T)
{
  __shared__ uint4 sJ[KW][TJ + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int sb = blockIdx.y * TI;
  const int64_t xb = (int64_t)blockIdx.x * TJ;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  suint32 *rowp[RPW];
#pragma unroll
  for (int a = 0; a < RPW; a++) {
    const int sp = min(sb + wv * RPW + a, ns - 1);  // past the list: its last sequence (counts not stored)
    rowp[a] = (suint32 *)(A.planes + ids[sp] * A.W);
  }
  uint32_t c[RPW][16];
#pragma unroll
  for (int a = 0; a < RPW; a++)
#pragma unroll
    for (int k = 0; k < 16; k++) c[a][k] = 0;
  for (int64_t w0 = 0; w0 < A.W; w0 += KW) {
    const int kw = (int)min((int64_t)KW, A.W - w0);
    __syncthreads();
    for (int k = tid; k < KW * TJ; k += 256) {
      const int q = k >> 4, w = k & (KW - 1);
      const int64_t seq = x0 + xb + q;
      sJ[w][q] = w < kw && xb + q < nx ? A.planes[seq * A.W + w0 + w] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    for (int k0 = 0; k0 < kw; k0 += 2) {
      uint4 rw[RPW][2];
#pragma unroll
      for (int a = 0; a < RPW; a++)
#pragma unroll
        for (int q = 0; q < 2; q++) rw[a][q] = sload4(rowp[a] + 4 * (w0 + k0 + q));
#pragma unroll
      for (int q = 0; q < 2; q++) {
        if (k0 + q >= kw) break;
        const uint4 b = sJ[k0 + q][lane];
        const uint32_t bj[4] = {~(b.y | b.z), b.y & ~b.z, b.z & ~b.y, b.w};
#pragma unroll
        for (int a = 0; a < RPW; a++) {
          const uint4 p = rw[a][q];
          const uint32_t M = p.x & b.x;
          const uint32_t ai[4] = {M & ~(p.y | p.z), M & p.y & ~p.z, M & p.z & ~p.y, M & p.w};
#pragma unroll
          for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) c[a][4 * x + y] += __popc(ai[x] & bj[y]);
        }
      }
    }
  }
  const int64_t xl = xb + lane;
  if (xl < nx)
#pragma unroll
    for (int a = 0; a < RPW; a++) {
      const int sp = sb + wave * RPW + a;
      if (sp < ns)
#pragma unroll
        for (int x = 0; x < 4; x++)
          *reinterpret_cast<uint4 *>(T + (((size_t)sp * 4 + x) * (size_t)nx + (size_t)xl) * 4) =
              make_uint4(c[a][4 * x], c[a][4 * x + 1], c[a][4 * x + 2], c[a][4 * x + 3]);
    }
}

// Operands of the GEMM counts for sequences s0 .. s0 + nq (one thread per 16 sites of one sequence): X1 row q =
// [A indicator | G indicator] (2 Ls bytes), X2 = [C | T], X3 = purine (A or G), X4 = pyrimidine (C or T); 0 at
// every other code.  X1.X1 = AA + GG, X2.X2 = CC + TT, X3.X3 = RR, X4.X4 = YY over the clean/clean sites.
__global__ __launch_bounds__(256) void k_onehot_counts(const uint8_t *__restrict__ codes, int64_t Ls, int64_t s0,
                                                       int64_t nq, int8_t *__restrict__ X1, int8_t *__restrict__ X2,
                                                       int8_t *__restrict__ X3, int8_t *__restrict__ X4)
{
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x, chunks = Ls / 16;
  if (t >= nq * chunks) return;
  const int64_t q = t / chunks, ch = t - q * chunks;
  const uint4 v = *reinterpret_cast<const uint4 *>(codes + (s0 + q) * Ls + ch * 16);
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  uint32_t ind[4][4];  // [base][word]: byte y of word k is 1 where site 4k + y holds that base
#pragma unroll
  for (uint32_t b = 0; b < 4; b++)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint32_t r = 0;
#pragma unroll
      for (int y = 0; y < 4; y++) r |= ((w[k] >> (8 * y) & 255u) == b ? 1u : 0u) << (8 * y);
      ind[b][k] = r;
    }
  auto put = [&](int8_t *row, const uint32_t(&x)[4]) {
    *reinterpret_cast<uint4 *>(row + ch * 16) = make_uint4(x[0], x[1], x[2], x[3]);
  };
  const uint32_t pur[4] = {ind[0][0] | ind[2][0], ind[0][1] | ind[2][1], ind[0][2] | ind[2][2], ind[0][3] | ind[2][3]};
  const uint32_t pyr[4] = {ind[1][0] | ind[3][0], ind[1][1] | ind[3][1], ind[1][2] | ind[3][2], ind[1][3] | ind[3][3]};
  put(X1 + q * 2 * Ls, ind[0]);
  put(X1 + q * 2 * Ls + Ls, ind[2]);
  put(X2 + q * 2 * Ls, ind[1]);
  put(X2 + q * 2 * Ls + Ls, ind[3]);
  put(X3 + q * Ls, pur);
  put(X4 + q * Ls, pyr);
}

// One-hot rows of sequences for the special-sequence tables as an int8 GEMM: out[(q * 4 + b) * Kp + p] = 1 where
// sequence seq(q) (ids[q], or x0 + q without ids) holds the clean base b at site p, else 0 (ambiguity codes, N, gaps
// and the padding of the codes rows past L).  One thread per 16 sites of one sequence.
__global__ __launch_bounds__(256) void k_onehot_rows(const uint8_t *__restrict__ codes, int64_t Ls,
                                                     const int64_t *__restrict__ ids, int64_t x0, int64_t nq,
                                                     int8_t *__restrict__ out)
{
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x, chunks = Ls / 16;
  if (t >= nq * chunks) return;
  const int64_t q = t / chunks, ch = t - q * chunks;
  const int64_t seq = ids ? ids[q] : x0 + q;
  const uint4 v = *reinterpret_cast<const uint4 *>(codes + seq * Ls + ch * 16);
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (uint32_t b = 0; b < 4; b++) {
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint32_t r = 0;
#pragma unroll
      for (int y = 0; y < 4; y++) r |= ((w[k] >> (8 * y) & 255u) == b ? 1u : 0u) << (8 * y);
      o[k] = r;
    }
    *reinterpret_cast<uint4 *>(out + (q * 4 + b) * Ls + ch * 16) = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// The clean/clean table of pair (i, j) from the special-sequence tables when i or j is special.
__device__ inline bool cc_table_lookup(const RunArgs &A, int64_t i, int64_t j, uint32_t (&icc)[16])
{
  if (!A.spec_row) return false;
  const int32_t si = A.spec_row[i - A.r0];
  if (si >= 0) {  // icc[4a + b] = [si][a][j][b]
    const uint32_t *t = A.tab_row + ((size_t)si * 4 * (size_t)A.ncols + (size_t)(j - A.c0)) * 4;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const uint4 v = *reinterpret_cast<const uint4 *>(t + (size_t)a * (size_t)A.ncols * 4);
      icc[4 * a] = v.x, icc[4 * a + 1] = v.y, icc[4 * a + 2] = v.z, icc[4 * a + 3] = v.w;
    }
    return true;
  }
  const int32_t sj = A.spec_col[j - A.c0];
  if (sj >= 0) {  // icc[4a + b] = [sj][b][i][a] (the special sequence is the column)
    const int64_t nr = A.r1 - A.r0;
    const uint32_t *t = A.tab_col + ((size_t)sj * 4 * (size_t)nr + (size_t)(i - A.r0)) * 4;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const uint4 v = *reinterpret_cast<const uint4 *>(t + (size_t)b * (size_t)nr * 4);
      icc[b] = v.x, icc[4 + b] = v.y, icc[8 + b] = v.z, icc[12 + b] = v.w;
    }
    return true;
  }
  return false;
}

// One sequence's event list read in site order for the replay, with the partner's code at each site.  The loads run
// ahead of their use: the entry two places on and the partner's code one place on are in flight while the current
// site is applied, so no step waits for a load issued in the same step.  Past the end the entries read NONE, whose
// site (NONE >> 5) lies after every site of an alignment (L < 2^26).
struct EvStream {
  static constexpr uint32_t NONE = 0xffffffffu;
  uint32_t w0, w1, w2;  // entries k, k + 1, k + 2 (site << 5 | code)
  uint32_t c0, c1;      // the partner's codes at the sites of w0 and w1
  int k, e;
  __device__ inline void init(const uint32_t *ev, int k_, int e_, const uint8_t *partner)
  {
    k = k_;
    e = e_;
    w0 = k < e ? ev[k] : NONE;
    w1 = k + 1 < e ? ev[k + 1] : NONE;
    w2 = k + 2 < e ? ev[k + 2] : NONE;
    c0 = w0 != NONE ? partner[w0 >> 5] : 0u;
    c1 = w1 != NONE ? partner[w1 >> 5] : 0u;
  }
  __device__ inline void next(const uint32_t *ev, const uint8_t *partner)
  {
    w0 = w1;
    c0 = c1;
    w1 = w2;
    c1 = w1 != NONE ? partner[w1 >> 5] : 0u;
    k++;
    w2 = k + 2 < e ? ev[k + 2] : NONE;
  }
};

// The sites where either sequence is ambiguous (under GAPMM also gapped: its lists hold the gap sites), in
// increasing site order: the two sorted event lists merged.  A site in one list only takes the partner's code from
// the codes array; at a site in both lists that array holds the partner's own event code, so either list's entry
// gives both codes.
__device__ void replay_events(const RunArgs &A, int64_t i, int64_t j, const ReplayEnt *tab, ReplayState &s)
{
  const uint8_t *si = A.codes + i * A.Ls, *sj = A.codes + j * A.Ls;
  EvStream a, b;
  a.init(A.ev, A.ev_off[i], A.ev_off[i + 1], sj);
  b.init(A.ev, A.ev_off[j], A.ev_off[j + 1], si);
  while (a.w0 != EvStream::NONE || b.w0 != EvStream::NONE) {
    const uint32_t pa = a.w0 >> 5, pb = b.w0 >> 5;
    const bool ta = pa <= pb, tb = pb <= pa;
    const unsigned c1 = ta ? (a.w0 & 31u) : b.c0, c2 = ta ? a.c0 : (b.w0 & 31u);
    if (ta) a.next(A.ev, sj);
    if (tb) b.next(A.ev, si);
    replay_apply(s, tab[c1 * NCODE + c2]);
  }
}

// m additions of v to a cell (tn93g_add_repeat): the cases that come to one FP64 addition (every partial sum exact,
// or one or two powers of two passed) inline, the rest (three or more powers passed, or x < v) a call: the crossing
// loop inlined at every cell spills the replay's registers.
__device__ __noinline__ double add_repeat_slow(double x, double v, double inv, uint32_t m)
{
  return tn93g_add_repeat(x, v, inv, m);
}
__device__ inline double add_repeat(double x, double v, double inv, uint32_t m)
{
  double y;
  if (tn93g_add_repeat_exact(x, v, inv, m, &y) || (x >= v && tn93g_add_repeat_two(x, v, inv, m, &y))) return y;
  return add_repeat_slow(x, v, inv, m);
}

// 1 / v for a normal power of two v = 2^e: the exponent field 2046 - (e + 1023)
__device__ inline double pow2_inv(double v) { return tn93g_from_bits((2046ull - ((tn93g_bits(v) >> 52) & 0x7ffu)) << 52); }

__device__ inline void replay_apply_rep(ReplayState &s, const ReplayEnt &E, uint32_t m)
{
  if (!m) return;
  if (E.mask) {
    const double inv = pow2_inv(E.v);
#pragma unroll
    for (int k = 0; k < 16; k++)
      if (E.mask >> k & 1u) s.fc[k] = add_repeat(s.fc[k], E.v, inv, m);
  }
#pragma unroll
  for (int d = 0; d < 4; d++) s.dgi[d] += E.icd == d ? m : 0u;
  s.amb += E.amb * m;
}

// A segment of sites where one sequence is in an N run and the other holds cnt[x] sites of clean base x.  by_col:
// the row sequence is in N, and entry (N, x) marks cells of column x only; else the column sequence is, and (x, N)
// marks cells of row x only (checked when the tables are built).  So each cell takes at most one entry's run, cell
// (a, b) the run of (N, b) or of (a, N), and the 16 cells are one pass.
__device__ inline void replay_apply_seg(ReplayState &s, const ReplayEnt *tab, const uint32_t (&cnt)[4], bool by_col)
{
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const int x = by_col ? (k & 3) : (k >> 2);
    const ReplayEnt &E = by_col ? tab[NC * NCODE + x] : tab[x * NCODE + NC];
    const uint32_t m = (E.mask >> k & 1u) ? cnt[x] : 0u;
    if (m) s.fc[k] = add_repeat(s.fc[k], E.v, pow2_inv(E.v), m);
  }
#pragma unroll
  for (int x = 0; x < 4; x++) {
    const ReplayEnt &E = by_col ? tab[NC * NCODE + x] : tab[x * NCODE + NC];
#pragma unroll
    for (int d = 0; d < 4; d++) s.dgi[d] += E.icd == d ? cnt[x] : 0u;
    s.amb += E.amb * cnt[x];
  }
}

// Counts of A, C, G, T of sequence q over sites [p, e): differences of its plane prefix counts (clean, C|T, G|T, T)
__device__ inline void base_counts(const RunArgs &A, int64_t q, uint32_t p, uint32_t e, uint32_t (&n)[4])
{
  const uint4 hi = plane_prefix(A.pre, A.W, q, e), lo = plane_prefix(A.pre, A.W, q, p);
  const uint32_t c = hi.x - lo.x, b0 = hi.y - lo.y, b1 = hi.z - lo.z, t = hi.w - lo.w;
  n[0] = c - b0 - b1 + t;
  n[1] = b0 - t;
  n[2] = b1 - t;
  n[3] = t;
}

// Site-order replay for a dense-N set: the ambiguity events of both sequences plus their N runs, swept as segments
// between breakpoints (events, run starts and ends).  Inside a segment neither sequence has an event, so its sites
// are N against a base, a gap or N: the same effect repeated, column (row) b's cells only for partner base b, so the
// segment is applied per base with replay_apply_rep.  An event inside the other sequence's N run is the site
// (N, code).
__device__ void replay_events_dense(const RunArgs &A, int64_t i, int64_t j, const ReplayEnt *tab, ReplayState &s)
{
  constexpr uint32_t INF = 0xffffffffu, NOEV = 0xffffffffu;  // no event word is all ones (sites < 2^26)
  const uint2 NORUN = make_uint2(INF, INF);
  int a = A.ev_off[i], ae = A.ev_off[i + 1], b = A.ev_off[j], be = A.ev_off[j + 1];
  int ra = A.nrun_off[i], rae = A.nrun_off[i + 1], rb = A.nrun_off[j], rbe = A.nrun_off[j + 1];
  const uint8_t *si = A.codes + i * A.Ls, *sj = A.codes + j * A.Ls;
  const ReplayEnt &nn = tab[NC * NCODE + NC];
  // the heads of the four lists, each read again only when it advances
  uint32_t wa = a < ae ? A.ev[a] : NOEV, wb = b < be ? A.ev[b] : NOEV;
  uint2 ri = ra < rae ? A.nruns[ra] : NORUN, rj = rb < rbe ? A.nruns[rb] : NORUN;
  uint32_t p = 0;
  for (;;) {
    const uint32_t ea = wa == NOEV ? INF : wa >> 5, eb = wb == NOEV ? INF : wb >> 5;
    const bool ni = ri.x <= p, nj = rj.x <= p;  // inside an N run at p
    const uint32_t q = min(min(ea, eb), min(ni ? ri.y : ri.x, nj ? rj.y : rj.x));
    if (q == INF) break;
    if (q > p && (ni || nj)) {  // the segment [p, q)
      if (ni && nj) {
        replay_apply_rep(s, nn, q - p);
      } else {
        uint32_t cnt[4];
        base_counts(A, ni ? j : i, p, q, cnt);
        replay_apply_seg(s, tab, cnt, ni);
      }
    }
    p = q;
    if (ri.y <= p) ri = ++ra < rae ? A.nruns[ra] : NORUN;  // runs are maximal: the next starts after p
    if (rj.y <= p) rj = ++rb < rbe ? A.nruns[rb] : NORUN;
    if (ea == p || eb == p) {  // the site p
      const unsigned c1 = ea == p ? (wa & 31u) : ri.x <= p ? (unsigned)NC : si[p];
      const unsigned c2 = eb == p ? (wb & 31u) : rj.x <= p ? (unsigned)NC : sj[p];
      replay_apply(s, tab[c1 * NCODE + c2]);
      if (ea == p) wa = ++a < ae ? A.ev[a] : NOEV;
      if (eb == p) wb = ++b < be ? A.ev[b] : NOEV;
      p++;
      if (ri.y <= p) ri = ++ra < rae ? A.nruns[ra] : NORUN;
      if (rj.y <= p) rj = ++rb < rbe ? A.nruns[rb] : NORUN;
    }
  }
}

// A wave's LDS in the replay: 64 row slots (the wave's row parts, summed per row) or, for long alignments, the 64
// lanes' clean/clean tables (built by the wave before the walk, read before any slot is used).
struct RowSlot {
  double slo, shi, sabs;
  unsigned long long mlo, mhi;  // bit patterns of max thisD (non-negative doubles order like their bits); 0: none
  unsigned long long wsum, skip, linkw;
};
union ReplayLds {
  uint32_t scc[64][16];
  RowSlot slot[64];
};

// Adds each lane's row part to its row's accumulator, one set of global atomics per distinct row of the wave (lanes
// of one row hitting one address with global atomics serialise in the L2): a ballot loop gives every distinct row a
// slot of the wave's LDS, the lanes add their parts into their slots with LDS atomics, and one lane per slot sends
// the slot on.  All 64 lanes must call it; lanes without a part pass row < 0.
__device__ void rowacc_add_wave(const RunArgs &A, int64_t row, const RowPart &rp, RowSlot *slot)
{
  const int lane = threadIdx.x & 63;
  bool pend = row >= 0, lead = false;
  int my = 0, ns = 0;
  for (;;) {
    const unsigned long long m = __ballot(pend);
    if (!m) break;
    const int leader = __ffsll((long long)m) - 1;
    const int64_t r = (int64_t)((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(row >> 32), leader) << 32 |
                                (uint32_t)__builtin_amdgcn_readlane((int)row, leader));
    if (pend && row == r) {
      my = ns;
      lead = lane == leader;
      pend = false;
    }
    ns++;
  }
  if (lead) slot[my] = RowSlot{0., 0., 0., 0ull, 0ull, 0ull, 0ull, 0ull};
  __builtin_amdgcn_wave_barrier();  // LDS operations of one wave complete in order
  if (row >= 0) {
    RowSlot *q = slot + my;
    if (rp.slo != 0.) atomicAdd(&q->slo, rp.slo);
    if (rp.shi != 0.) atomicAdd(&q->shi, rp.shi);
    if (rp.sabs != 0.) atomicAdd(&q->sabs, rp.sabs);
    if (rp.mlo > 0.) atomicMax(&q->mlo, (unsigned long long)tn93g_bits(rp.mlo));
    if (rp.mhi > 0.) atomicMax(&q->mhi, (unsigned long long)tn93g_bits(rp.mhi));
    if (rp.wsum) atomicAdd(&q->wsum, (unsigned long long)rp.wsum);
    if (rp.skip) atomicAdd(&q->skip, (unsigned long long)rp.skip);
    if (rp.linkw) atomicAdd(&q->linkw, (unsigned long long)rp.linkw);
  }
  __builtin_amdgcn_wave_barrier();
  if (lead) {
    const RowSlot v = slot[my];
    RowPart t;
    t.slo = v.slo;
    t.shi = v.shi;
    t.sabs = v.sabs;
    t.mlo = v.mlo ? tn93g_from_bits(v.mlo) : -1.;
    t.mhi = v.mhi ? tn93g_from_bits(v.mhi) : -1.;
    t.wsum = (long long)v.wsum;
    t.skip = (long long)v.skip;
    t.linkw = (long long)v.linkw;
    rowacc_add(A, row, t);
  }
}

// one thread per pair that needs the site-order replay (the flat list the pair kernel's tiles appended to); every
// lane of a wave runs the same iterations so the row parts can be combined across the wave.  The replay is latency
// bound: a register budget for 4 waves per SIMD measured fastest on the MI355X (also with dense N: C2 470 -> 460 ms
// against 3 waves), although the compiler then spills some registers to scratch.
// The wave's 64 replay-list entries kb + lane (of `total`, from entry `beg` of slow_list) of run A, with A's replay
// tables in `tabs`; histogram counts and certification counters go to the block's lhist / bs.  Every lane of the
// wave calls it (the long-alignment count tables are built by the whole wave).
template <bool DENSE>
__device__ __attribute__((always_inline)) inline void replay_entries(const RunArgs &A, const ReplayEnt *tabs,
                                                                      unsigned long long *lhist, BlockStats *bs,
                                                                      ReplayLds *wl, int64_t beg, int64_t total,
                                                                      int64_t kb)
{
  const int lane = threadIdx.x & 63;
  const int64_t k = kb + lane;
  int64_t row = -1;
  RowPart rp = rowpart_zero();
  const bool coop = DENSE && A.W >= 128;  // uniform
  unsigned long long pr = 0;
  if (k < total) pr = A.slow_list[beg + k];
  uint32_t icc[16];
  const bool looked = k < total && cc_table_lookup(A, (int64_t)(pr >> 32 & 0x7fffffffu),
                                                   (int64_t)(pr & 0xffffffffull), icc);
  if (coop)
    cc_table_wave(A, (int64_t)(pr >> 32 & 0x7fffffffu), (int64_t)(pr & 0xffffffffull), k < total && !looked, wl->scc);
  const bool avg_first = pr >> 63;
  pr &= ~(1ull << 63);
  if (k < total) {
    const int64_t i = (int64_t)(pr >> 32), j = (int64_t)(pr & 0xffffffffull);
    if (looked) {
    } else if (coop) {
#pragma unroll
      for (int q = 0; q < 16; q++) icc[q] = wl->scc[lane][q];
    } else {
      cc_table(A, i, j, icc);
    }
#pragma unroll
    for (int b = 0; b < 4; b++) icc[5 * b] += A.kc[b];  // the folded constant columns
    ReplayState s;
    replay_reset(s);
    // the main mode first, unless the pair kernel's -g test (on exact, dyadic sums) already chose AVERAGE
    const ReplayEnt *t0 = avg_first ? tabs + NCODE * NCODE : tabs;
    if constexpr (DENSE)
      replay_events_dense(A, i, j, t0, s);
    else
      replay_events(A, i, j, t0, s);  // GAPMM: the lists hold the gap sites too
    tn93g_tailin tin;
    double tng;
    replay_merge(s, icc, tin, tng);
    bool skip = tng <= (double)A.min_overlap;
    if (!avg_first && !skip && A.avg_fallback &&
        tng * A.rf <= (double)s.amb) {  // recompute in AVERAGE (refkernel.cpp:742-750)
      replay_reset(s);
      if constexpr (DENSE)
        replay_events_dense(A, i, j, tabs + NCODE * NCODE, s);
      else
        replay_events(A, i, j, tabs + NCODE * NCODE, s);
      replay_merge(s, icc, tin, tng);
      skip = tng <= (double)A.min_overlap;
    }
    if (skip)
      skip_pair(A, i, j, rp);
    else
      A.fast ? finish_pair<true>(A, i, j, tin, lhist, bs, rp) : finish_pair<false>(A, i, j, tin, lhist, bs, rp);
    row = i;
  }
  if (A.fast) rowacc_add_wave(A, row, rp, wl->slot);
}

template <bool DENSE>
__global__ __launch_bounds__(RB, 4) void k_replay(RunArgs A)
{
  __shared__ ReplayEnt tabs[2 * NCODE * NCODE];
  __shared__ unsigned long long lhist[3 * 200];
  __shared__ BlockStats bs;
  __shared__ ReplayLds wl[RB / 64];
  const int tid = threadIdx.x, wave = tid >> 6;
  const int64_t total = (int64_t)*A.slow_total;
  if ((int64_t)blockIdx.x * RB >= total) return;  // no pair for this block (uniform): a run without replays is cheap
  for (int k = tid; k < 2 * NCODE * NCODE; k += RB) tabs[k] = A.rtab[k];
  for (int k = tid; k < 600; k += RB) lhist[k] = 0;
  if (tid == 0) bs.unc = bs.undecided = 0;
  __syncthreads();
  const int64_t nw = (int64_t)gridDim.x * (RB / 64);
  for (int64_t kb = ((int64_t)blockIdx.x * (RB / 64) + wave) * 64; kb < total; kb += nw * 64)
    replay_entries<DENSE>(A, tabs, lhist, &bs, wl + wave, 0, total, kb);
  flush_block(A, &bs, lhist, tid, RB);
}

// The replays of several jobs (a batch) in one launch.  The jobs' lists are numbered one after another in units of
// RB entries; block b takes the units u = b (mod gridDim.x), a job's units together: for each job with units for it,
// the block loads that job's tables, replays its units and flushes its counters into that job's stripes.
template <bool DENSE>
__global__ __launch_bounds__(RB, 4) void k_replay_multi(const RunArgs *__restrict__ jobs,
                                                        const int32_t *__restrict__ ids, int nj)
{
  __shared__ ReplayEnt tabs[2 * NCODE * NCODE];
  __shared__ unsigned long long lhist[3 * 200];
  __shared__ BlockStats bs;
  __shared__ ReplayLds wl[RB / 64];
  const int tid = threadIdx.x, wave = tid >> 6;
  const int64_t G = gridDim.x, b = blockIdx.x;
  int64_t cum = 0;  // units of the jobs before q (uniform)
  for (int q = 0; q < nj; q++) {
    const RunArgs &A = jobs[ids[q]];
    const int64_t total = (int64_t)*A.slow_total;
    const int64_t units = (total + RB - 1) / RB;
    const int64_t u0 = cum + (((b - cum) % G) + G) % G;  // the block's first unit at or after cum
    cum += units;
    if (u0 >= cum) continue;
    __syncthreads();  // the previous job's tables and counters are no longer read
    for (int k = tid; k < 2 * NCODE * NCODE; k += RB) tabs[k] = A.rtab[k];
    for (int k = tid; k < 600; k += RB) lhist[k] = 0;
    if (tid == 0) bs.unc = bs.undecided = 0;
    __syncthreads();
    for (int64_t u = u0; u < cum; u += G)
      replay_entries<DENSE>(A, tabs, lhist, &bs, wl + wave, 0, total, (u - (cum - units)) * RB + wave * 64);
    flush_block(A, &bs, lhist, tid, RB);
  }
}

// the links of the NSTRIPE regions, concatenated into `out` (one block per region)
__device__ __attribute__((always_inline)) inline void link_gather(const RunArgs &A, LinkRec *__restrict__ out,
                                                                  const int64_t *__restrict__ row_off,
                                                                  uint32_t *__restrict__ slot, const int st)
{
  __shared__ int64_t soff;
  if (threadIdx.x < 64) {  // the links of the stripes before st, summed across one wave (NSTRIPE = 64)
    int64_t v = (int)threadIdx.x < st ? (int64_t)min(A.stripes[threadIdx.x * 32 + ST_LINKS], A.link_cap) : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (threadIdx.x == 0) soff = v;
  }
  __syncthreads();
  const int64_t off = soff;
  const unsigned cnt = min(A.stripes[st * 32 + ST_LINKS], A.link_cap);
  const int64_t cap = (int64_t)A.link_cap * NSTRIPE;
  for (unsigned k = threadIdx.x; k < cnt; k += blockDim.x) {
    const LinkRec L = A.links[(size_t)st * A.link_cap + k];
    out[off + k] = L;
    const int64_t r = L.i - A.r0;
    const int64_t pos = row_off[r] + atomicSub(&A.row_links[r], 1) - 1;  // the row's count serves as its cursor
    if (pos < cap) slot[pos] = (uint32_t)(off + k);
  }
}

__global__ void k_link_gather(RunArgs A) { link_gather(A, A.links_flat, A.link_off, A.lslot, blockIdx.x); }

// the same for several jobs: block (st, k) is stripe st of job k
__global__ void k_link_gather_multi(const RunArgs *__restrict__ jobs)
{
  const RunArgs &A = jobs[blockIdx.y];
  link_gather(A, A.links_flat, A.link_off, A.lslot, blockIdx.x);
}

// The links in tn93's emission order (rows ascending, then columns ascending), into out.  Row r's links are
// slot[row_off[r] .. row_off[r + 1]) (indices into flat); their rank within the row is their column's rank.  Up to 64
// links: ranks by comparison across one wave.  More: a bitmap of the row's link columns, BMW words per window of
// columns, and a prefix count of its words.  Columns of a row are distinct (one record per pair).  A run whose link
// regions overflowed is repeated, so indices are checked against the capacity only to stay in bounds.
constexpr int BMW = 4096;
__device__ __attribute__((always_inline)) inline void link_order(const RunArgs &A, const LinkRec *__restrict__ flat,
                                                                 const int64_t *__restrict__ row_off,
                                                                 const uint32_t *__restrict__ slot,
                                                                 LinkRec *__restrict__ out, int64_t rbeg, int64_t rstep)
{
  __shared__ uint32_t bits[BMW];
  __shared__ uint32_t pre[BMW];
  __shared__ uint32_t wtot[4];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t cap = (int64_t)A.link_cap * NSTRIPE;
  const int64_t nrows = A.r1 - A.r0;
  for (int64_t r = rbeg; r < nrows; r += rstep) {
    const int64_t b = row_off[r], e = min(row_off[r + 1], cap);
    const int64_t cnt = e - b;
    if (cnt <= 0) continue;
    if (cnt <= 64) {
      if (wave == 0) {
        const uint32_t idx = lane < cnt ? slot[b + lane] : 0u;
        const bool ok = lane < cnt && (int64_t)idx < cap;
        const int64_t jj = ok ? flat[idx].j : INT64_MAX;
        int rank = 0;
        for (int m = 0; m < cnt; m++) rank += __shfl(jj, m, 64) < jj ? 1 : 0;
        if (ok) out[b + rank] = flat[idx];
      }
      continue;
    }
    int64_t base = 0;  // links of the row in earlier windows
    for (int64_t w0 = A.c0; w0 < A.c1; w0 += (int64_t)BMW * 32) {
      for (int k = tid; k < BMW; k += 256) bits[k] = 0;
      __syncthreads();
      for (int64_t k = b + tid; k < e; k += 256) {
        const uint32_t idx = slot[k];
        if ((int64_t)idx >= cap) continue;
        const int64_t c = flat[idx].j - w0;
        if (c >= 0 && c < (int64_t)BMW * 32) atomicOr(&bits[c >> 5], 1u << (c & 31));
      }
      __syncthreads();
      // exclusive prefix of the words' popcounts: 16 words per thread, then across the block
      uint32_t loc[BMW / 256], sum = 0;
#pragma unroll
      for (int q = 0; q < BMW / 256; q++) {
        loc[q] = sum;
        sum += __popc(bits[tid * (BMW / 256) + q]);
      }
      uint32_t x = sum;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane == 63) wtot[wave] = x;
      __syncthreads();
      uint32_t wpre = 0;
      for (int v = 0; v < wave; v++) wpre += wtot[v];
      const uint32_t tpre = wpre + x - sum;
#pragma unroll
      for (int q = 0; q < BMW / 256; q++) pre[tid * (BMW / 256) + q] = tpre + loc[q];
      const uint32_t wincnt = wtot[0] + wtot[1] + wtot[2] + wtot[3];
      __syncthreads();
      for (int64_t k = b + tid; k < e; k += 256) {
        const uint32_t idx = slot[k];
        if ((int64_t)idx >= cap) continue;
        const LinkRec L = flat[idx];
        const int64_t c = L.j - w0;
        if (c < 0 || c >= (int64_t)BMW * 32) continue;
        const int64_t rank = base + pre[c >> 5] + __popc(bits[c >> 5] & ((1u << (c & 31)) - 1u));
        if (b + rank < e) out[b + rank] = L;
      }
      base += wincnt;
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void k_link_order(RunArgs A)
{
  link_order(A, A.links_flat, A.link_off, A.lslot, A.links, blockIdx.x, gridDim.x);
}

// the same for several jobs: job k owns blocks [boff[k], boff[k + 1])
__global__ __launch_bounds__(256) void k_link_order_multi(const RunArgs *__restrict__ jobs,
                                                          const int64_t *__restrict__ boff, int nj)
{
  const int k = job_of(boff, nj, blockIdx.x);
  const RunArgs &A = jobs[k];
  link_order(A, A.links_flat, A.link_off, A.lslot, A.links, (int64_t)blockIdx.x - boff[k], boff[k + 1] - boff[k]);
}

// Zeroes a list of ranges in one launch (a batch's per-job buffers): grid row y takes ranges y, y + gridDim.y, ...
struct FillRange {
  uint32_t *p;
  int64_t n;  // 32-bit words
};
// Copies a list of ranges (8-byte units) in one launch: a batch's read-back staged contiguously for one transfer.
struct CopyRange {
  const uint2 *src;
  uint2 *dst;
  int64_t n;
};
__global__ void k_copy_ranges(const CopyRange *__restrict__ r, int nr)
{
  for (int k = blockIdx.y; k < nr; k += gridDim.y) {
    const uint2 *src = r[k].src;
    uint2 *dst = r[k].dst;
    const int64_t n = r[k].n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
      dst[q] = src[q];
  }
}

// A single run's read-back (at most 4 ranges, passed by value) written by the kernel straight into the caller's
// pinned host buffer (its device mapping): one launch after the run's kernels instead of a copy per range, each of which
// costs a switch to the copy engine and back (~10 us).  Grid row y copies range y.
// With nlk > 0, the last grid row writes the first nlk link records straight into a result's five link arrays (a
// pinned block's device mapping).
struct CopyList {
  CopyRange r[4];
  const LinkRec *lk;
  int64_t *row, *col;
  double *d, *lo, *hi;
  int64_t nlk;
};
__global__ __launch_bounds__(256) void k_copy_list(CopyList L, int nr)
{
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
  if ((int)blockIdx.y < nr) {
    const CopyRange g = L.r[blockIdx.y];
    for (int64_t q = t0; q < g.n; q += step) g.dst[q] = g.src[q];
    return;
  }
  for (int64_t q = t0; q < L.nlk; q += step) {
    const LinkRec r = L.lk[q];
    L.row[q] = r.i;
    L.col[q] = r.j;
    L.d[q] = r.d;
    L.lo[q] = r.lo;
    L.hi[q] = r.hi;
  }
}

// The device's constant-rate wall clock into *t (one lane): stamps inside a captured graph, where events time nothing.
__global__ void k_stamp(unsigned long long *t)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *t = (unsigned long long)wall_clock64();
}

// A single run's zero-started buffers (at most 4 ranges, passed by value): grid row y clears range y.
struct FillList {
  FillRange r[4];
};
__global__ void k_fill_list(FillList L)
{
  const unsigned y = blockIdx.y;  // uniform: the range is picked without indexing the parameter by a variable
  const FillRange f = y == 0 ? L.r[0] : y == 1 ? L.r[1] : y == 2 ? L.r[2] : L.r[3];
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < f.n; q += (int64_t)gridDim.x * blockDim.x)
    f.p[q] = 0u;
}

__global__ void k_fill_ranges(const FillRange *__restrict__ r, int nr)
{
  for (int k = blockIdx.y; k < nr; k += gridDim.y) {
    uint32_t *p = r[k].p;
    const int64_t n = r[k].n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
      p[q] = 0u;
  }
}

// histogram replicas summed into replica 0
__device__ inline void hist_sum(const RunArgs &A, int k)
{
  if (k >= 600) return;
  unsigned long long v = 0;
  for (int st = 0; st < NSTRIPE; st++) v += A.hist[st * 600 + k];
  A.hist[k] = v;
}
__global__ void k_hist_sum(RunArgs A) { hist_sum(A, blockIdx.x * blockDim.x + threadIdx.x); }
// A single fast run's link order and histogram sum in one launch: blocks [0, nob) order the links as k_link_order,
// the three after them sum the histogram replicas as k_hist_sum (both follow the tail; neither reads the other's data).
__global__ __launch_bounds__(256) void k_link_order_hist(RunArgs A, unsigned nob)
{
  if (blockIdx.x >= nob) {
    hist_sum(A, (int)(blockIdx.x - nob) * 256 + (int)threadIdx.x);
    return;
  }
  link_order(A, A.links_flat, A.link_off, A.lslot, A.links, blockIdx.x, nob);
}
// the same for several jobs: grid row y is job y
__global__ void k_hist_sum_multi(const RunArgs *__restrict__ jobs)
{
  hist_sum(jobs[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x);
}

// host-exact distances back into D (after TN93G_FLAG_EXACT)
__global__ void k_patch(RunArgs A, const PatchRec *__restrict__ P, int64_t np)
{
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const PatchRec r = P[k];
  const int64_t off = (r.i - A.r0) * A.ncols + (r.j - A.c0);
  A.D[off] = r.thisD;
  A.U[off] = 0;
  const int64_t w = A.weights[r.i] * A.weights[r.j];
  atomicAdd(&A.hist[fst_bin(A, r.i, r.j) * 200 + r.idx], (unsigned long long)w);
  if (r.link) atomicAdd(&A.row_links[r.i - A.r0], 1);
}

// Exclusive prefix sum of n int32 values (the per-row link counts) into out[0..n), total in out[n]; one block of
// 1024 threads walks the array in chunks of 4096 (4 consecutive values per thread, wave scans, then a scan of the 16
// wave totals).
__device__ __attribute__((always_inline)) inline void scan_rows(const int32_t *__restrict__ in,
                                                                int64_t *__restrict__ out, int64_t n)
{
  __shared__ int64_t wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int64_t carry = 0;
  for (int64_t base = 0; base < n; base += 4096) {
    const int64_t k0 = base + 4 * t;
    int64_t v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = k0 + q < n ? in[k0 + q] : 0;
    const int64_t s = v[0] + v[1] + v[2] + v[3];
    int64_t x = s;  // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
      int64_t z = lane < 16 ? wsum[lane] : 0;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        const int64_t y = __shfl_up(z, o, 64);
        if (lane >= o) z += y;
      }
      if (lane < 16) wsum[lane] = z;
    }
    __syncthreads();
    int64_t run = carry + (w ? wsum[w - 1] : 0) + x - s;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (k0 + q < n) out[k0 + q] = run;
      run += v[q];
    }
    carry += wsum[15];
    __syncthreads();
  }
  if (t == 0) out[n] = carry;
}
__global__ __launch_bounds__(1024) void k_scan(const int32_t *__restrict__ in, int64_t *__restrict__ out, int64_t n)
{
  scan_rows(in, out, n);
}
// the link offsets of several jobs (fast path): block k scans job k's row counts
__global__ __launch_bounds__(1024) void k_scan_multi(const RunArgs *__restrict__ jobs)
{
  const RunArgs &A = jobs[blockIdx.x];
  scan_rows(A.row_links, A.link_off, A.r1 - A.r0);
}

// ------------------------------------------------------------------------------------------ row kernel
// One thread per row: tn93's per-row reductions in column order (swifttn93.cpp:2547-2603): links in emission
// order, the skip count, and lower/upper enclosures of local_sum and local_max; local_weighted exactly.
struct RowOut {
  double *sum_lo, *sum_hi, *wsum, *max_lo, *max_hi;  // [3][nrows]
  int64_t *skipped, *linkw;                          // [nrows]
  int64_t *link_row, *link_col;
  double *link_d, *link_lo, *link_hi;
  const int64_t *link_off;
};

constexpr int RT = 64;  // rows per block (one wave) and columns per tile of the row kernel

__global__ __launch_bounds__(RT) void k_rows(RunArgs A, RowOut O)
{
  __shared__ double sD[RT][RT + 1];
  __shared__ uint16_t sU[RT][RT + 2];
  __shared__ int64_t sW[RT];
  const int lane = threadIdx.x;
  const int64_t nrows = A.r1 - A.r0;
  const int64_t rb = (int64_t)blockIdx.x * RT, r = rb + lane;
  const bool live = r < nrows;
  const int64_t i = A.r0 + (live ? r : nrows - 1);
  const int64_t lower = A.triangle ? max(i + 1, A.c0) : A.c0;  // an append run's columns start at c0
  const int64_t wi = A.weights[i];
  double slo[3] = {0., 0., 0.}, shi[3] = {0., 0., 0.}, ws[3] = {0., 0., 0.}, mlo[3] = {0., 0., 0.},
         mhi[3] = {0., 0., 0.};
  if (A.triangle && i >= A.c0 && wi > 1) {  // self_v_self (swifttn93.cpp:2406-2415), the run's own sequences
    unsigned long long svs = (unsigned long long)((wi * (wi - 1)) >> 1);
    ws[(A.fst && i >= A.n1) ? 1 : 0] += (double)svs;
  }
  int64_t skipped = 0, linkw = 0, k = O.link_off[live ? r : 0];
  const int64_t kend = live ? O.link_off[r + 1] : k;
  // column tiles: from the tile holding the block's first column to the end
  const int64_t cfirst = A.triangle ? max(A.r0 + rb + 1, A.c0) : A.c0;
  const int64_t jt0 = cfirst & ~(int64_t)(RT - 1);
  // tile loader: lane = column; rows rb .. rb + 63 of the tile
  double nd[RT];
  uint16_t nu[RT];
  auto load = [&](int64_t jt) {
    const int64_t col = jt + lane;
    const bool cok = col >= A.c0 && col < A.c1;
#pragma unroll
    for (int rr = 0; rr < RT; rr++) {
      const int64_t row = rb + rr;
      const bool ok = cok && row < nrows;
      const int64_t off = ok ? row * A.ncols + (col - A.c0) : 0;
      nd[rr] = ok ? A.D[off] : 0.0;
      nu[rr] = ok ? A.U[off] : (uint16_t)0;
    }
  };
  if (jt0 < A.c1) load(jt0);
  for (int64_t jt = jt0; jt < A.c1; jt += RT) {
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < RT; rr++) {
      sD[rr][lane] = nd[rr];
      sU[rr][lane] = nu[rr];
    }
    sW[lane] = (jt + lane < A.c1) ? A.weights[jt + lane] : 0;
    __syncthreads();
    if (jt + RT < A.c1) load(jt + RT);  // next tile in flight while this one is reduced
    if (!live) continue;
    const int64_t jb = max(jt, lower), je = min(jt + RT, A.c1);
    for (int64_t j = jb; j < je; j++) {
      const int jj = (int)(j - jt);
      const double x = sD[lane][jj];
      const uint16_t u = sU[lane][jj];
      const int bin = fst_bin(A, i, j);
      const int64_t w = wi * sW[jj];
      const double lo = u ? tn93g_from_bits(tn93g_bits(x) - (u & 255u)) : x;
      const double hi = u ? tn93g_from_bits(tn93g_bits(x) + (u >> 8)) : x;
      if (x >= A.mind && x <= A.thr && k < kend) {
        linkw += w;
        O.link_row[k] = i;
        O.link_col[k] = j;
        O.link_d[k] = x;
        O.link_lo[k] = lo;
        O.link_hi[k] = hi;
        k++;
      }
      if (x <= -0.5) {
        skipped++;
      } else {
        const double wd = (double)w;
        double tl = lo * wd, th = hi * wd;
        if (w < 0) {
          const double q = tl;
          tl = th;
          th = q;
        }
        slo[bin] += tl;
        shi[bin] += th;
        ws[bin] += wd;
        if (lo > mlo[bin]) mlo[bin] = lo;
        if (hi > mhi[bin]) mhi[bin] = hi;
      }
    }
  }
  if (!live) return;
  for (int bb = 0; bb < 3; bb++) {
    O.sum_lo[bb * nrows + r] = slo[bb];
    O.sum_hi[bb * nrows + r] = shi[bb];
    O.wsum[bb * nrows + r] = ws[bb];
    O.max_lo[bb * nrows + r] = mlo[bb];
    O.max_hi[bb * nrows + r] = mhi[bb];
  }
  O.skipped[r] = skipped;
  O.linkw[r] = linkw;
}

// ------------------------------------------------------------------------------------------ host side
std::string fmt(const char *f, ...)
{
  char b[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(b, sizeof b, f, ap);
  va_end(ap);
  return b;
}

}  // namespace

// A set of sequences on the device, sized for `cap` sequences; appends fill it in order.  Runs take a snapshot of
// (d, n) under the mutex and only look at the first n sequences, so a set may grow while runs on it are in flight.
// Event arrays that must grow are replaced, and the old ones kept until the set is freed (in-flight runs may still
// read them; their prefix is unchanged).
struct tn93g_set {
  tn93g_ctx *ctx = nullptr;
  SetDev d{};
  int64_t cap = 0;
  int64_t ev_total[4] = {0, 0, 0, 0}, ev_cap[4] = {0, 0, 0, 0};  // entries used / allocated: ev, evg, nruns, gruns
  bool decided = false;                                  // ndense fixed by the first append
  std::vector<int64_t> hw;                            // copy numbers (host), cap entries
  int64_t nonunit = 0;                                // sequences whose copy number is not 1
  std::vector<uint8_t> spec;                          // 1: the sequence has a B, D, H or V code (host), cap entries
  std::vector<void *> retired;                        // replaced event arrays
  void *listcnt = nullptr;                            // the list lengths of an append (5 x n int32), grown on demand
  size_t listcnt_bytes = 0;
  std::mutex mu;
};

static double tn93g_host_from_bits(unsigned long long u)
{
  double x;
  memcpy(&x, &u, 8);
  return x;
}

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};
struct HostBuf {  // pinned host memory (hipHostMalloc), grown on demand
  void *p = nullptr;
  size_t n = 0;
};

// The device buffers of one run.  A single run uses slot 0; a batch gives each of its jobs a slot of its own, so all
// of a batch's kernels can be enqueued before any result is read back.
struct RunBufs {
  DevBuf D, U, hist, row_links, link_off, slow_list, counters, exact_list, rowout,
      links, links_flat, lslot, evtab, rtab, patch, chdr, cdata, spec_idx, spec_ids, spec_tab;
  std::vector<int32_t> h_spec_idx;  // host staging of spec_idx / spec_ids (kept until the next run of the slot)
  std::vector<int64_t> h_spec_ids;
  int64_t link_cap_hint = 0;  // link region size that sufficed last time
  int64_t exact_hint = 0;     // host-exact list size that sufficed last time
  size_t link_spec = 4096;     // link records a single run reads back with its counters
  int tab_mode = -1;           // mode and subset of the tables in evtab/rtab (uploaded again only when they change)
  uint32_t tab_mask = 0;
  bool tab_fold = false;       // the folded tables are valid (the marginal form of the pair kernel may run)
  bool tab_gcert = false;      // every order-dependent site of the main table weighs 0 or 1 (see event_entry)
  void release()
  {
    for (DevBuf *b : {&D, &U, &hist, &row_links, &link_off, &slow_list, &counters, &exact_list, &rowout, &links,
                      &links_flat, &lslot, &evtab, &rtab, &patch, &chdr, &cdata, &spec_idx, &spec_ids, &spec_tab}) {
      if (b->p) (void)hipFree(b->p);
      b->p = nullptr;
      b->n = 0;
    }
  }
};

constexpr int NBRANCH = 8;  // parallel branches of a batch graph

// The work memory of tn93g_run_noalloc: one file-scope device array, of which every GPU the library runs on holds its
// own copy from the loading of the library until the process ends.  In host code its name is only a stand-in: the
// array's address on a device comes from hipGetSymbolAddress with that device current (g_scratch below, taken once
// per device when a context opens), and only that address is ever passed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[TN93G_NOALLOC_SCRATCH_BYTES];
constexpr int MAX_SCRATCH_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[MAX_SCRATCH_DEVICES];
static hipError_t g_scratch_err[MAX_SCRATCH_DEVICES];

// The device address of `scratch` on `device` (which must be current), looked up once; nullptr and the HIP error
// when the lookup failed.
static void *scratch_address(int device, hipError_t *err)
{
  *err = hipErrorInvalidDevice;
  if (device < 0 || device >= MAX_SCRATCH_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device] && g_scratch_err[device] == hipSuccess) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess && ptr)
      g_scratch[device] = ptr;
    else
      g_scratch_err[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
  }
  *err = g_scratch_err[device];
  return g_scratch[device];
}

struct tn93g_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side[NBRANCH] = {};  // the branches of a batch, forked from and joined back into `stream`
  hipEvent_t fork = nullptr, join[NBRANCH] = {};
  // per branch (a single run uses 0): the stream on which a job's tail kernel runs beside its replay
  hipStream_t rside[NBRANCH] = {};
  hipEvent_t rfork[NBRANCH] = {}, rjoin[NBRANCH] = {};
  hipEvent_t ev[5];  // pair kernel start, its end, run end, tail end, run start
  rocblas_handle blas = nullptr;  // single runs' special-sequence tables as int8 GEMMs (made on first use)
  int64_t table_runs = 0;         // single runs that built special-sequence tables
  int64_t count_runs = 0;         // single runs that could take the GEMM counts
  void *gemm_out = nullptr;       // the GEMM counts' four products
  size_t gemm_out_n = 0;
  void *gemm_ops = nullptr;       // their one-hot operands
  size_t gemm_ops_n = 0;
  std::string err;
  std::vector<RunBufs *> slots;
  HostBuf pin;  // staging of the read-backs (counters, links, row accumulators, histogram)
  HostBuf pin_links;  // the links of a single run that produced more than its read-back expected
  HostBuf pin_up;     // staging of appended sequences (sets owned by this context)
  // the instantiated graph of the last batch, launched again when a batch enqueues exactly the same work
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;
  // autotuner: the fastest pair-kernel variant measured for each job shape (see tune_shape)
  std::map<uint64_t, int> tuned;
  // a batch's merged pair-kernel launches: per launch the jobs' RunArgs, then their tile offsets; the fill and copy
  // range lists; and the staging of the batch's read-back
  DevBuf multi, stage;
  int wall_khz = 0;  // rate of the device's constant wall clock (the stamps around a batch's merged pair kernels)
  void *scratch = nullptr;  // device address of this device's `scratch` (tn93g_run_noalloc), set when opened
  hipError_t scratch_err = hipSuccess;

  RunBufs *slot(size_t k)
  {
    while (slots.size() <= k) slots.push_back(new RunBufs());
    return slots[k];
  }
  int fail(int code, const std::string &m)
  {
    err = m;
    return code;
  }
};

#define HIPCHK(ctx, x)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (x);                                                                                 \
    if (e_ != hipSuccess) return (ctx)->fail(TN93G_E_HIP, fmt("%s: %s", #x, hipGetErrorString(e_))); \
  } while (0)

static int ensure_pinned(tn93g_ctx *c, HostBuf &b, size_t bytes)
{
  if (b.n >= bytes && b.p) return 0;
  if (b.p) (void)hipHostFree(b.p);
  b.p = nullptr;
  b.n = 0;
  size_t want = std::max(bytes, (size_t)65536);
  hipError_t e = hipHostMalloc(&b.p, want);
  if (e != hipSuccess) return c->fail(TN93G_E_NOMEM, fmt("hipHostMalloc(%zu): %s", want, hipGetErrorString(e)));
  b.n = want;
  return 0;
}

static int ensure(tn93g_ctx *c, DevBuf &b, size_t bytes)
{
  if (b.n >= bytes && b.p) return 0;
  // a buffer that has to grow grows by at least a quarter, so that runs on a growing set reallocate rarely
  size_t want = std::max({bytes, (size_t)256, b.n + b.n / 4});
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.n = 0;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) return c->fail(TN93G_E_NOMEM, fmt("hipMalloc(%zu): %s", want, hipGetErrorString(e)));
  b.n = want;
  return 0;
}

// tn93's tail with the host libm (refkernel.cpp:752-815), for the host-exact path
static double host_tail(const tn93g_tailin &in)
{
  const double *nucFreq = in.nf;
  double auxd, fY, fR, K1 = 0, K2 = 0, K3 = 0, AG, CT, ti, tv, tng2, nucF[4];
  bool useK2P = false;
  tng2 = 2. / (nucFreq[0] + nucFreq[1] + nucFreq[2] + nucFreq[3]);
  auxd = 1. / (nucFreq[0] + nucFreq[1] + nucFreq[2] + nucFreq[3]);
  for (int k = 0; k < 4; k++) nucF[k] = nucFreq[k] * auxd;
  fR = nucF[0] + nucF[2];
  fY = nucF[1] + nucF[3];
  if (nucFreq[0] == 0 || nucFreq[1] == 0 || nucFreq[2] == 0 || nucFreq[3] == 0)
    useK2P = true;
  else {
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
    return (AG > 0. && CT > 0.) ? -0.5 * log(AG) - 0.25 * log(CT) : 1000.0;
  }
  AG = 1. - AG / K1 - 0.5 * tv / fR;
  CT = 1. - CT / K2 - 0.5 * tv / fY;
  tv = 1. - 0.5 * tv / fY / fR;
  return (AG > 0. && CT > 0. && tv > 0) ? -K1 * log(AG) - K2 * log(CT) - K3 * log(tv) : 1000.0;
}

static int host_hist_index(double d)
{
  double f = floor(d * 200.0);
  if (f >= 0.0 && f < 200.0) return (int)f;
  return 199;
}

constexpr size_t STRIPE_BYTES = (size_t)NSTRIPE * 32 * 4;
constexpr size_t COUNTER_STRIDE = STRIPE_BYTES + 256;  // pinned bytes per job: the stripes (replay total included)

// One run: a set, its parameters, its buffers and what the host learns about it along the way.
struct Job {
  const tn93g_set *S = nullptr;
  tn93g_params P{};
  RunBufs *B = nullptr;
  RunArgs A{};
  bool fast = false;
  bool sparse_gaps = true;  // the marginal form's gap-run corrections cost less than counting the base sums
  int64_t nrows = 0, ntiles = 0, gx = 0, gy = 0, npairs = 0;
  int rp = 4;  // rows per wave of the pair kernel's tiles (tiles of 4 rp rows)
  size_t nr1 = 0;
  unsigned link_cap = 1, exact_cap = 1;
  uint64_t n_slow = 0, n_unc = 0, n_undecided = 0, n_linkrec = 0;
  unsigned max_region = 0, n_exact = 0;
  int64_t ra = -1, rb = -1;  // a part of a split run: its rows [ra, rb) (the whole run's rows when ra < 0)
  bool gemm_counts = false;  // the pair kernel reads its clean/clean counts from int8 GEMMs (a single run)
};

// Pairs (i, j) with r0 <= i < r1, c0 <= j < c1 and j > i.
static int64_t tri_pairs(int64_t r0, int64_t r1, int64_t c0, int64_t c1)
{
  int64_t p = 0;
  const int64_t a = std::min(r1, c0);  // rows below c0 pair with every column
  if (a > r0) p += (a - r0) * (c1 - c0);
  const int64_t b0 = std::max(r0, c0), b1 = std::min(r1, c1);  // row i in [c0, c1) pairs with (i, c1)
  if (b1 > b0) p += (b1 - b0) * (c1 - 1) - (b1 * (b1 - 1) / 2 - b0 * (b0 - 1) / 2);
  return p;
}

// Checks the parameters, uploads the job's event tables and sizes its buffers: everything before its kernels.
// The entry of the marginal form of the pair kernel: a site where exactly one code is clean and the other is an
// event code (4..16, or the gap under GAPMM) has the clean base subtracted from its nf field, because that base is
// counted in the sequence's totals.  Such a site always adds at least 1 (x16) to that field (the clean row or column
// of the site's cells sums to 1, or the site resolves to the base's diagonal cell, 2), so the result stays a
// non-negative 16-bit field; false when it would not.  Order-dependent entries are folded the same way: their nf
// fields hold only the site's weight (event_entry), which the fold keeps consistent with the base sums.
static bool fold_entry(int a, int b, bool gapmm, uint32_t q[4])
{
  const bool ca = a < 4, cb = b < 4;
  if (ca == cb) return true;
  const int other = ca ? b : a;
  if (other == GAPC && !gapmm) return true;  // never an event: gap runs cover these sites
  const int x = ca ? a : b;
  uint32_t &w = q[x >> 1];
  const int sh = (x & 1) * 16;
  if (((w >> sh) & 0xffffu) < 16u) return false;
  w -= 16u << sh;
  return true;
}

// A fixed device region that a run's buffers are carved from in order (the no-allocation entry point).
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
};

static int job_prepare(tn93g_ctx *c, Job &J, size_t *dry_bytes = nullptr, Arena *arena = nullptr)
{
  // with dry_bytes: only the sizes of the run's buffers are added up, nothing is allocated or uploaded; with arena:
  // every buffer is carved from the arena (no allocation), and a run that does not fit fails
  auto need = [&](DevBuf &b, size_t bytes) -> int {
    if (dry_bytes) {
      *dry_bytes += std::max(bytes, (size_t)256);
      return 0;
    }
    if (arena) {
      const size_t want = (std::max(bytes, (size_t)256) + 255) & ~(size_t)255;
      if (arena->used + want > arena->size)
        return c->fail(TN93G_E_ARG, fmt("the run needs more than the %zu bytes of scratch of the no-allocation entry "
                                        "point (%zu already placed, %zu more asked); use tn93g_run",
                                        arena->size, arena->used, want));
      b.p = arena->base + arena->used;
      b.n = want;
      arena->used += want;
      return 0;
    }
    return ensure(c, b, bytes);
  };
  tn93g_set *S = const_cast<tn93g_set *>(J.S);
  const tn93g_params *P = &J.P;
  SetDev d;
  int64_t nonunit;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    d = S->d;
    nonunit = S->nonunit;  // counts later appends too: only ever conservative
  }
  RunBufs *B = J.B;
  if (P->mode < 0 || P->mode > 4 || P->kind < 0 || P->kind > 3) return c->fail(TN93G_E_ARG, "bad mode or kind");
  if (P->kind == TN93G_APPEND) {
    const int64_t end = P->n2 ? P->n2 : d.n;
    if (P->n1 < 0 || P->n1 > end || end > d.n) return c->fail(TN93G_E_ARG, "bad appended range n1..n2");
    d.n = end;  // the run sees the set as it was when [n1, n2) had just been appended
  } else if (P->kind != TN93G_TRIANGLE && (P->n1 < 0 || P->n1 > d.n))
    return c->fail(TN93G_E_ARG, "bad n1");
  RunArgs &A = J.A;
  memset(&A, 0, sizeof A);
  A.planes = d.planes;
  A.onehot = d.onehot;
  A.codes = d.codes;
  A.codesT = d.codesT;
  A.ev_off = P->mode == TN93G_GAPMM ? d.evg_off : d.ev_off;
  A.ev = P->mode == TN93G_GAPMM ? d.evg : d.ev;
  // dense N: the ambiguity lists leave N out; GAPMM's lists (with gaps) keep it
  A.nplane = d.ndense && P->mode != TN93G_GAPMM ? 1 : 0;
  A.nmask = d.nmask;
  A.nrun_off = d.nrun_off;
  A.nruns = d.nruns;
  A.pre = d.pre;
  A.tb = d.tb;
  A.npre = d.npre;
  A.grun_off = d.grun_off;
  A.gruns = d.gruns;
  // The marginal form of the pair kernel trades 13 of 24 operations per word pair for, per pair, two prefix
  // look-ups (scattered loads) per gap run of either sequence; measured on the MI355X one gap run costs about as
  // much as 105 words of counting, so it is used when 2 x 105 x (gap runs per sequence) < W.  Under GAPMM gaps are
  // events and cost nothing extra.
  J.sparse_gaps = P->mode == TN93G_GAPMM || 210 * d.total_gruns < d.W * std::max<int64_t>(d.n, 1);
  A.weights = d.weights;
  A.unitw = nonunit == 0 ? 1 : 0;
  A.n = d.n;
  A.Ls = d.Ls;
  A.npad = d.npad;
  A.W = d.W;
  A.W64 = d.W64;
  if (P->kind == TN93G_CROSS) {
    A.r0 = 0;
    A.r1 = P->n1;
    A.c0 = P->n1;
    A.c1 = d.n;
    A.triangle = 0;
  } else if (P->kind == TN93G_APPEND) {  // pairs i < j with j in [n1, n): rows [0, n), columns [n1, n)
    A.r0 = 0;
    A.r1 = d.n;
    A.c0 = P->n1;
    A.c1 = d.n;
    A.triangle = 1;
  } else {
    A.r0 = 0;
    A.r1 = d.n;
    A.c0 = 0;
    A.c1 = d.n;
    A.triangle = 1;
  }
  if (J.ra >= 0) {  // one part of a split run: its rows (a triangle's part needs only the columns from its first row)
    if (J.ra < A.r0 || J.rb > A.r1 || J.ra >= J.rb) return c->fail(TN93G_E_ARG, "bad row range");
    A.r0 = J.ra;
    A.r1 = J.rb;
    if (P->kind == TN93G_TRIANGLE) A.c0 = J.ra;
  }
  A.fst = P->kind == TN93G_FST;
  A.n1 = (P->kind == TN93G_TRIANGLE || P->kind == TN93G_APPEND) ? d.n : P->n1;
  A.ncols = A.c1 - A.c0;
  const int64_t nrows = A.r1 - A.r0;
  J.nrows = nrows;
  A.gapmm = P->mode == TN93G_GAPMM;
  A.skip_events = P->mode == TN93G_SKIP;
  A.avg_fallback = (P->mode == TN93G_RESOLVE || P->mode == TN93G_SUBSET) && P->resolve_fraction < 1.;
  A.exact = (P->flags & TN93G_FLAG_EXACT) ? 1 : 0;
  A.all_slow = (P->flags & TN93G_FLAG_ALL_SLOW) ? 1 : 0;
  A.rf = P->resolve_fraction;
  A.thr = P->threshold;
  A.mind = P->min_distance;
  A.window = TN93G_LOG_WINDOW;
  A.min_overlap = P->min_overlap;
  A.kt = d.kt;
  for (int b = 0; b < 4; b++) A.kc[b] = d.kc[b];
  A.c16 = 32 * (A.Ls + A.kt) < 65536 ? 1 : 0;
  A.cstride = (A.c16 ? 4 : 7) * 64;
  const int64_t npairs_box = nrows * A.ncols;
  J.npairs = A.triangle ? tri_pairs(A.r0, A.r1, A.c0, A.c1) : npairs_box;
  if (J.npairs >= (1ll << 32)) return c->fail(TN93G_E_ARG, "more than 2^32 pairs in one run");
  // tables: [4][18*18] packed event entries (main mode, AVERAGE; then both folded), [2][18*18] replay entries
  if (!dry_bytes && (B->tab_mode != P->mode || B->tab_mask != P->subset_mask || !B->evtab.p || !B->rtab.p)) {
    // built on the stack (31 KB): copied in the context's stream order, which is synchronised before they go
    std::array<uint4, 4 * NCODE * NCODE> et;
    std::array<ReplayEnt, 2 * NCODE * NCODE> rt;
    bool fold_ok = true, gcert = true;
    for (int a = 0; a < NCODE; a++)
      for (int b = 0; b < NCODE; b++) {
        SiteEffect E = site_effect(a, b, P->mode, P->subset_mask);
        for (int t = 0; t < 2; t++) {
          SiteEffect F = t ? site_effect(a, b, TN93G_AVERAGE, 0) : E;
          uint32_t q[4];
          const bool unit = event_entry(F, q, a, b);
          if (t == 0) gcert = gcert && unit;
          et[(size_t)(t * NCODE * NCODE + a * NCODE + b)] = make_uint4(q[0], q[1], q[2], q[3]);
          fold_ok = fold_entry(a, b, P->mode == TN93G_GAPMM, q) && fold_ok;
          et[(size_t)((2 + t) * NCODE * NCODE + a * NCODE + b)] = make_uint4(q[0], q[1], q[2], q[3]);
          ReplayEnt &r = rt[(size_t)(t * NCODE * NCODE + a * NCODE + b)];
          r.v = F.v;
          r.mask = F.mask;
          r.icd = (int16_t)F.icd;
          r.amb = (uint16_t)F.amb;
        }
      }
    // the replay applies a run of N sites facing clean bases cell by cell: (N, x) may mark cells of column x only,
    // (x, N) cells of row x only (true of tn93's handler in every mode)
    for (int t = 0; t < 2; t++)
      for (int x = 0; x < 4; x++) {
        const uint32_t mc = rt[(size_t)(t * NCODE * NCODE + NC * NCODE + x)].mask;
        const uint32_t mr = rt[(size_t)(t * NCODE * NCODE + x * NCODE + NC)].mask;
        if ((mc & ~(0x1111u << x)) || (mr & ~(0xfu << (4 * x))))
          return c->fail(TN93G_E_ARG, "replay table: an N effect outside its row or column");
      }
    int rc;
    if ((rc = need(B->evtab, et.size() * sizeof(uint4))) || (rc = need(B->rtab, rt.size() * sizeof(ReplayEnt))))
      return rc;
    HIPCHK(c, hipMemcpyAsync(B->evtab.p, et.data(), et.size() * sizeof(uint4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(B->rtab.p, rt.data(), rt.size() * sizeof(ReplayEnt), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    B->tab_mode = P->mode;
    B->tab_mask = P->subset_mask;
    B->tab_fold = fold_ok;
    B->tab_gcert = gcert;
  }
  A.gcert = B->tab_gcert ? 1 : 0;
  {
    const double Lf = 32.0 * (double)A.W + (double)A.kt;
    A.gE = (16.0 * Lf * Lf + 64.0 * Lf) * 0x1p-52;
  }
  A.evtab = (const uint4 *)B->evtab.p;
  A.rtab = (const ReplayEnt *)B->rtab.p;
  // buffers
  J.fast = !A.exact && P->kind != TN93G_FST;  // F_ST combines the means non-monotonically: exact path
  A.fast = J.fast ? 1 : 0;
  const size_t nbox = (size_t)std::max<int64_t>(npairs_box, 1);
  J.nr1 = (size_t)std::max<int64_t>(nrows, 1);
  const size_t nr1 = J.nr1;
  const int64_t c0a = A.c0 & ~(int64_t)63;
  J.gx = std::max<int64_t>((A.c1 - c0a + TJ - 1) / TJ, 1);
  // tile rows: the caller's (8, 16, 32), else 16 until the autotuner or its earlier choice sets them
  if (P->tile_rows != 0 && P->tile_rows != 8 && P->tile_rows != 16 && P->tile_rows != 32)
    return c->fail(TN93G_E_ARG, "tile_rows must be 0, 8, 16 or 32");
  if (P->tile_rows != 0 && P->tile_rows != 16 && P->backend == TN93G_BACKEND_MFMA)
    return c->fail(TN93G_E_ARG, "the MFMA backend tiles 16 rows");
  J.rp = P->tile_rows ? P->tile_rows / 4 : 4;
  J.gy = std::max<int64_t>((nrows + 4 * J.rp - 1) / (4 * J.rp), 1);
  J.ntiles = J.gx * J.gy;
  const int64_t nslots = J.fast ? (int64_t)nr1 * J.gx : 1;
  if (J.gx >= (1ll << 31) || J.gy >= 65536) return c->fail(TN93G_E_ARG, "too many pair tiles in one run");
  if ((!J.fast && (need(B->D, nbox * 8) || need(B->U, nbox * 2))) ||
      need(B->hist, (size_t)NSTRIPE * 600 * 8) || need(B->row_links, nr1 * 4) ||
      need(B->link_off, (size_t)(nrows + 1) * 8) ||
      need(B->slow_list, (size_t)std::max<int64_t>(J.npairs, 1) * 8) ||
      need(B->counters, STRIPE_BYTES) ||
      need(B->rowout, nr1 * std::max<size_t>(15 * 8 + 16, sizeof(RowAcc))) ||
      need(B->chdr, (size_t)nslots * sizeof(ChunkHdr)) ||
      need(B->cdata, (size_t)nslots * (A.c16 ? 4 : 7) * 64 * sizeof(uint32_t)))
    return arena ? TN93G_E_ARG : TN93G_E_NOMEM;
  // host-exact list: only the pairs with an uncertain logarithm are stored; it starts at 2^20 records (72 MB) or what
  // sufficed last time, and a run that overflows it is repeated with the list it needed
  // (TN93G_EXACT_LIST sets the starting size, for tests of the growth)
  const char *es = getenv("TN93G_EXACT_LIST");
  const int64_t start = es ? std::max<int64_t>(1, atoll(es)) : (1 << 20);
  J.exact_cap = A.exact ? (unsigned)std::min<int64_t>({J.npairs, 1ll << 31, std::max<int64_t>(B->exact_hint, start)})
                        : 1u;
  if (need(B->exact_list, (size_t)J.exact_cap * sizeof(ExactRec))) return arena ? TN93G_E_ARG : TN93G_E_NOMEM;
  J.link_cap = 1;
  if (J.fast) {
    // links per stripe region; a run that overflows a region is repeated with larger regions (in an arena: the
    // regions take what is left of it, up to the usual size, and an overflow is an error)
    const size_t per_link = NSTRIPE * (2 * sizeof(LinkRec) + sizeof(uint32_t)), slack = 3 * 256;
    int64_t want = std::max<int64_t>(B->link_cap_hint, 1 << 15);
    if (arena) {
      const size_t left = arena->size - arena->used;
      want = std::min<int64_t>(want, left > slack ? (int64_t)((left - slack) / per_link) : 0);
      if (want < 1) return need(B->links, per_link + slack + left);  // fails with the arena's message
    }
    J.link_cap = (unsigned)std::min<int64_t>(std::max<int64_t>(J.npairs, 1), want);
    if (need(B->links, (size_t)J.link_cap * NSTRIPE * sizeof(LinkRec)) ||
        need(B->links_flat, (size_t)J.link_cap * NSTRIPE * sizeof(LinkRec)) ||
        need(B->lslot, (size_t)J.link_cap * NSTRIPE * sizeof(uint32_t)))
      return arena ? TN93G_E_ARG : TN93G_E_NOMEM;
  }
  A.D = (double *)B->D.p;
  A.U = (uint16_t *)B->U.p;
  A.hist = (unsigned long long *)B->hist.p;
  A.row_links = (int32_t *)B->row_links.p;
  A.slow_list = (unsigned long long *)B->slow_list.p;
  A.slow_total = (unsigned long long *)((unsigned *)B->counters.p + ST_SLOW);
  A.stripes = (unsigned *)B->counters.p;
  A.exact_cnt = A.stripes + ST_EXACT;
  A.exact_list = (ExactRec *)B->exact_list.p;
  A.exact_cap = J.exact_cap;
  A.racc = (RowAcc *)B->rowout.p;
  A.links = (LinkRec *)B->links.p;
  A.links_flat = (LinkRec *)B->links_flat.p;
  A.link_off = (int64_t *)B->link_off.p;
  A.lslot = (uint32_t *)B->lslot.p;
  A.link_cap = J.link_cap;
  A.chdr = (ChunkHdr *)B->chdr.p;
  A.cdata = (uint32_t *)B->cdata.p;
  A.gx = J.gx;
  A.ti = 4 * J.rp;
  // count tables of the special sequences for the replay: rows against the run's columns and, unless the run is a
  // triangle over one range (then a column's table is its row table), columns against the run's rows.  Long
  // alignments only (128 words or more): below that the replay's own per-pair count is cheaper than 16 gathers.
  A.spec_row = A.spec_col = nullptr;
  A.tab_row = A.tab_col = nullptr;
  A.spec_rows = nullptr;
  A.nsr = A.nsc = 0;
  if (J.fast && !A.exact && !A.all_slow && P->mode != TN93G_SKIP && nrows > 0 && A.ncols > 0 && A.W >= 128 &&
      !arena) {  // (in an arena the replay counts per pair: the tables' host index lists would allocate)
    const bool same = A.r0 == A.c0 && A.r1 == A.c1;
    std::vector<int32_t> &idx = B->h_spec_idx;
    std::vector<int64_t> &ids = B->h_spec_ids;
    idx.assign((size_t)(nrows + (same ? 0 : A.ncols)), -1);
    ids.clear();
    for (int64_t r = 0; r < nrows; r++)
      if (S->spec[(size_t)(A.r0 + r)]) {
        idx[(size_t)r] = (int32_t)ids.size();
        ids.push_back(A.r0 + r);
      }
    const int64_t nsr = (int64_t)ids.size();
    if (!same)
      for (int64_t q = 0; q < A.ncols; q++)
        if (S->spec[(size_t)(A.c0 + q)]) {
          idx[(size_t)(nrows + q)] = (int32_t)(ids.size() - (size_t)nsr);
          ids.push_back(A.c0 + q);
        }
    const int64_t nsc = (int64_t)ids.size() - nsr;
    const size_t cells = ((size_t)nsr * (size_t)A.ncols + (size_t)nsc * (size_t)nrows) * 16;
    // beyond 8 GB (or 2^20 special sequences on a side) the replay counts per pair instead
    if (nsr + nsc > 0 && cells * 4 <= ((size_t)8 << 30) && nsr < 65535 * TI && nsc < 65535 * TI) {
      if (need(B->spec_idx, idx.size() * 4) || need(B->spec_ids, ids.size() * 8) || need(B->spec_tab, cells * 4))
        return TN93G_E_NOMEM;
      if (!dry_bytes) {
        HIPCHK(c, hipMemcpyAsync(B->spec_idx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(B->spec_ids.p, ids.data(), ids.size() * 8, hipMemcpyHostToDevice, c->stream));
        const int32_t *ix = (const int32_t *)B->spec_idx.p;
        const uint32_t *tb = (const uint32_t *)B->spec_tab.p;
        A.spec_row = ix;
        A.spec_col = same ? ix : ix + nrows;
        A.tab_row = tb;
        A.tab_col = same ? tb : tb + (size_t)nsr * (size_t)A.ncols * 16;
        A.spec_rows = (const int64_t *)B->spec_ids.p;
        A.nsr = (int)nsr;
        A.nsc = (int)nsc;
      }
    }
  }
  return 0;
}

// The pair-kernel variant of a job: 0-7 the popcount backend (bit 0 fast path, bit 1 dense-N planes, bit 2 marginal
// form), 8 / 9 the MFMA backend (fast / exact path), 10 / 11 the same for dense-N sets; -1 when the run has no pair.
static int pairs_variant(const Job &J)
{
  const RunArgs &A = J.A;
  if (J.nrows <= 0 || A.ncols <= 0) return -1;
  if (J.P.backend == TN93G_BACKEND_MFMA) return (J.fast ? 8 : 9) + (A.nplane ? 2 : 0);
  if (J.gemm_counts) return 12;
  const bool marg = J.B->tab_fold && !A.skip_events && J.sparse_gaps;
  return (J.fast ? 1 : 0) | (A.nplane ? 2 : 0) | (marg ? 4 : 0);
}

// The popcount fast-path variants (1, 3, 5, 7) come in tiles of 8, 16 and 32 rows (rp 2, 4, 8); the exact-path
// variants and the MFMA backend in tiles of 16.
template <int RP>
static void launch_pairs_rp(int v, dim3 g, hipStream_t st, const RunArgs &A)
{
  constexpr int PC = TN93G_BACKEND_POPC;
  switch (v) {
    case 1: hipLaunchKernelGGL((k_pairs<PC, true, false, false, RP>), g, dim3(256), 0, st, A); break;
    case 3: hipLaunchKernelGGL((k_pairs<PC, true, true, false, RP>), g, dim3(256), 0, st, A); break;
    case 5: hipLaunchKernelGGL((k_pairs<PC, true, false, true, RP>), g, dim3(256), 0, st, A); break;
    case 7: hipLaunchKernelGGL((k_pairs<PC, true, true, true, RP>), g, dim3(256), 0, st, A); break;
    default: break;
  }
}
static void launch_pairs(int v, int rp, dim3 g, hipStream_t st, const RunArgs &A)
{
  constexpr int PC = TN93G_BACKEND_POPC, MF = TN93G_BACKEND_MFMA;
  if (rp == 2) return launch_pairs_rp<2>(v, g, st, A);
  if (rp == 8) return launch_pairs_rp<8>(v, g, st, A);
  switch (v) {
    case 0: hipLaunchKernelGGL((k_pairs<PC, false, false, false, 4>), g, dim3(256), 0, st, A); break;
    case 1: hipLaunchKernelGGL((k_pairs<PC, true, false, false, 4>), g, dim3(256), 0, st, A); break;
    case 2: hipLaunchKernelGGL((k_pairs<PC, false, true, false, 4>), g, dim3(256), 0, st, A); break;
    case 3: hipLaunchKernelGGL((k_pairs<PC, true, true, false, 4>), g, dim3(256), 0, st, A); break;
    case 4: hipLaunchKernelGGL((k_pairs<PC, false, false, true, 4>), g, dim3(256), 0, st, A); break;
    case 5: hipLaunchKernelGGL((k_pairs<PC, true, false, true, 4>), g, dim3(256), 0, st, A); break;
    case 6: hipLaunchKernelGGL((k_pairs<PC, false, true, true, 4>), g, dim3(256), 0, st, A); break;
    case 7: hipLaunchKernelGGL((k_pairs<PC, true, true, true, 4>), g, dim3(256), 0, st, A); break;
    case 8: hipLaunchKernelGGL((k_pairs<MF, true, false, false, 4>), g, dim3(256), 0, st, A); break;
    case 9: hipLaunchKernelGGL((k_pairs<MF, false, false, false, 4>), g, dim3(256), 0, st, A); break;
    case 10: hipLaunchKernelGGL((k_pairs<MF, true, true, false, 4>), g, dim3(256), 0, st, A); break;
    case 11: hipLaunchKernelGGL((k_pairs<MF, false, true, false, 4>), g, dim3(256), 0, st, A); break;
    case 12: hipLaunchKernelGGL((k_pairs<BACKEND_GEMMC, true, true, true, 4>), g, dim3(256), 0, st, A); break;
    default: break;
  }
}

// One launch over the tiles of several jobs of fast-path variant v and tile rows 4 rp (a batch); false for other
// variants.
template <int RP>
static bool launch_pairs_multi_rp(int v, const dim3 g, hipStream_t st, const RunArgs *jobs, const int64_t *toff,
                                  int nj)
{
  constexpr int PC = TN93G_BACKEND_POPC, MF = TN93G_BACKEND_MFMA;
  const dim3 b(256);
  switch (v) {
    case 1: hipLaunchKernelGGL((k_pairs_multi<PC, true, false, false, RP>), g, b, 0, st, jobs, toff, nj); return true;
    case 3: hipLaunchKernelGGL((k_pairs_multi<PC, true, true, false, RP>), g, b, 0, st, jobs, toff, nj); return true;
    case 5: hipLaunchKernelGGL((k_pairs_multi<PC, true, false, true, RP>), g, b, 0, st, jobs, toff, nj); return true;
    case 7: hipLaunchKernelGGL((k_pairs_multi<PC, true, true, true, RP>), g, b, 0, st, jobs, toff, nj); return true;
    case 8:
    case 10:
      if constexpr (RP == 4) {
        if (v == 8)
          hipLaunchKernelGGL((k_pairs_multi<MF, true, false, false, 4>), g, b, 0, st, jobs, toff, nj);
        else
          hipLaunchKernelGGL((k_pairs_multi<MF, true, true, false, 4>), g, b, 0, st, jobs, toff, nj);
        return true;
      }
      return false;
    default: return false;
  }
}
static bool launch_pairs_multi(int v, int rp, unsigned nblocks, hipStream_t st, const RunArgs *jobs,
                               const int64_t *toff, int nj)
{
  const dim3 g(nblocks);
  if (rp == 2) return launch_pairs_multi_rp<2>(v, g, st, jobs, toff, nj);
  if (rp == 8) return launch_pairs_multi_rp<8>(v, g, st, jobs, toff, nj);
  return launch_pairs_multi_rp<4>(v, g, st, jobs, toff, nj);
}

// Sets a job's pair-kernel shape: backend and tile rows (4 rp), with the grid that follows.  The exact path and the
// MFMA backend take tiles of 16 rows.
static int job_set_shape(tn93g_ctx *c, Job &J, int backend, int rp)
{
  if (backend == TN93G_BACKEND_MFMA || !J.fast) rp = 4;
  J.P.backend = backend;
  J.rp = rp;
  J.gy = std::max<int64_t>((J.nrows + 4 * rp - 1) / (4 * rp), 1);
  J.ntiles = J.gx * J.gy;
  J.A.ti = 4 * rp;
  if (J.gy >= 65536) return c->fail(TN93G_E_ARG, "too many pair tiles in one run");
  return 0;
}

// A job's kernels in three parts on stream `st`, with no host synchronisation (a batch captures them into one graph):
// before the pair kernel (fills, special-sequence tables), the pair kernel, and after it.  The replay and the tail
// kernel depend only on the pair kernel: the tail runs on the branch's side stream beside the replay (both fold
// their results in with atomics), joined back before the link gather.  With `timing`, events around the pair and
// tail kernels give the device times of a single run.
// The buffers a run starts from zero (histogram replicas, counters, per-row link counts, row accumulators).
static int job_fill_list(const Job &J, FillList &L)
{
  RunBufs *B = J.B;
  L.r[0] = FillRange{(uint32_t *)B->hist.p, (int64_t)NSTRIPE * 600 * 2};
  L.r[1] = FillRange{(uint32_t *)B->counters.p, (int64_t)(STRIPE_BYTES / 4)};
  L.r[2] = FillRange{(uint32_t *)B->row_links.p, (int64_t)J.nr1};
  if (!J.fast) return 3;
  L.r[3] = FillRange{(uint32_t *)B->rowout.p, (int64_t)(J.nr1 * sizeof(RowAcc) / 4)};
  return 4;
}
static void job_fill_ranges(const Job &J, std::vector<FillRange> &out)
{
  FillList L{};
  const int nf = job_fill_list(J, L);
  for (int k = 0; k < nf; k++) out.push_back(L.r[k]);
}

// The special-sequence tables of a single run as int8 GEMMs on the matrix cores (rocBLAS): table [s][a][x][b] is
// entry (4 x + b, 4 s + a) of the column-major product X^T Y, X the one-hot rows of the other side's sequences and Y
// those of the special ones ([4 q + base][site], site-contiguous, so both operands are K-major: a "TN" product), with
// int32 accumulation: exact counts of the clean/clean sites.  False (nothing enqueued) when the operands cannot be
// allocated or the library cannot be initialised: the caller then runs k_spec16.
static bool spec_gemm_one(tn93g_ctx *c, const RunArgs &A, const int64_t *ids, int64_t ns, int64_t x0, int64_t nx,
                          uint32_t *T, hipStream_t st)
{
  const int64_t Ls = A.Ls;  // a multiple of 16; sites from L on are gaps, zero in every one-hot row
  int8_t *X = (int8_t *)c->gemm_ops, *Y = X + 4 * nx * Ls;
  hipLaunchKernelGGL(k_onehot_rows, dim3((unsigned)((nx * (Ls / 16) + 255) / 256)), dim3(256), 0, st, A.codes, Ls,
                     (const int64_t *)nullptr, x0, nx, X);
  hipLaunchKernelGGL(k_onehot_rows, dim3((unsigned)((ns * (Ls / 16) + 255) / 256)), dim3(256), 0, st, A.codes, Ls,
                     ids, (int64_t)0, ns, Y);
  const int32_t one = 1, zero = 0;
  return rocblas_gemm_ex(c->blas, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)(4 * nx),
                         (rocblas_int)(4 * ns), (rocblas_int)Ls, &one, X, rocblas_datatype_i8_r, (rocblas_int)Ls, Y,
                         rocblas_datatype_i8_r, (rocblas_int)Ls, &zero, T, rocblas_datatype_i32_r,
                         (rocblas_int)(4 * nx), T, rocblas_datatype_i32_r, (rocblas_int)(4 * nx),
                         rocblas_datatype_i32_r, rocblas_gemm_algo_standard, 0, 0) == rocblas_status_success;
}
// The standard streams' formatting (flags, precision, width, fill), saved on construction and put back on
// destruction: rocBLAS changes the formatting of std::cout when it is set up, and the caller's output (tn93's
// JSON, written with the default formatting) must not see that.
struct StreamFormatGuard {
  std::ios *s[3] = {&std::cout, &std::cerr, &std::clog};
  std::ios::fmtflags f[3];
  std::streamsize p[3], w[3];
  char fill[3];
  StreamFormatGuard()
  {
    for (int k = 0; k < 3; k++) f[k] = s[k]->flags(), p[k] = s[k]->precision(), w[k] = s[k]->width(), fill[k] = s[k]->fill();
  }
  ~StreamFormatGuard()
  {
    for (int k = 0; k < 3; k++) s[k]->flags(f[k]), s[k]->precision(p[k]), s[k]->width(w[k]), s[k]->fill(fill[k]);
  }
};

static bool spec_gemm(tn93g_ctx *c, const Job &J, hipStream_t st)
{
  const StreamFormatGuard keep_format;
  const RunArgs &A = J.A;
  const bool cols = A.nsc > 0;  // a triangle over one range has none: its column tables are its row tables
  const int64_t rows_side = 4 * (A.ncols + A.nsr) * A.Ls, cols_side = cols ? 4 * (J.nrows + A.nsc) * A.Ls : 0;
  // GEMM dimensions are 32-bit: 4 x (sequences) and the sites
  if (4 * std::max(A.ncols, J.nrows) >= (1ll << 31) || A.Ls >= (1ll << 31)) return false;
  const size_t need_bytes = (size_t)std::max(rows_side, cols_side);
  if (c->gemm_ops_n < need_bytes) {
    if (c->gemm_ops) (void)hipFree(c->gemm_ops);
    c->gemm_ops = nullptr;
    c->gemm_ops_n = 0;
    if (hipMalloc(&c->gemm_ops, need_bytes) != hipSuccess) {
      (void)hipGetLastError();
      c->gemm_ops = nullptr;
      return false;
    }
    c->gemm_ops_n = need_bytes;
  }
  if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success) {
    c->blas = nullptr;
    return false;
  }
  if (rocblas_set_stream(c->blas, st) != rocblas_status_success) return false;
  if (A.nsr && !spec_gemm_one(c, A, A.spec_rows, A.nsr, A.c0, A.ncols, const_cast<uint32_t *>(A.tab_row), st))
    return false;
  // the second product reuses the operand buffer after the first on the same stream
  if (cols && !spec_gemm_one(c, A, A.spec_rows + A.nsr, A.nsc, A.r0, J.nrows, const_cast<uint32_t *>(A.tab_col), st))
    return false;
  return true;
}

// The pair kernel's clean/clean counts as four int8 GEMMs (rocBLAS) over one-hot operands of the run's sequences
// (k_onehot_counts): product k is [ncols][nrows]-shaped, entry (j - c0) + (i - r0) ncols the dot product of column
// j's and row i's operand rows, so that a wave's row reads its 64 columns' counts contiguously.  The rows and columns
// of a run lie in [lo, hi), whose operands are built once.  A triangle's products are computed in blocks of rows, each
// against the columns from the block's first row on.  False, with nothing of the counts enqueued, when memory or rocBLAS is not available: the
// pair kernel then counts with popcount.
static bool gemm_counts(tn93g_ctx *c, Job &J, hipStream_t st)
{
  const StreamFormatGuard keep_format;
  RunArgs &A = J.A;
  const int64_t lo = std::min(A.r0, A.c0), hi = std::max(A.r1, A.c1), nq = hi - lo, Ls = A.Ls;
  if (2 * Ls >= (1ll << 31) || A.ncols >= (1ll << 31) || J.nrows >= (1ll << 31)) return false;
  const size_t ops = (size_t)nq * 6 * (size_t)Ls, out = (size_t)A.ncols * (size_t)J.nrows * 16;
  auto grow = [](void *&p, size_t &n, size_t want) {
    if (n >= want) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return false;
    }
    n = want;
    return true;
  };
  if (!grow(c->gemm_ops, c->gemm_ops_n, ops) || !grow(c->gemm_out, c->gemm_out_n, out)) return false;
  if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success) {
    c->blas = nullptr;
    return false;
  }
  if (rocblas_set_stream(c->blas, st) != rocblas_status_success) return false;
  int8_t *X[4];
  X[0] = (int8_t *)c->gemm_ops;
  X[1] = X[0] + nq * 2 * Ls;
  X[2] = X[1] + nq * 2 * Ls;
  X[3] = X[2] + nq * Ls;
  const int64_t K[4] = {2 * Ls, 2 * Ls, Ls, Ls};
  hipLaunchKernelGGL(k_onehot_counts, dim3((unsigned)((nq * (Ls / 16) + 255) / 256)), dim3(256), 0, st, A.codes, Ls,
                     lo, nq, X[0], X[1], X[2], X[3]);
  uint32_t *C = (uint32_t *)c->gemm_out;
  const int64_t gcs = A.ncols * J.nrows;
  const int32_t one = 1, zero = 0;
  // a triangle in blocks of rows, each against the columns from its first row on: (nb + 1) / 2nb of the square
  int nb = A.triangle ? (int)std::min<int64_t>(std::max<int64_t>(J.nrows / 4096, 1), 8) : 1;
  if (const char *e = getenv("TN93G_GEMM_BLOCKS")) nb = std::max(1, atoi(e));
  for (int b = 0; b < nb; b++) {
    const int64_t i0 = A.r0 + J.nrows * b / nb, i1 = A.r0 + J.nrows * (b + 1) / nb;
    const int64_t cs = A.triangle ? std::min(std::max(A.c0, i0), A.c1) : A.c0;
    if (i1 <= i0 || cs >= A.c1) continue;
    for (int k = 0; k < 4; k++)
      if (rocblas_gemm_ex(c->blas, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)(A.c1 - cs),
                          (rocblas_int)(i1 - i0), (rocblas_int)K[k], &one, X[k] + (cs - lo) * K[k],
                          rocblas_datatype_i8_r, (rocblas_int)K[k], X[k] + (i0 - lo) * K[k], rocblas_datatype_i8_r,
                          (rocblas_int)K[k], &zero, C + k * gcs + (i0 - A.r0) * A.ncols + (cs - A.c0),
                          rocblas_datatype_i32_r, (rocblas_int)A.ncols,
                          C + k * gcs + (i0 - A.r0) * A.ncols + (cs - A.c0), rocblas_datatype_i32_r,
                          (rocblas_int)A.ncols, rocblas_datatype_i32_r, rocblas_gemm_algo_standard, 0,
                          0) != rocblas_status_success)
        return false;
  }
  A.gcnt = C;
  A.gcs = gcs;
  return true;
}

// Whether a single run's pair kernel takes the GEMM counts: the fast dense-N marginal form of long alignments (the
// counts of 32 sites cost the popcount form 9 VALU operations per pair, 7e-15 s per pair and site; the six one-hot
// units of the products 12 int8 operations, ~5e-15 s at the ~2.4 POPS rocBLAS reaches on these shapes, with a
// triangle's blocks ~0.6 of the square), within 32 GB of operands and products.  From a context's second such run
// on (rocBLAS's set-up costs ~0.15 s per process, more than the products save on one 20,000 x 30,000-site triangle),
// or at once past 1e13 word pairs, unless the caller asked for a backend.  TN93G_COUNTS=gemm|popc forces one.
static bool use_gemm_counts(tn93g_ctx *c, const Job &J, int asked)
{
  const char *e = getenv("TN93G_COUNTS");
  if (e && !strcmp(e, "popc")) return false;
  const bool force = e && !strcmp(e, "gemm");
  const RunArgs &A = J.A;
  if (J.gemm_counts || pairs_variant(J) != 7 || A.all_slow) return false;
  if (asked != TN93G_BACKEND_AUTO && !force) return false;  // a backend asked for by name is kept
  const int64_t lo = std::min(A.r0, A.c0), hi = std::max(A.r1, A.c1);
  const double bytes = (double)(hi - lo) * 6.0 * (double)A.Ls + 16.0 * (double)A.ncols * (double)J.nrows;
  if (bytes > 32e9) return false;
  if (force) return true;
  if (A.W < 256) return false;
  const bool first = c->count_runs++ == 0;
  return c->blas || !first || (double)J.npairs * (double)A.W > 1e13;
}

// Whether a single run's tables come from rocBLAS: setting the library up costs about 0.15 s once per process, so a
// context takes the GEMM from its second run with tables on (a process that runs many), or at once when k_spec16
// would take longer than that (about 1e12 word pairs per second).  TN93G_TABLES=blas|kernel forces one (tests).
static bool use_gemm_tables(tn93g_ctx *c, const Job &J)
{
  const char *e = getenv("TN93G_TABLES");
  if (e && !strcmp(e, "kernel")) return false;
  const bool first = c->table_runs++ == 0;
  if (e && !strcmp(e, "blas")) return true;
  const RunArgs &A = J.A;
  const double word_pairs = ((double)A.nsr * (double)A.ncols + (double)A.nsc * (double)J.nrows) * (double)A.W;
  return c->blas || !first || word_pairs > 1.5e11;
}

// fills: false when a batch clears every job's buffers with one k_fill_ranges launch; blas: a single run (not
// captured into a batch graph), whose special-sequence tables may come from rocBLAS
static void job_enqueue_pre(tn93g_ctx *c, Job &J, bool timing, hipStream_t st, bool fills = true, bool blas = false)
{
  RunArgs &A = J.A;
  const int64_t nrows = J.nrows;
  if (timing) (void)hipEventRecord(c->ev[4], st);
  if (fills) {  // one launch for the (at most 4) ranges
    FillList L{};
    const int nf = job_fill_list(J, L);
    hipLaunchKernelGGL(k_fill_list, dim3(64, (unsigned)nf), dim3(256), 0, st, L);
  }
  if ((A.nsr || A.nsc) && blas && use_gemm_tables(c, J) && spec_gemm(c, J, st)) return;
  if (A.nsr)
    hipLaunchKernelGGL(k_spec16, dim3((unsigned)((A.ncols + TJ - 1) / TJ), (unsigned)((A.nsr + TI - 1) / TI)),
                       dim3(256), 0, st, A, A.spec_rows, A.nsr, A.c0, A.ncols, const_cast<uint32_t *>(A.tab_row));
  if (A.nsc)
    hipLaunchKernelGGL(k_spec16, dim3((unsigned)((nrows + TJ - 1) / TJ), (unsigned)((A.nsc + TI - 1) / TI)), dim3(256),
                       0, st, A, A.spec_rows + A.nsr, A.nsc, A.r0, nrows, const_cast<uint32_t *>(A.tab_col));
}

static void job_enqueue_pairs(tn93g_ctx *c, Job &J, bool timing, hipStream_t st)
{
  if (timing) (void)hipEventRecord(c->ev[0], st);
  if (J.gemm_counts && !gemm_counts(c, J, st)) J.gemm_counts = false;  // (the products are part of the pair time)
  const int v = pairs_variant(J);
  if (v >= 0) launch_pairs(v, J.rp, dim3((unsigned)J.gx, (unsigned)J.gy), st, J.A);
  if (timing) (void)hipEventRecord(c->ev[1], st);
}

// grid sizes of a job's k_tail and k_link_order (the kernels stride over the rows)
static unsigned tail_blocks(const Job &J)
{
  return J.fast && J.nrows > 0 && J.A.ncols > 0 ? (unsigned)std::min<int64_t>(J.nrows, 65535) : 0u;
}
static unsigned order_blocks(const Job &J) { return (unsigned)std::min<int64_t>(std::max<int64_t>(J.nrows, 1), 65535); }

static void job_enqueue_replay(const Job &J, hipStream_t st)
{
  if (J.A.nplane)  // 3072 blocks measured best for the dense replay (C2: 482 -> 470 ms per run), 2048 otherwise
    hipLaunchKernelGGL(k_replay<true>, dim3(3072), dim3(RB), 0, st, J.A);
  else
    hipLaunchKernelGGL(k_replay<false>, dim3(2048), dim3(RB), 0, st, J.A);
}

static void job_enqueue_post(tn93g_ctx *c, Job &J, bool timing, hipStream_t st, int branch)
{
  RunArgs &A = J.A;
  RunBufs *B = J.B;
  const int64_t nrows = J.nrows;
  // the replay is enqueued first so that its long, latency-bound waves are resident before the tail's blocks fill
  // the machine; the tail runs on the side stream and is joined back before the link gather
  // (with timing, the pair kernel's end event, just recorded on st, is the fork, and the tail's end event the join:
  // each event record is a marker in the queue, a few microseconds of the device's time)
  hipStream_t rs = c->rside[branch];
  if (!timing) (void)hipEventRecord(c->rfork[branch], st);
  (void)hipStreamWaitEvent(rs, timing ? c->ev[1] : c->rfork[branch], 0);
  job_enqueue_replay(J, st);
  if (tail_blocks(J)) {
    if (A.c16)
      hipLaunchKernelGGL(k_tail<true>, dim3(tail_blocks(J)), dim3(256), 0, rs, A);
    else
      hipLaunchKernelGGL(k_tail<false>, dim3(tail_blocks(J)), dim3(256), 0, rs, A);
  }
  hipEvent_t join = timing ? c->ev[3] : c->rjoin[branch];
  (void)hipEventRecord(join, rs);
  (void)hipStreamWaitEvent(st, join, 0);
  if (J.fast) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, (const int32_t *)B->row_links.p, A.link_off, nrows);
    hipLaunchKernelGGL(k_link_gather, dim3(NSTRIPE), dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_link_order_hist, dim3(order_blocks(J) + 3), dim3(256), 0, st, A, order_blocks(J));
  } else {
    hipLaunchKernelGGL(k_hist_sum, dim3(3), dim3(256), 0, st, A);
  }
}

static void job_enqueue(tn93g_ctx *c, Job &J, bool timing, hipStream_t st, int branch, bool fills = true,
                        bool blas = false)
{
  job_enqueue_pre(c, J, timing, st, fills, blas);
  job_enqueue_pairs(c, J, timing, st);
  job_enqueue_post(c, J, timing, st, branch);
}

// Device-to-host copy of the job's counters (the replay total among them) into `pin` (COUNTER_STRIDE bytes).
static int job_enqueue_counters(tn93g_ctx *c, Job &J, char *pin)
{
  HIPCHK(c, hipMemcpyAsync(pin, J.B->counters.p, STRIPE_BYTES, hipMemcpyDeviceToHost, c->stream));
  return 0;
}

static void job_parse_counters(Job &J, const char *pin)
{
  unsigned st[NSTRIPE * 32];
  memcpy(st, pin, STRIPE_BYTES);
  int64_t slow_total = 0;
  memcpy(&slow_total, pin + ST_SLOW * 4, 8);
  J.n_slow = (uint64_t)slow_total;
  J.n_unc = J.n_undecided = J.n_linkrec = 0;
  J.max_region = 0;
  for (int s = 0; s < NSTRIPE; s++) {
    J.n_unc += st[(size_t)s * 32 + ST_UNCERTAIN];
    J.n_undecided += st[(size_t)s * 32 + ST_UNDECIDED];
    J.n_linkrec += st[(size_t)s * 32 + ST_LINKS];
    J.max_region = std::max(J.max_region, st[(size_t)s * 32 + ST_LINKS]);
  }
  J.n_exact = st[ST_EXACT];
}

// Read-back layout of a fast-path job in pinned memory: row accumulators, histogram, then nl link records.
struct FastLayout {
  size_t off_ra, off_h, off_lk, bytes;
};
static FastLayout fast_layout(const Job &J, size_t nl)
{
  FastLayout F;
  F.off_ra = 0;
  F.off_h = (J.nr1 * sizeof(RowAcc) + 255) & ~(size_t)255;
  F.off_lk = F.off_h + ((600 * 8 + 255) & ~(size_t)255);
  F.bytes = F.off_lk + ((nl * sizeof(LinkRec) + 255) & ~(size_t)255);
  return F;
}

// Copies of the row accumulators, the histogram and the first nl link records (nl may exceed the job's count: the
// rest is not read).
static int job_enqueue_fast_readback(tn93g_ctx *c, Job &J, char *pin, size_t nl)
{
  const FastLayout F = fast_layout(J, nl);
  HIPCHK(c, hipMemcpyAsync(pin + F.off_ra, J.B->rowout.p, J.nr1 * sizeof(RowAcc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(pin + F.off_h, J.B->hist.p, 600 * 8, hipMemcpyDeviceToHost, c->stream));
  if (nl) HIPCHK(c, hipMemcpyAsync(pin + F.off_lk, J.B->links.p, nl * sizeof(LinkRec), hipMemcpyDeviceToHost,
                                   c->stream));
  return 0;
}

// A single run's counters and, for the fast path, its row accumulators, histogram and first nl link records, read
// back by one k_copy_list launch into the pinned buffer's device mapping; the copies of job_enqueue_counters and
// job_enqueue_fast_readback when the buffer has no device mapping.
// With R (a result whose link block is pinned, holding at least nl links), the links go straight into R's arrays
// instead of the pinned buffer.
static int job_enqueue_readback(tn93g_ctx *c, Job &J, char *pin, size_t nl, tn93g_result *R = nullptr)
{
  char *d = nullptr;
  if (hipHostGetDevicePointer((void **)&d, pin, 0) != hipSuccess || !d) {
    (void)hipGetLastError();
    if (R) return TN93G_E_HIP;  // (the caller checked R's mapping first)
    int rc = job_enqueue_counters(c, J, pin);
    if (!rc && J.fast) rc = job_enqueue_fast_readback(c, J, pin + COUNTER_STRIDE, nl);
    return rc;
  }
  CopyList L{};
  int n = 0;
  L.r[n++] = CopyRange{(const uint2 *)J.B->counters.p, (uint2 *)d, (int64_t)(STRIPE_BYTES / 8)};
  bool big_rows = false;  // row sums of more than 4 MB (large APPEND runs) go by the copy engine, which moves them faster
  if (J.fast) {
    const FastLayout F = fast_layout(J, nl);
    char *f = d + COUNTER_STRIDE;
    big_rows = J.nr1 * sizeof(RowAcc) > ((size_t)4 << 20);
    if (!big_rows)
      L.r[n++] = CopyRange{(const uint2 *)J.B->rowout.p, (uint2 *)(f + F.off_ra), (int64_t)(J.nr1 * sizeof(RowAcc) / 8)};
    L.r[n++] = CopyRange{(const uint2 *)J.B->hist.p, (uint2 *)(f + F.off_h), 600};
    if (nl && R) {
      int64_t *row = nullptr;
      HIPCHK(c, hipHostGetDevicePointer((void **)&row, R->link_row, 0));
      const int64_t cap = R->link_col - R->link_row;  // the block's capacity (arrays of cap entries, in order)
      L.lk = (const LinkRec *)J.B->links.p;
      L.row = row;
      L.col = row + cap;
      L.d = (double *)(row + 2 * cap);
      L.lo = (double *)(row + 3 * cap);
      L.hi = (double *)(row + 4 * cap);
      L.nlk = (int64_t)nl;
    } else if (nl) {
      L.r[n++] = CopyRange{(const uint2 *)J.B->links.p, (uint2 *)(f + F.off_lk), (int64_t)(nl * sizeof(LinkRec) / 8)};
    }
  }
  hipLaunchKernelGGL(k_copy_list, dim3(64, (unsigned)n + (L.nlk ? 1u : 0u)), dim3(256), 0, c->stream, L, n);
  HIPCHK(c, hipGetLastError());
  if (big_rows)
    HIPCHK(c, hipMemcpyAsync(pin + COUNTER_STRIDE + fast_layout(J, nl).off_ra, J.B->rowout.p, J.nr1 * sizeof(RowAcc),
                             hipMemcpyDeviceToHost, c->stream));
  return 0;
}

// What every result gets at the end: histogram replicas, self_v_self entries, diagnostics.
static void job_finish_common(const Job &J, const unsigned long long *hrep, tn93g_result *R)
{
  for (int k = 0; k < 600; k++) (&R->hist[0][0])[k] += hrep[k];  // replicas summed by k_hist_sum
  R->exact = (!J.fast && J.n_unc == 0) ? 1 : 0;
  R->nrows = (int32_t)J.nrows;
  R->n_replayed = (int64_t)J.n_slow;
  R->n_uncertain = (int64_t)J.n_unc;
  R->n_host_exact = J.n_exact;
  R->backend_used = J.gemm_counts                         ? TN93G_BACKEND_GEMM
                    : J.P.backend == TN93G_BACKEND_AUTO ? TN93G_BACKEND_POPC
                                                        : J.P.backend;
  R->folded_columns = J.A.kt;
  R->tile_rows_used = 4 * J.rp;
  // self_v_self histogram entries (swifttn93.cpp:2406-2415); a split run adds them once, for all its parts
  if (J.A.triangle && J.ra < 0)
    for (int64_t i = J.A.c0; i < J.A.n; i++) {
      const int64_t wi = J.S->hw[(size_t)i];
      if (wi > 1) R->hist[(J.A.fst && i >= J.A.n1) ? 1 : 0][0] += (unsigned long long)((wi * (wi - 1)) >> 1);
    }
}

// The five link arrays of a result are one block, with its capacity and kind (0 host memory, 1 pinned) in a 64-byte
// header.  Freed blocks are kept (at most LINK_CACHE host blocks) and handed out again, so a caller running the same
// workload repeatedly does not map and fault in fresh pages for every result.  Single fast-path runs take pinned
// blocks, into which the read-back kernel writes the links directly (no copy on the host): at most PIN_BLOCKS of them
// exist, in use or kept, each with half again the capacity first asked for, and none is unpinned (pinning and
// unpinning cost far more than the copy they save; unpinning may wait for the device); a run that finds none free
// and large enough takes a host block and the copy.
constexpr int LINK_CACHE = 4, PIN_BLOCKS = 8;
static std::mutex g_link_mu;
static std::vector<void *> g_link_cache[2];
static int g_pin_blocks = 0;

static int alloc_links(tn93g_result *R, size_t nl, size_t cap_want = 0, int kind = 0)
{
  R->n_links = (int64_t)nl;
  size_t cap = std::max<size_t>({nl, cap_want, 1});
  char *blk = nullptr;
  bool fresh = false;
  {
    std::lock_guard<std::mutex> lk(g_link_mu);
    auto &cache = g_link_cache[kind];
    for (size_t k = 0; k < cache.size(); k++) {
      size_t have;
      memcpy(&have, cache[k], sizeof have);
      if (have >= cap && (kind || have <= 2 * cap + 4096)) {
        blk = (char *)cache[k];
        cache.erase(cache.begin() + (long)k);
        break;
      }
    }
    if (!blk && kind) {
      if (g_pin_blocks >= PIN_BLOCKS) return TN93G_E_NOMEM;  // every one in use or too small: a host block
      cap += cap / 2;  // room for the next runs' growth
      if (hipHostMalloc((void **)&blk, 64 + 5 * cap * 8) != hipSuccess) {
        (void)hipGetLastError();
        return TN93G_E_NOMEM;
      }
      g_pin_blocks++;
      fresh = true;
    }
  }
  if (!blk) {
    blk = (char *)malloc(64 + 5 * cap * 8);
    if (!blk) return TN93G_E_NOMEM;
    fresh = true;
  }
  if (fresh) {
    const size_t k64 = (size_t)kind;
    memcpy(blk, &cap, sizeof cap);
    memcpy(blk + 8, &k64, sizeof k64);
  }
  size_t have;
  memcpy(&have, blk, sizeof have);
  int64_t *base = (int64_t *)(blk + 64);
  R->link_row = base;
  R->link_col = base + have;
  R->link_dist = (double *)(base + 2 * have);
  R->link_dist_lo = (double *)(base + 3 * have);
  R->link_dist_hi = (double *)(base + 4 * have);
  return 0;
}

static void free_links(int64_t *link_row)
{
  if (!link_row) return;
  char *blk = (char *)link_row - 64;
  size_t kind;
  memcpy(&kind, blk + 8, sizeof kind);
  {
    std::lock_guard<std::mutex> lk(g_link_mu);
    if (kind || g_link_cache[0].size() < LINK_CACHE) {  // pinned blocks are all kept (at most PIN_BLOCKS)
      g_link_cache[kind].push_back(blk);
      return;
    }
  }
  free(blk);
}

// Enclosure of tn93's mean (swifttn93.cpp:2594-2602, 2839-2848), the maximum and the integer sums, from the fast
// path's row accumulators added in row order.  Row r: tn93 adds its n_r terms t in column order in FP64; here they
// were added in another order.  Both results are within gamma_{n_r} * sum|t| of the exact sum (Higham, Accuracy and
// Stability of Numerical Algorithms, 2nd ed., eq. 4.4), so tn93's row sum lies in [slo - 2 g sabs', shi + 2 g sabs']
// with sabs' an upper bound of sum|t|.  The rows are then added in row order: again within gamma_R * sum |row sums|.
// Host arithmetic is double.  For n u <= 2^-22 (n < 2^31), gamma_n <= n u (1 + 2^-21) <= fl(n cu) with
// cu = u (1 + 2^-20); K = 1 + 2^-20 covers the factor (1 + g) and the roundings of each product.  Adding the
// bounds lo_r, hi_r up in double is again within gamma_{R+1} sum max(|lo_r|, |hi_r|) (their own rounding
// included), so eR below covers both tn93's row-order sum and this one; mabs itself is inflated by K.
struct RowMerge {
  double mlo = 0, mhi = 0, mabs = 0, maxlo = 0., maxhi = 0.;
  long long wsum = 0, skip = 0, linkw = 0;
  static constexpr double cu = 0x1.00001p-53, K = 1.0 + 0x1p-20;
  void add(const RowAcc &q, double n_r)  // one row, summed over n_r pairs; rows must come in order
  {
    const double g = std::max(n_r, 1.0) * cu;
    const double e = (2.0 * g) * q.sabs * K;
    const double lo = q.slo - e, hi = q.shi + e;
    mlo += lo;
    mhi += hi;
    mabs += std::max(std::fabs(lo), std::fabs(hi));
    const double a = tn93g_host_from_bits(q.mlo), b = tn93g_host_from_bits(q.mhi);
    if (a > maxlo) maxlo = a;
    if (b > maxhi) maxhi = b;
    wsum += q.wsum;
    skip += q.skip;
    linkw += q.linkw;
  }
  // nrows: the rows added; self_w: the weights of the self pairs (copy numbers above 1)
  void finish(tn93g_result *R, int64_t nrows, long long self_w) const
  {
    const double gR = ((double)std::max<int64_t>(nrows, 1) + 1.0) * cu;
    const double eR = (2.0 * gR) * (mabs * K) * K + 0x1p-50 * mabs;
    R->mean_lo[0] = std::nextafter(mlo - eR, -INFINITY);
    R->mean_hi[0] = std::nextafter(mhi + eR, INFINITY);
    R->max_lo[0] = maxlo;
    R->max_hi[0] = maxhi;
    R->skipped = skip;
    R->found_links = linkw;
    R->weighted[0] = (double)(wsum + self_w);
  }
};

// Pairs of row r of a run.
static double row_pairs(const RunArgs &A, int64_t r)
{
  const int64_t i = A.r0 + r;
  return (double)(A.triangle ? std::max<int64_t>(A.c1 - std::max(A.c0, i + 1), 0) : A.ncols);
}

// The weights of the self pairs of sequences [c0, n) (a copy number w adds w (w - 1) / 2 pairs at distance 0).
static long long self_weights(const tn93g_set *S, int64_t c0, int64_t n)
{
  long long w = 0;
  for (int64_t i = c0; i < n; i++) {
    const int64_t wi = S->hw[(size_t)i];
    if (wi > 1) w += (wi * (wi - 1)) >> 1;
  }
  return w;
}

// The rows of the parts of a split run, gathered in row order: the enclosure merge, and tn93's own row-order sums
// while every part is exact.
struct RowSink {
  RowMerge M;
  int64_t rows = 0;
  bool exact = true;
  double mean[3] = {0., 0., 0.}, weighted[3] = {0., 0., 0.}, maxv[3] = {0., 0., 0.};
};

// The fast path's result from its read-back: links in tn93's order, the enclosures of the mean and the maximum.
static int job_fast_result(const Job &J, const char *pin, const LinkRec *lk, tn93g_result *R,
                           RowSink *sink = nullptr)
{
  const FastLayout F = fast_layout(J, 0);
  const RowAcc *ra = (const RowAcc *)(pin + F.off_ra);
  const size_t nl = J.n_linkrec;
  // the records arrive in tn93's emission order (rows ascending, columns ascending: k_link_order); lk null: the
  // read-back kernel wrote them into R's arrays already
  if (!lk)
    R->n_links = (int64_t)nl;
  else if (alloc_links(R, nl))
    return TN93G_E_NOMEM;
  for (size_t k = 0; lk && k < nl; k++) {
    const LinkRec &q = lk[k];
    R->link_row[k] = q.i;
    R->link_col[k] = q.j;
    R->link_dist[k] = q.d;
    R->link_dist_lo[k] = q.lo;
    R->link_dist_hi[k] = q.hi;
  }
  if (sink) {  // a part of a split run: the rows go on to the whole run's merge
    for (int64_t r = 0; r < J.nrows; r++) sink->M.add(ra[(size_t)r], row_pairs(J.A, r));
    sink->rows += J.nrows;
    sink->exact = false;
  } else {
    RowMerge M;
    for (int64_t r = 0; r < J.nrows; r++) M.add(ra[(size_t)r], row_pairs(J.A, r));
    M.finish(R, J.nrows, J.A.triangle ? self_weights(J.S, J.A.c0, J.A.n) : 0);
  }
  unsigned long long hrep[600];
  memcpy(hrep, pin + F.off_h, sizeof hrep);
  job_finish_common(J, hrep, R);
  return 0;
}

// The exact path after the run's kernels: host-exact tails patched in, rows summed in column order by k_rows.
static int job_exact_result(tn93g_ctx *c, Job &J, tn93g_result *R, RowSink *sink = nullptr)
{
  RunArgs &A = J.A;
  RunBufs *B = J.B;
  const int64_t nrows = J.nrows;
  const size_t nr1 = J.nr1;
  const unsigned n_exact = J.n_exact;
  if (A.exact && n_exact) {
    if (n_exact > J.exact_cap) return c->fail(TN93G_E_NOMEM, "host-exact list overflow");  // tn93g_run grew it
    std::vector<ExactRec> ex(n_exact);
    HIPCHK(c, hipMemcpyAsync(ex.data(), B->exact_list.p, n_exact * sizeof(ExactRec), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<PatchRec> pr(n_exact);
    auto work = [&](size_t b0, size_t e0) {
      for (size_t k = b0; k < e0; k++) {
        double dd = host_tail(ex[k].tin);
        double x = dd <= 0. ? 0. : dd;
        pr[k].i = ex[k].i;
        pr[k].j = ex[k].j;
        pr[k].thisD = x;
        pr[k].idx = host_hist_index(dd);
        pr[k].link = x >= J.P.min_distance && x <= J.P.threshold;
      }
    };
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n_exact < 4096) nth = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++)
      th.emplace_back(work, (size_t)n_exact * t / nth, (size_t)n_exact * (t + 1) / nth);
    for (auto &t : th) t.join();
    if (ensure(c, B->patch, pr.size() * sizeof(PatchRec))) return TN93G_E_NOMEM;
    HIPCHK(c, hipMemcpyAsync(B->patch.p, pr.data(), pr.size() * sizeof(PatchRec), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_patch, dim3((unsigned)((n_exact + 255) / 256)), dim3(256), 0, c->stream, A,
                       (const PatchRec *)B->patch.p, (int64_t)n_exact);
  }
  // link offsets, then the rows in column order
  int64_t *loff = (int64_t *)B->link_off.p;
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, (const int32_t *)B->row_links.p, loff, nrows);
  int64_t n_links = 0;
  HIPCHK(c, hipMemcpyAsync(&n_links, loff + nrows, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ensure(c, B->links, (size_t)std::max<int64_t>(n_links, 1) * 40)) return TN93G_E_NOMEM;
  RowOut O;
  {
    double *base = (double *)B->rowout.p;
    O.sum_lo = base;
    O.sum_hi = base + 3 * nr1;
    O.wsum = base + 6 * nr1;
    O.max_lo = base + 9 * nr1;
    O.max_hi = base + 12 * nr1;
    O.skipped = (int64_t *)(base + 15 * nr1);
    O.linkw = (int64_t *)(base + 16 * nr1);
    const size_t nl = (size_t)std::max<int64_t>(n_links, 1);
    O.link_row = (int64_t *)B->links.p;
    O.link_col = O.link_row + nl;
    O.link_d = (double *)(O.link_col + nl);
    O.link_lo = O.link_d + nl;
    O.link_hi = O.link_lo + nl;
    O.link_off = loff;
  }
  if (nrows > 0) hipLaunchKernelGGL(k_rows, dim3((unsigned)((nrows + RT - 1) / RT)), dim3(RT), 0, c->stream, A, O);
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  HIPCHK(c, hipGetLastError());
  const size_t nl = (size_t)n_links;
  if (alloc_links(R, nl)) return TN93G_E_NOMEM;
  std::vector<double> rows(nr1 * 17);
  std::vector<unsigned long long> hrep(600);
  if (nl) {
    HIPCHK(c, hipMemcpyAsync(R->link_row, O.link_row, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_col, O.link_col, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_dist, O.link_d, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_dist_lo, O.link_lo, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_dist_hi, O.link_hi, nl * 8, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(rows.data(), B->rowout.p, rows.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hrep.data(), B->hist.p, hrep.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // row-order merge (swifttn93.cpp:2832-2848)
  for (int b = 0; b < 3; b++) R->mean_lo[b] = R->mean_hi[b] = R->weighted[b] = R->max_lo[b] = R->max_hi[b] = 0.;
  if (sink) {  // a part of a split run (never F_ST: one histogram): rows into the whole run's merges
    for (int64_t r = 0; r < nrows; r++) {
      const size_t q = (size_t)r;
      RowAcc a{};
      a.slo = rows[q];
      a.shi = rows[3 * nr1 + q];
      a.sabs = 0.;  // tn93's own row sum (column order) when exact, else an enclosure of it
      double mlo = rows[9 * nr1 + q], mhi = rows[12 * nr1 + q];
      memcpy(&a.mlo, &mlo, 8);
      memcpy(&a.mhi, &mhi, 8);
      // k_rows counts the row's own self pairs into its weights; the merge adds them once for the whole run
      const int64_t i = A.r0 + r, wi = J.S->hw[(size_t)i];
      const double self = A.triangle && i >= A.c0 && wi > 1 ? (double)((wi * (wi - 1)) >> 1) : 0.;
      a.wsum = (long long)(rows[6 * nr1 + q] - self);
      memcpy(&a.skip, &rows[15 * nr1 + q], 8);
      memcpy(&a.linkw, &rows[16 * nr1 + q], 8);
      sink->M.add(a, row_pairs(A, r));
      sink->mean[0] += a.slo;
      sink->weighted[0] += rows[6 * nr1 + q] - self;
      if (mlo > sink->maxv[0]) sink->maxv[0] = mlo;
    }
    sink->rows += nrows;
  }
  for (int64_t r = 0; r < nrows && !sink; r++) {
    for (int b = 0; b < 3; b++) {
      const double mlo = rows[9 * nr1 + b * nr1 + (size_t)r], mhi = rows[12 * nr1 + b * nr1 + (size_t)r];
      if (mlo > R->max_lo[b]) R->max_lo[b] = mlo;
      if (mhi > R->max_hi[b]) R->max_hi[b] = mhi;
      R->mean_lo[b] += rows[b * nr1 + (size_t)r];
      R->mean_hi[b] += rows[3 * nr1 + b * nr1 + (size_t)r];
      R->weighted[b] += rows[6 * nr1 + b * nr1 + (size_t)r];
    }
    int64_t sk, lw;
    memcpy(&sk, &rows[15 * nr1 + (size_t)r], 8);
    memcpy(&lw, &rows[16 * nr1 + (size_t)r], 8);
    R->skipped += sk;
    R->found_links += lw;
  }
  job_finish_common(J, hrep.data(), R);
  return 0;
}

static void set_times(tn93g_ctx *c, tn93g_result *R, bool single)
{
  float ms_all = 0, ms_pair = 0, ms_tail = 0;
  (void)hipEventElapsedTime(&ms_all, c->ev[4], c->ev[2]);
  if (single) {
    (void)hipEventElapsedTime(&ms_pair, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&ms_tail, c->ev[1], c->ev[3]);
  }
  R->gpu_ms = ms_all;
  R->pair_kernel_ms = ms_pair;
  R->tail_kernel_ms = ms_tail;
}

extern "C" {

const char *tn93g_version(void) { return TN93G_VERSION; }

// The streams and events of a batch's NBRANCH branches (created on a context's first batch).
static int ensure_branches(tn93g_ctx *c)
{
  if (c->fork) return 0;
  for (int b = 0; b < NBRANCH; b++) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->side[b], hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->join[b], hipEventDisableTiming));
    if (b == 0) continue;  // branch 0's replay side stream exists from the start
    HIPCHK(c, hipStreamCreateWithFlags(&c->rside[b], hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->rfork[b], hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->rjoin[b], hipEventDisableTiming));
  }
  HIPCHK(c, hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  return 0;
}

int tn93g_open(int device, tn93g_ctx **out)
{
  if (!out) return TN93G_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return TN93G_E_NODEV;
  tn93g_ctx *c = new tn93g_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return TN93G_E_NODEV;
  }
  for (auto &e : c->ev) (void)hipEventCreate(&e);
  if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) c->wall_khz = 0;
  // a single run needs branch 0's side stream (its tail beside its replay); the batch branches come on first use
  if (hipStreamCreateWithFlags(&c->rside[0], hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->rfork[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->rjoin[0], hipEventDisableTiming) != hipSuccess) {
    tn93g_close(c);
    return TN93G_E_NODEV;
  }
  // the no-allocation entry point's scratch on this device (a failed lookup only disables that entry point)
  c->scratch = scratch_address(device, &c->scratch_err);
  *out = c;
  return 0;
}

void tn93g_close(tn93g_ctx *c)
{
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  for (RunBufs *b : c->slots) {
    b->release();
    delete b;
  }
  for (auto &e : c->ev) (void)hipEventDestroy(e);
  for (int b = 0; b < NBRANCH; b++) {
    if (c->side[b]) (void)hipStreamDestroy(c->side[b]);
    if (c->join[b]) (void)hipEventDestroy(c->join[b]);
    if (c->rside[b]) (void)hipStreamDestroy(c->rside[b]);
    if (c->rfork[b]) (void)hipEventDestroy(c->rfork[b]);
    if (c->rjoin[b]) (void)hipEventDestroy(c->rjoin[b]);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->blas) (void)rocblas_destroy_handle(c->blas);
  if (c->gemm_ops) (void)hipFree(c->gemm_ops);
  if (c->gemm_out) (void)hipFree(c->gemm_out);
  if (c->pin.p) (void)hipHostFree(c->pin.p);
  if (c->pin_links.p) (void)hipHostFree(c->pin_links.p);
  if (c->pin_up.p) (void)hipHostFree(c->pin_up.p);
  if (c->multi.p) (void)hipFree(c->multi.p);
  if (c->stage.p) (void)hipFree(c->stage.p);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *tn93g_error(const tn93g_ctx *c) { return c ? c->err.c_str() : "no context"; }

int tn93g_set_create(tn93g_ctx *c, int64_t L, int64_t capacity, tn93g_set **out)
{
  if (!c || !out || capacity < 0 || L <= 0 || capacity >= (1ll << 31) || L >= (1ll << 26)) return TN93G_E_ARG;
  *out = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  tn93g_set *S = new tn93g_set();
  S->ctx = c;
  S->cap = capacity;
  SetDev &d = S->d;
  memset(&d, 0, sizeof d);
  d.n = 0;
  d.L = L;
  d.Ls = (L + 15) / 16 * 16;
  d.npad = (capacity + 63) / 64 * 64 + 64;
  d.W = (L + 31) / 32;
  d.W64 = (L + 63) / 64;
  const int64_t nn = std::max<int64_t>(capacity, 1);
  S->hw.assign((size_t)nn, 1);
  S->spec.assign((size_t)nn, 0);
  auto fail = [&](int code, const std::string &m) {
    tn93g_set_free(S);
    return c->fail(code, m);
  };
#define SCHK(x)                                                                              \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail(TN93G_E_HIP, fmt("%s: %s", #x, hipGetErrorString(e_))); \
  } while (0)
  SCHK(hipMalloc(&d.codes, (size_t)(nn * d.Ls)));
  SCHK(hipMalloc(&d.codesT, (size_t)(d.Ls * d.npad)));
  SCHK(hipMalloc(&d.planes, (size_t)(nn * d.W + 4) * sizeof(uint4)));  // + 4: the pair kernel's 4-word row loads
  SCHK(hipMalloc(&d.onehot, (size_t)(d.npad * d.W64 * 256)));
  SCHK(hipMalloc(&d.ev_off, (size_t)(nn + 1) * sizeof(int32_t)));
  SCHK(hipMalloc(&d.evg_off, (size_t)(nn + 1) * sizeof(int32_t)));
  SCHK(hipMalloc(&d.nrun_off, (size_t)(nn + 1) * sizeof(int32_t)));
  SCHK(hipMalloc(&d.grun_off, (size_t)(nn + 1) * sizeof(int32_t)));
  SCHK(hipMalloc(&d.pre, (size_t)(nn * (d.W + 1)) * sizeof(PrefixRec)));
  SCHK(hipMalloc(&d.tb, (size_t)nn * sizeof(uint4)));
  SCHK(hipMalloc(&d.npre, (size_t)(nn * (d.W + 1)) * sizeof(uint32_t)));
  SCHK(hipMalloc(&d.nmask, (size_t)(nn * d.W + 4) * sizeof(uint32_t)));
  SCHK(hipMalloc(&d.weights, (size_t)nn * sizeof(int64_t)));
  // padding reads as gaps (codes, site-major codes) and as nothing (planes, one-hot operands, event offsets)
  SCHK(hipMemsetAsync(d.codes, GAPC, (size_t)(nn * d.Ls), c->stream));
  SCHK(hipMemsetAsync(d.codesT, GAPC, (size_t)(d.Ls * d.npad), c->stream));
  SCHK(hipMemsetAsync(d.planes, 0, (size_t)(nn * d.W + 4) * sizeof(uint4), c->stream));
  SCHK(hipMemsetAsync(d.onehot, 0, (size_t)(d.npad * d.W64 * 256), c->stream));
  SCHK(hipMemsetAsync(d.ev_off, 0, (size_t)(nn + 1) * sizeof(int32_t), c->stream));
  SCHK(hipMemsetAsync(d.evg_off, 0, (size_t)(nn + 1) * sizeof(int32_t), c->stream));
  SCHK(hipMemsetAsync(d.nrun_off, 0, (size_t)(nn + 1) * sizeof(int32_t), c->stream));
  SCHK(hipMemsetAsync(d.grun_off, 0, (size_t)(nn + 1) * sizeof(int32_t), c->stream));
  SCHK(hipMemsetAsync(d.nmask, 0, (size_t)(nn * d.W + 4) * sizeof(uint32_t), c->stream));
  SCHK(hipStreamSynchronize(c->stream));
#undef SCHK
  *out = S;
  return 0;
}

}  // extern "C"

// Appends n sequences of srcL columns each; `cols` (the set's L entries, or null when srcL == L) names the columns
// that are kept, in order.
static int set_append(tn93g_set *S, const uint8_t *codes, int64_t n, const int64_t *weights, const int32_t *cols,
                      int64_t srcL)
{
  if (!S || n < 0 || (n > 0 && !codes)) return TN93G_E_ARG;
  tn93g_ctx *c = S->ctx;
  std::lock_guard<std::mutex> lk(S->mu);
  SetDev &d = S->d;
  const int64_t L = d.L, n0 = d.n;
  if (!cols && d.kt) return c->fail(TN93G_E_ARG, "the set's constant columns were folded: it cannot grow");
  if (n0 + n > S->cap) return c->fail(TN93G_E_ARG, "set capacity exceeded");
  if (n == 0) return 0;
  // the codes checked, the B/D/H/V flags set and the gap-padded copy made in one pass, blocks of sequences on up to
  // 16 host threads (one per 1 MB copied, one per 256 KB gathered from kept columns)
  // staged in the context's pinned buffer (one copy at full PCIe rate; an append ends synchronised, so the buffer is
  // free again when the next append of a set of this context starts)
  if (ensure_pinned(c, c->pin_up, (size_t)(n * d.Ls))) return TN93G_E_NOMEM;
  uint8_t *padded = (uint8_t *)c->pin_up.p;
  std::atomic<bool> bad{false};
  {
    const int64_t bytes = n * L;
    const int nt = (int)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)std::max(1u, std::thread::hardware_concurrency()), 16,
                              cols ? bytes >> 18 : bytes >> 20, n}));
    auto work = [&](int t) {
      for (int64_t q = n * t / nt; q < n * (t + 1) / nt && !bad.load(std::memory_order_relaxed); q++) {
        const uint8_t *src = codes + q * srcL;
        uint8_t *dst = &padded[(size_t)(q * d.Ls)];
        unsigned mx = 0, sp = 0;
        if (cols)
          for (int64_t k = 0; k < L; k++) dst[k] = src[cols[k]];
        else
          memcpy(dst, src, (size_t)L);
        for (int64_t k = 0; k < L; k++) {
          const unsigned v = dst[k];
          mx |= v > (unsigned)GAPC ? 1u : 0u;
          sp |= (v - 11u) < 4u ? 1u : 0u;  // B, D, H, V: the only codes whose effects are not dyadic
        }
        memset(dst + L, GAPC, (size_t)(d.Ls - L));
        if (mx) bad = true;
        S->spec[(size_t)(n0 + q)] = (uint8_t)sp;
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
  }
  if (bad) return c->fail(TN93G_E_ARG, "codes must be ValidChars indices 0..17");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<int64_t> wts((size_t)n, 1);
  if (weights)
    for (int64_t s = 0; s < n; s++) wts[(size_t)s] = weights[s];
  HIPCHK(c, hipMemcpyAsync(d.codes + n0 * d.Ls, padded, (size_t)(n * d.Ls), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d.weights + n0, wts.data(), wts.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           c->stream));
  const int64_t n1 = n0 + n;
  {
    const int64_t sb = n0 & ~(int64_t)63;
    dim3 g((unsigned)((d.Ls + 63) / 64), (unsigned)((n1 - sb + 63) / 64));
    hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, c->stream, d.codes, d.codesT, n1, d.Ls, d.npad, sb);
    const int64_t tot = n * d.W;
    hipLaunchKernelGGL(k_planes, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, d.codes, d.planes,
                       d.nmask, n0, n1, L, d.Ls, d.W);
    hipLaunchKernelGGL(k_prefix, dim3((unsigned)n), dim3(64), 0, c->stream, (const uint4 *)d.planes,
                       (const uint32_t *)d.nmask, d.pre, d.tb, d.npre, d.W,
                       n0);
    const int64_t ob = n0 & ~(int64_t)3, oe = (n1 + 3) & ~(int64_t)3;
    const int64_t toh = (oe - ob) * d.W64 * 64;
    hipLaunchKernelGGL(k_onehot, dim3((unsigned)((toh + 255) / 256)), dim3(256), 0, c->stream, d.codes, d.onehot, n1,
                       L, d.Ls, d.W64, ob, oe);
  }
  // every list's lengths in one kernel and one read-back; then per list the offsets, the array (grown when needed)
  // and the fill kernel, with one synchronisation at the end
  int rc = 0;
  if (S->listcnt_bytes < (size_t)(5 * n) * sizeof(int32_t)) {
    if (S->listcnt) (void)hipFree(S->listcnt);
    S->listcnt = nullptr;
    S->listcnt_bytes = 0;
    if (hipMalloc(&S->listcnt, (size_t)(5 * n) * sizeof(int32_t)) != hipSuccess)
      return c->fail(TN93G_E_NOMEM, "list count allocation failed");
    S->listcnt_bytes = (size_t)(5 * n) * sizeof(int32_t);
  }
  int32_t *cnt_d = (int32_t *)S->listcnt;
  hipLaunchKernelGGL(k_listcount, dim3((unsigned)n), dim3(64), 0, c->stream, d.codes, cnt_d, L, d.Ls, n0, n);
  std::vector<int32_t> cnt((size_t)(5 * n));
  HIPCHK(c, hipMemcpyAsync(cnt.data(), cnt_d, cnt.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!S->decided) {
    // dense N: on average 32 or more N sites per sequence of the first batch; then N stays out of the ambiguity
    // event lists for the life of the set
    int64_t nsites = 0;
    for (int64_t k = 0; k < n; k++) nsites += cnt[(size_t)k] - cnt[(size_t)(n + k)];
    d.ndense = nsites >= 32 * n ? 1 : 0;
    S->decided = true;
  }
  // lists: 0 = ambiguity events (without N when dense), 1 = with gap sites (GAPMM), 2 = N runs (dense N only),
  // 3 = gap runs
  std::vector<int32_t> offs((size_t)(4 * n), 0);  // kept until the synchronisation below (pageable source)
  int64_t totals[4] = {S->ev_total[0], S->ev_total[1], S->ev_total[2], S->ev_total[3]};
  for (int li = 0; li < 4 && !rc; li++) {
    if (li == 2 && !d.ndense) continue;
    int32_t *eoff = li == 0 ? d.ev_off : li == 1 ? d.evg_off : li == 2 ? d.nrun_off : d.grun_off;
    const size_t esz = li >= 2 ? sizeof(uint2) : sizeof(uint32_t);
    void *arr = li == 0 ? (void *)d.ev : li == 1 ? (void *)d.evg : li == 2 ? (void *)d.nruns : (void *)d.gruns;
    const int mode = li == 1 ? EV_GAPS : d.ndense ? EV_NO_N : EV_AMB;
    const int rcode = li == 2 ? NC : GAPC;
    const int32_t *lc = cnt.data() + (size_t)n * (li == 0 ? (d.ndense ? 1 : 0) : li == 1 ? 2 : li == 2 ? 3 : 4);
    int32_t *off = offs.data() + (size_t)li * n;
    int64_t total = totals[li];
    for (int64_t k = 0; k < n; k++) {
      total += lc[(size_t)k];
      off[(size_t)k] = (int32_t)std::min<int64_t>(total, INT32_MAX);  // offsets of sequences n0 + 1 + k
    }
    if (total >= (1ll << 31)) {
      rc = c->fail(TN93G_E_ARG, "too many ambiguity events");
      break;
    }
    if (total > S->ev_cap[li] || !arr) {
      // grow: a new array with the old entries copied; the old one stays valid for runs in flight
      const int64_t want = std::max<int64_t>({total, 2 * S->ev_cap[li], 1024});
      void *na = nullptr;
      if (hipMalloc(&na, (size_t)want * esz) != hipSuccess) {
        rc = c->fail(TN93G_E_NOMEM, "event array allocation failed");
        break;
      }
      if (S->ev_total[li])
        (void)hipMemcpyAsync(na, arr, (size_t)S->ev_total[li] * esz, hipMemcpyDeviceToDevice, c->stream);
      if (arr) S->retired.push_back(arr);
      arr = na;
      if (li == 0) d.ev = (uint32_t *)na;
      else if (li == 1) d.evg = (uint32_t *)na;
      else if (li == 2) d.nruns = (uint2 *)na;
      else d.gruns = (uint2 *)na;
      S->ev_cap[li] = want;
    }
    if (hipMemcpyAsync(eoff + n0 + 1, off, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream) !=
        hipSuccess) {
      rc = c->fail(TN93G_E_HIP, "event offsets upload failed");
      break;
    }
    if (li < 2)
      hipLaunchKernelGGL(k_evfill, dim3((unsigned)n), dim3(64), 0, c->stream, d.codes, eoff, (uint32_t *)arr, L, d.Ls,
                         mode, n0);
    else
      hipLaunchKernelGGL(k_runfill, dim3((unsigned)n), dim3(64), 0, c->stream, d.codes, eoff, (uint2 *)arr, L, d.Ls,
                         n0, rcode);
    totals[li] = total;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = c->fail(TN93G_E_HIP, "event packing failed");
  if (!rc)
    for (int li = 0; li < 4; li++) S->ev_total[li] = totals[li];
  if (rc) return rc;
  HIPCHK(c, hipGetLastError());
  for (int64_t s = 0; s < n; s++) {
    S->hw[(size_t)(n0 + s)] = wts[(size_t)s];
    S->nonunit += wts[(size_t)s] != 1 ? 1 : 0;
  }
  d.total_events = S->ev_total[0];
  d.total_gruns = S->ev_total[3];
  d.n = n1;  // published: runs that start from now on see the new sequences
  return 0;
}

// The constant columns of n sequences of L columns: cst[p] = 1 where every sequence holds the same clean base (the
// columns swifttn93.cpp:591-660 leaves out of its counting).  Host threads take blocks of columns and leave a block
// as soon as none of its columns can still be constant.
static void constant_columns(const uint8_t *codes, int64_t n, int64_t L, std::vector<uint8_t> &cst)
{
  cst.assign((size_t)L, 0);
  if (n <= 0) return;
  for (int64_t p = 0; p < L; p++) cst[(size_t)p] = codes[p] < 4 ? 1 : 0;
  const int64_t B = 4096, nb = (L + B - 1) / B;
  const int nt = (int)std::max<int64_t>(
      1, std::min<int64_t>({(int64_t)std::max(1u, std::thread::hardware_concurrency()), 16, nb}));
  std::atomic<int64_t> next{0};
  auto work = [&] {
    uint8_t *cs = cst.data();
    for (int64_t b; (b = next.fetch_add(1)) < nb;) {
      const int64_t p0 = b * B, p1 = std::min(L, p0 + B);
      for (int64_t i = 1; i < n; i++) {
        const uint8_t *r = codes + i * L;
        unsigned any = 0;
        for (int64_t p = p0; p < p1; p++) {
          cs[p] &= (uint8_t)(r[p] == codes[p]);
          any |= cs[p];
        }
        if (!any) break;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

extern "C" {

int tn93g_set_append(tn93g_set *S, const uint8_t *codes, int64_t n, const int64_t *weights)
{
  return set_append(S, codes, n, weights, nullptr, S ? S->d.L : 0);
}

int64_t tn93g_set_size(const tn93g_set *S)
{
  if (!S) return 0;
  std::lock_guard<std::mutex> lk(const_cast<tn93g_set *>(S)->mu);
  return S->d.n;
}

int tn93g_set_upload(tn93g_ctx *c, const uint8_t *codes, int64_t n, int64_t L, const int64_t *weights,
                     tn93g_set **out)
{
  if (!c || !out || (n > 0 && !codes) || n < 0 || L <= 0 || n >= (1ll << 31) || L >= (1ll << 26))
    return TN93G_E_ARG;
  *out = nullptr;
  // Constant columns add one clean/clean site of their base to every pair: when they are at least 1/16 of the
  // columns they are left out of the device arrays and their counts added to each pair's totals.  TN93G_NOFOLD=1
  // keeps every column (a diagnostic: the output is the same).
  std::vector<int32_t> keep;
  int64_t kt = 0;
  uint32_t kc[4] = {0, 0, 0, 0};
  const char *nf = getenv("TN93G_NOFOLD");
  if (n > 1 && !(nf && *nf && strcmp(nf, "0"))) {
    std::vector<uint8_t> cst;
    constant_columns(codes, n, L, cst);
    for (int64_t p = 0; p < L; p++)
      if (cst[(size_t)p]) kt++;
    if (kt < L && 16 * kt >= L) {
      keep.reserve((size_t)(L - kt));
      for (int64_t p = 0; p < L; p++)
        if (cst[(size_t)p])
          kc[codes[p]]++;
        else
          keep.push_back((int32_t)p);
    } else {
      kt = 0;
    }
  }
  tn93g_set *S = nullptr;
  int rc = tn93g_set_create(c, L - kt, n, &S);
  if (rc) return rc;
  S->d.kt = kt;
  for (int b = 0; b < 4; b++) S->d.kc[b] = kc[b];
  if ((rc = set_append(S, codes, n, weights, kt ? keep.data() : nullptr, L))) {
    tn93g_set_free(S);
    return rc;
  }
  *out = S;
  return 0;
}

void tn93g_set_free(tn93g_set *S)
{
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  (void)hipDeviceSynchronize();  // runs of any context may still read the set
  void *ps[] = {S->d.codes,   S->d.codesT, S->d.planes,   S->d.onehot, S->d.ev_off,   S->d.ev,    S->d.evg_off,
                S->d.evg,     S->d.nrun_off, S->d.nruns, S->d.nmask,  S->d.grun_off, S->d.gruns, S->d.pre, S->d.tb, S->d.npre,
                S->d.weights};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  for (void *p : S->retired) (void)hipFree(p);
  if (S->listcnt) (void)hipFree(S->listcnt);
  delete S;
}

void tn93g_result_free(tn93g_result *r)
{
  if (!r) return;
  free_links(r->link_row);
  free(r);
}

}  // extern "C"

// True when lo and hi print the same with "%.{prec}g".  Most enclosures are a few ulps wide and far from a
// rounding midpoint of the printed digits: then the scaled values lo 10^(prec-1-e) and hi 10^(prec-1-e) (e the
// decimal exponent of lo) lie in [10^(prec-1), 10^prec) and round to the same integer with a margin of 1e-6, far
// above the error of the scaling (a few ulps of 10^prec); anything else is decided by formatting both.
static double pow10_tab(int k)  // 10^k for |k| <= 40 (the nearest doubles; their error is covered by the margin)
{
  static const std::array<double, 81> t = [] {
    std::array<double, 81> a{};
    for (int i = 0; i <= 80; i++) a[(size_t)i] = std::pow(10.0, (double)(i - 40));
    return a;
  }();
  return t[(size_t)(k + 40)];
}

static bool same_printed(double lo, double hi, int prec)
{
  if (lo == hi) return true;
  if (lo > 0. && hi > lo && hi < 1e20 && lo > 1e-20) {
    const int e = (int)std::floor(std::log10(lo));
    const double p = pow10_tab(prec - 1 - e), lo_d = pow10_tab(prec - 1), hi_d = 10.0 * lo_d;
    const double a = lo * p, b = hi * p;
    if (a >= lo_d * (1. + 1e-9) && b <= hi_d * (1. - 1e-9)) {
      const double fa = std::floor(a + 0.5), fb = std::floor(b + 0.5);
      if (fa == fb && std::fabs(a - fa) < 0.5 - 1e-6 && std::fabs(b - fb) < 0.5 - 1e-6) return true;
    }
  }
  char x[64], y[64];
  snprintf(x, sizeof x, "%.*g", prec, lo);
  snprintf(y, sizeof y, "%.*g", prec, hi);
  return strcmp(x, y) == 0;
}

// Runs of 2^31 pairs or more, or of more than 2^25 (row, column tile) slots, are split into parts of consecutive
// rows, each of at most 2^30 pairs and 2^24 slots (the device memory of a part stays below about 25 GB).  The parts
// run one after the other; their links follow each other in tn93's row order, their histograms add, and their rows
// enter one row-order merge.  F_ST runs are not split.
constexpr int64_t SPLIT_PAIRS = 1ll << 31, PART_PAIRS = 1ll << 30, SPLIT_SLOTS = 1ll << 25, PART_SLOTS = 1ll << 24;

// The run's rows [r0, r1) and columns [c0, c1) (triangle: only j > i), from the set's size and the parameters.
static void run_space(const tn93g_set *S, const tn93g_params *P, int64_t *r0, int64_t *r1, int64_t *c0, int64_t *c1,
                      bool *tri)
{
  const int64_t n = S->d.n;
  *r0 = 0;
  *r1 = P->kind == TN93G_CROSS ? P->n1 : n;
  *c0 = P->kind == TN93G_CROSS || P->kind == TN93G_APPEND ? P->n1 : 0;
  *c1 = n;
  *tri = P->kind != TN93G_CROSS;
}

// The parts of a split run, or none when the run is done in one piece.
static std::vector<std::pair<int64_t, int64_t>> split_parts(const tn93g_set *S, const tn93g_params *P)
{
  std::vector<std::pair<int64_t, int64_t>> parts;
  if (P->kind == TN93G_FST || (P->kind != TN93G_TRIANGLE && (P->n1 < 0 || P->n1 > S->d.n))) return parts;
  int64_t r0, r1, c0, c1;
  bool tri;
  run_space(S, P, &r0, &r1, &c0, &c1, &tri);
  if (r1 <= r0 || c1 <= c0) return parts;
  const int64_t gx = (c1 - (c0 & ~(int64_t)63) + TJ - 1) / TJ;
  const int64_t pairs = tri ? tri_pairs(r0, r1, c0, c1) : (r1 - r0) * (c1 - c0);
  const char *e = getenv("TN93G_PART_PAIRS");  // smaller parts (tests of the split at small sizes)
  const int64_t part_pairs = e && *e ? std::max<int64_t>(atoll(e), 1) : PART_PAIRS;
  if (pairs < SPLIT_PAIRS && (r1 - r0) * gx <= SPLIT_SLOTS && !(e && *e && pairs > part_pairs)) return parts;
  const int64_t max_rows = std::max<int64_t>(PART_SLOTS / std::max<int64_t>(gx, 1) / 32 * 32, 32);
  int64_t a = r0;
  while (a < r1) {
    int64_t b = a, p = 0;
    while (b < r1 && b - a < max_rows) {  // rows in blocks of 32 while the part's pairs fit
      const int64_t nb = std::min(b + 32, r1);
      const int64_t q = tri ? tri_pairs(b, nb, c0, c1) : (nb - b) * (c1 - c0);
      if (b > a && p + q > part_pairs) break;
      p += q;
      b = nb;
    }
    parts.push_back({a, b});
    a = b;
  }
  return parts;
}

extern "C" {

int tn93g_result_certify(const tn93g_result *r, int precision)
{
  if (!r || precision < 1 || precision > 17) return TN93G_E_ARG;
  if (r->exact) return 1;
  const int64_t nl = r->n_links;
  const unsigned nth = nl < 65536 ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<bool> ok{true};
  auto work = [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e && ok.load(std::memory_order_relaxed); k++)
      if (!same_printed(r->link_dist_lo[k], r->link_dist_hi[k], precision)) ok = false;
  };
  if (nth == 1)
    work(0, nl);
  else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++) th.emplace_back(work, nl * t / nth, nl * (t + 1) / nth);
    for (auto &t : th) t.join();
  }
  return ok ? 1 : 0;
}


// Autotuner.  A job's shape is its words per sequence, its size class (log2 of rows and of columns), its mode
// class and its pair kind.  The pair kernel's own shape has two choices: the backend (popcount or MFMA, when the
// caller leaves it to TN93G_BACKEND_AUTO) and the tile rows (8, 16 or 32 rows by 64 columns, when tile_rows is 0).
// The first job of a job shape with TN93G_FLAG_AUTOTUNE runs its kernels once with every candidate (the counts are
// exact either way, so the results are the same) and the fastest pair kernel is kept for the context's lifetime.
// Without the flag a job takes the tuned choice when there is one, else popcount in tiles of 16 rows.
static uint64_t tune_shape(const Job &J)
{
  auto lg = [](int64_t v) { return (uint64_t)(v > 0 ? 64 - __builtin_clzll((unsigned long long)v) : 0); };
  const uint64_t mode = J.A.gapmm ? 3 : J.A.all_slow ? 2 : J.A.skip_events ? 1 : 0;
  const uint64_t free_dims = (J.P.backend == TN93G_BACKEND_AUTO ? 1u : 0u) | (J.P.tile_rows == 0 ? 2u : 0u);
  return free_dims << 60 | (uint64_t)J.A.W << 24 | lg(J.nrows) << 16 | lg(J.A.ncols) << 8 | mode << 2 |
         (uint64_t)J.P.kind;
}

// A tuned choice is backend | rp << 8.
static int tuned_choice(const tn93g_ctx *c, const Job &J, int *backend, int *rp)
{
  *backend = J.P.backend == TN93G_BACKEND_AUTO ? TN93G_BACKEND_POPC : J.P.backend;
  *rp = J.rp;
  auto it = c->tuned.find(tune_shape(J));
  if (it == c->tuned.end()) return 0;
  *backend = it->second & 255;
  *rp = it->second >> 8;
  return 1;
}

static int autotune(tn93g_ctx *c, Job &J)
{
  const bool free_backend = J.P.backend == TN93G_BACKEND_AUTO, free_rows = J.P.tile_rows == 0;
  const uint64_t key = tune_shape(J);  // before the trials set J.P.backend
  int backend, rp;
  const bool known = tuned_choice(c, J, &backend, &rp) != 0;
  if (known || !(free_backend || free_rows) || !(J.P.flags & TN93G_FLAG_AUTOTUNE))
    return job_set_shape(c, J, backend, rp);
  // the candidates: popcount in every tile height allowed, MFMA in tiles of 16; the exact path tiles 16 rows
  std::vector<std::pair<int, int>> cand;
  for (int b : {TN93G_BACKEND_POPC, TN93G_BACKEND_MFMA}) {
    if (!free_backend && b != J.P.backend) continue;
    for (int r : {2, 4, 8}) {
      if (!free_rows && r != J.rp) continue;
      if ((b == TN93G_BACKEND_MFMA || !J.fast) && r != 4) continue;
      cand.push_back({b, r});
    }
  }
  if (cand.empty()) return job_set_shape(c, J, backend, rp);
  float best = 0;
  std::pair<int, int> pick = cand[0];
  for (size_t k = 0; k < cand.size(); k++) {
    int rc = job_set_shape(c, J, cand[k].first, cand[k].second);
    if (rc) return rc;
    job_enqueue(c, J, true, c->stream, 0, true, true);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    if (k == 0 || ms < best) {
      best = ms;
      pick = cand[k];
    }
  }
  c->tuned[key] = pick.first | pick.second << 8;
  return job_set_shape(c, J, pick.first, pick.second);
}

int tn93g_run_bytes(tn93g_ctx *c, const tn93g_set *S, const tn93g_params *P, int64_t *bytes)
{
  if (!c || !S || !P || !bytes) return TN93G_E_ARG;
  // a split run works in the memory of its largest part
  auto parts = split_parts(S, P);
  if (parts.empty()) parts.push_back({-1, -1});
  size_t most = 0;
  for (const auto &pr : parts) {
    Job J;
    J.S = S;
    J.P = *P;
    J.B = c->slot(0);
    J.ra = pr.first;
    J.rb = pr.second;
    size_t total = 0;
    const int rc = job_prepare(c, J, &total);
    if (rc) return rc;
    most = std::max(most, total);
  }
  *bytes = (int64_t)most;
  return 0;
}

}  // extern "C"

// One run, or the part [ra, rb) of the rows of a split run (ra >= 0: its rows go to `sink`, and its result holds
// its links, histogram and counters).
static int run_one(tn93g_ctx *c, const tn93g_set *S, const tn93g_params *P, int64_t ra, int64_t rb, RowSink *sink,
                   tn93g_result **out)
{
  Job J;
  J.S = S;
  J.P = *P;
  J.B = c->slot(0);
  J.ra = ra;
  J.rb = rb;
  int rc = job_prepare(c, J);
  if (rc) return rc;
  if ((rc = autotune(c, J))) return rc;
  if (use_gemm_counts(c, J, P->backend)) {
    if ((rc = job_set_shape(c, J, TN93G_BACKEND_POPC, 4))) return rc;
    J.gemm_counts = true;
  }
  job_enqueue(c, J, true, c->stream, 0, true, true);
  HIPCHK(c, hipGetLastError());
  // one synchronisation: the counters, and for the fast path the row accumulators, the histogram and as many link
  // records as the last run of this slot produced (plus a margin); more links take a second copy
  const size_t spec = J.fast ? std::min<size_t>(J.B->link_spec, (size_t)J.link_cap * NSTRIPE) : 0;
  const size_t fbytes = J.fast ? fast_layout(J, spec).bytes : 0;
  if (ensure_pinned(c, c->pin, COUNTER_STRIDE + fbytes)) return TN93G_E_NOMEM;
  char *pin = (char *)c->pin.p;
  // a single fast run's links go straight into its result's arrays: a pinned link block of spec entries
  tn93g_result *R = (tn93g_result *)calloc(1, sizeof(tn93g_result));
  if (!R) return TN93G_E_NOMEM;
  bool direct = false;
  if (J.fast && spec && !sink && alloc_links(R, 0, spec, 1) == 0) {
    void *dv = nullptr;
    direct = hipHostGetDevicePointer(&dv, R->link_row, 0) == hipSuccess && dv;
    if (!direct) {
      (void)hipGetLastError();
      free_links(R->link_row);
      R->link_row = R->link_col = nullptr;
      R->link_dist = R->link_dist_lo = R->link_dist_hi = nullptr;
    }
  }
  auto retry = [&](const tn93g_params *Q) {
    tn93g_result_free(R);
    return run_one(c, S, Q, ra, rb, sink, out);
  };
  if ((rc = job_enqueue_readback(c, J, pin, spec, direct ? R : nullptr))) {
    tn93g_result_free(R);
    return rc;
  }
  hipError_t e = hipEventRecord(c->ev[2], c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    tn93g_result_free(R);
    return c->fail(TN93G_E_HIP, fmt("run read-back: %s", hipGetErrorString(e)));
  }
  job_parse_counters(J, pin);
  if (J.n_undecided && !J.A.exact) {
    // an output decision could not be certified from the enclosures: run again resolving logs on the host
    tn93g_params Q = *P;
    Q.flags |= TN93G_FLAG_EXACT;
    return retry(&Q);
  }
  if (J.fast && J.max_region > J.link_cap) {  // more links than a region holds: grow the regions and run again
    J.B->link_cap_hint = (int64_t)J.max_region;
    return retry(P);
  }
  if (J.A.exact && J.n_exact > J.exact_cap) {  // more uncertain pairs than the host-exact list holds: grow it
    J.B->exact_hint = (int64_t)J.n_exact + J.n_exact / 4;
    return retry(P);
  }
  auto finish = [&]() -> int {
    if (J.fast) {
      const char *pf = pin + COUNTER_STRIDE;
      const LinkRec *lk = (const LinkRec *)(pf + fast_layout(J, 0).off_lk);
      J.B->link_spec = std::max<size_t>(4096, J.n_linkrec + J.n_linkrec / 4);
      if (direct && J.n_linkrec <= spec) return job_fast_result(J, pf, nullptr, R, sink);
      if (direct) {  // the pinned block is too small: the links are copied into a block of their size below
        free_links(R->link_row);
        R->link_row = R->link_col = nullptr;
        R->link_dist = R->link_dist_lo = R->link_dist_hi = nullptr;
      }
      if (J.n_linkrec > spec) {  // more links than copied: all of them again, into a buffer of their own
        if (ensure_pinned(c, c->pin_links, J.n_linkrec * sizeof(LinkRec))) return TN93G_E_NOMEM;
        HIPCHK(c, hipMemcpyAsync(c->pin_links.p, J.B->links.p, J.n_linkrec * sizeof(LinkRec),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        lk = (const LinkRec *)c->pin_links.p;
      }
      return job_fast_result(J, pf, lk, R, sink);
    }
    return job_exact_result(c, J, R, sink);
  };
  if ((rc = finish())) {
    tn93g_result_free(R);
    return rc;
  }
  set_times(c, R, true);
  *out = R;
  return 0;
}

// A split run: the parts in row order, merged into one result as one run would give it.
static int run_split(tn93g_ctx *c, const tn93g_set *S, const tn93g_params *P,
                     const std::vector<std::pair<int64_t, int64_t>> &parts, tn93g_result **out)
{
  RowSink sink;
  std::vector<tn93g_result *> res;
  auto drop = [&](int rc) {
    for (tn93g_result *r : res) tn93g_result_free(r);
    return rc;
  };
  bool exact = true;
  for (const auto &pr : parts) {
    tn93g_result *r = nullptr;
    const int rc = run_one(c, S, P, pr.first, pr.second, &sink, &r);
    if (rc) return drop(rc);
    res.push_back(r);
    exact = exact && r->exact;
  }
  tn93g_result *R = (tn93g_result *)calloc(1, sizeof(tn93g_result));
  if (!R) return drop(TN93G_E_NOMEM);
  size_t nl = 0;
  for (tn93g_result *r : res) nl += (size_t)r->n_links;
  if (alloc_links(R, nl)) {
    free(R);
    return drop(TN93G_E_NOMEM);
  }
  size_t k = 0;
  for (tn93g_result *r : res) {
    const size_t m = (size_t)r->n_links;
    if (m) {
      memcpy(R->link_row + k, r->link_row, m * 8);
      memcpy(R->link_col + k, r->link_col, m * 8);
      memcpy(R->link_dist + k, r->link_dist, m * 8);
      memcpy(R->link_dist_lo + k, r->link_dist_lo, m * 8);
      memcpy(R->link_dist_hi + k, r->link_dist_hi, m * 8);
    }
    k += m;
    for (int q = 0; q < 600; q++) (&R->hist[0][0])[q] += (&r->hist[0][0])[q];
    R->n_replayed += r->n_replayed;
    R->n_uncertain += r->n_uncertain;
    R->n_host_exact += r->n_host_exact;
    R->gpu_ms += r->gpu_ms;
    R->pair_kernel_ms += r->pair_kernel_ms;
    R->tail_kernel_ms += r->tail_kernel_ms;
  }
  R->backend_used = res[0]->backend_used;
  R->tile_rows_used = res[0]->tile_rows_used;
  R->folded_columns = res[0]->folded_columns;
  drop(0);
  int64_t r0, r1, c0, c1;
  bool tri;
  run_space(S, P, &r0, &r1, &c0, &c1, &tri);
  const long long self_w = tri ? self_weights(S, c0, S->d.n) : 0;
  if (tri)  // self_v_self histogram entries (swifttn93.cpp:2406-2415), once for the whole run
    for (int64_t i = c0; i < S->d.n; i++) {
      const int64_t wi = S->hw[(size_t)i];
      if (wi > 1) R->hist[0][0] += (unsigned long long)((wi * (wi - 1)) >> 1);
    }
  R->nrows = (int32_t)(r1 - r0);
  sink.M.finish(R, sink.rows, self_w);
  if (exact && sink.exact) {  // every part exact: tn93's row-order sums themselves
    R->mean_lo[0] = R->mean_hi[0] = sink.mean[0];
    R->weighted[0] = sink.weighted[0] + (double)self_w;
    R->max_lo[0] = R->max_hi[0] = sink.maxv[0];
  }
  R->exact = exact && sink.exact ? 1 : 0;
  *out = R;
  return 0;
}

extern "C" {

int tn93g_run(tn93g_ctx *c, const tn93g_set *S, const tn93g_params *P, tn93g_result **out)
{
  if (!c || !S || !P || !out) return TN93G_E_ARG;
  *out = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  const auto parts = split_parts(S, P);
  if (!parts.empty()) return run_split(c, S, P, parts, out);
  return run_one(c, S, P, -1, -1, nullptr, out);
}

int tn93g_run_noalloc(tn93g_ctx *c, const tn93g_set *S, const tn93g_params *P, tn93g_result *R,
                      int64_t link_capacity)
{
  if (!c || !S || !P || !R || link_capacity < 0) return TN93G_E_ARG;
  if (link_capacity > 0 &&
      (!R->link_row || !R->link_col || !R->link_dist || !R->link_dist_lo || !R->link_dist_hi))
    return c->fail(TN93G_E_ARG, "tn93g_run_noalloc: the result's five link arrays must hold link_capacity entries");
  if (!c->scratch)
    return c->fail(TN93G_E_HIP, fmt("tn93g_run_noalloc: no scratch on device %d (hipGetSymbolAddress: %s)", c->device,
                                    hipGetErrorString(c->scratch_err)));
  if ((P->flags & TN93G_FLAG_EXACT) || P->kind == TN93G_FST)
    return c->fail(TN93G_E_ARG, "tn93g_run_noalloc: exact-path runs (TN93G_FLAG_EXACT, TN93G_FST) need host and "
                                "device work memory: use tn93g_run");
  HIPCHK(c, hipSetDevice(c->device));
  Job J;
  J.S = S;
  J.P = *P;
  RunBufs B;  // every buffer points into the scratch; nothing to release
  J.B = &B;
  Arena ar;
  ar.base = (char *)c->scratch;
  ar.size = TN93G_NOALLOC_SCRATCH_BYTES;
  int rc = job_prepare(c, J, nullptr, &ar);
  if (rc) return rc;
  {  // the tuned choice of this shape when there is one (no trial runs here)
    int backend, rp;
    (void)tuned_choice(c, J, &backend, &rp);
    if ((rc = job_set_shape(c, J, backend, rp))) return rc;
  }
  job_enqueue(c, J, true, c->stream, 0);
  HIPCHK(c, hipGetLastError());
  // read-backs through fixed host buffers on the stack, each copy completed before the buffer is reused
  char cnt[STRIPE_BYTES];
  HIPCHK(c, hipMemcpyAsync(cnt, B.counters.p, STRIPE_BYTES, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  job_parse_counters(J, cnt);
  if (J.n_undecided)
    return c->fail(TN93G_E_ARG, "tn93g_run_noalloc: an output decision needs logarithms resolved on the host, which "
                                "needs work memory: use tn93g_run");
  if (J.max_region > J.link_cap)
    return c->fail(TN93G_E_ARG, fmt("tn93g_run_noalloc: %u links in one region, more than the scratch holds (%u): "
                                    "use tn93g_run", J.max_region, J.link_cap));
  if ((int64_t)J.n_linkrec > link_capacity)
    return c->fail(TN93G_E_ARG, fmt("tn93g_run_noalloc: %llu links, more than link_capacity (%lld)",
                                    (unsigned long long)J.n_linkrec, (long long)link_capacity));
  int64_t *lrow = R->link_row, *lcol = R->link_col;
  double *ld = R->link_dist, *llo = R->link_dist_lo, *lhi = R->link_dist_hi;
  memset(R, 0, sizeof *R);
  R->link_row = lrow;
  R->link_col = lcol;
  R->link_dist = ld;
  R->link_dist_lo = llo;
  R->link_dist_hi = lhi;
  const size_t nl = J.n_linkrec;
  R->n_links = (int64_t)nl;
  {
    constexpr size_t CH = 1536;  // 60 KB of link records per copy
    LinkRec buf[CH];
    for (size_t k0 = 0; k0 < nl; k0 += CH) {
      const size_t m = std::min(CH, nl - k0);
      HIPCHK(c, hipMemcpyAsync(buf, (const LinkRec *)B.links.p + k0, m * sizeof(LinkRec), hipMemcpyDeviceToHost,
                               c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t k = 0; k < m; k++) {
        lrow[k0 + k] = buf[k].i;
        lcol[k0 + k] = buf[k].j;
        ld[k0 + k] = buf[k].d;
        llo[k0 + k] = buf[k].lo;
        lhi[k0 + k] = buf[k].hi;
      }
    }
  }
  {
    constexpr int64_t CH = 1024;  // 64 KB of row accumulators per copy, merged in row order
    RowAcc buf[CH];
    RowMerge M;
    for (int64_t r0 = 0; r0 < J.nrows; r0 += CH) {
      const int64_t m = std::min(CH, J.nrows - r0);
      HIPCHK(c, hipMemcpyAsync(buf, (const RowAcc *)B.rowout.p + r0, (size_t)m * sizeof(RowAcc),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (int64_t r = 0; r < m; r++) M.add(buf[r], row_pairs(J.A, r0 + r));
    }
    M.finish(R, J.nrows, J.A.triangle ? self_weights(J.S, J.A.c0, J.A.n) : 0);
  }
  unsigned long long hrep[600];
  HIPCHK(c, hipMemcpyAsync(hrep, B.hist.p, sizeof hrep, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  job_finish_common(J, hrep, R);
  set_times(c, R, true);
  return 0;
}

int tn93g_run_batch(tn93g_ctx *c, int n, const tn93g_set *const *sets, const tn93g_params *params,
                    tn93g_result **out)
{
  if (!c || n < 0 || (n > 0 && (!sets || !params || !out))) return TN93G_E_ARG;
  for (int k = 0; k < n; k++) out[k] = nullptr;
  for (int k = 0; k < n; k++)
    if (!sets[k]) return TN93G_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (int rc = ensure_branches(c)) return rc;
  auto fail_all = [&](int rc) {
    for (int k = 0; k < n; k++) {
      tn93g_result_free(out[k]);
      out[k] = nullptr;
    }
    return rc;
  };
  // exact-path jobs (host work between kernels) run alone after the batch
  std::vector<int> idx, alone;
  for (int k = 0; k < n; k++)
    ((params[k].flags & TN93G_FLAG_EXACT) || params[k].kind == TN93G_FST || !split_parts(sets[k], &params[k]).empty()
         ? alone
         : idx)
        .push_back(k);
  const size_t m = idx.size();
  std::vector<Job> jobs(m);
  for (size_t t = 0; t < m; t++) {
    jobs[t].S = sets[idx[t]];
    jobs[t].P = params[idx[t]];
    jobs[t].B = c->slot(t);
    int rc = job_prepare(c, jobs[t]);
    if (!rc) rc = autotune(c, jobs[t]);
    if (rc) return fail_all(rc);
  }
  if (m) {
    // Pair kernels of jobs with the same fast-path variant run as one launch over all their tiles (small alignments
    // fill the device together): per group its RunArgs and tile offsets, in c->multi, uploaded before the graph.
    struct Group {
      int v, rp;
      std::vector<size_t> jobs;
      size_t off_args, off_toff;
      int64_t tiles;
    };
    std::vector<Group> groups;
    std::vector<char> merged(m, 0);
    std::vector<FillRange> fills;
    std::vector<CopyRange> copies;
    std::vector<size_t> spec(m, 0), soff(m, 0);
    std::vector<size_t> post;  // the merged jobs, whose tails and link kernels also run as one launch each
    size_t off_fills = 0, off_copies = 0, stage_bytes = 0, off_pargs = 0, off_tb = 0, off_ob = 0, off_rid = 0;
    std::vector<int32_t> rids[2];  // the merged jobs (indices into their RunArgs) whose replay is plain / dense
    int64_t tail_total = 0, order_total = 0;
    {
      std::map<int, std::vector<size_t>> byv;
      for (size_t t = 0; t < m; t++) {
        const int v = pairs_variant(jobs[t]);  // grouped by variant and tile rows
        if (v == 1 || v == 3 || v == 5 || v == 7 || v == 8 || v == 10) byv[v + 16 * jobs[t].rp].push_back(t);
      }
      size_t bytes = 0;
      for (auto &kv : byv) {
        if (kv.second.size() < 2) continue;
        Group G;
        G.v = kv.first % 16;
        G.rp = kv.first / 16;
        G.jobs = kv.second;
        G.off_args = bytes;
        bytes = (bytes + G.jobs.size() * sizeof(RunArgs) + 255) & ~(size_t)255;
        G.off_toff = bytes;
        bytes = (bytes + (G.jobs.size() + 1) * 8 + 255) & ~(size_t)255;
        G.tiles = 0;
        for (size_t t : G.jobs) G.tiles += jobs[t].gx * jobs[t].gy;
        if (G.tiles >= (1ll << 31)) continue;  // beyond one grid: these jobs launch one by one
        for (size_t t : G.jobs) merged[t] = 1;
        groups.push_back(G);
      }
      for (size_t t = 0; t < m; t++)
        if (merged[t]) post.push_back(t);
      off_pargs = bytes;
      bytes = (bytes + post.size() * sizeof(RunArgs) + 255) & ~(size_t)255;
      off_tb = bytes;
      bytes = (bytes + (post.size() + 1) * 8 + 255) & ~(size_t)255;
      off_ob = bytes;
      bytes = (bytes + (post.size() + 1) * 8 + 255) & ~(size_t)255;
      for (size_t q = 0; q < post.size(); q++) rids[jobs[post[q]].A.nplane ? 1 : 0].push_back((int32_t)q);
      off_rid = bytes;
      bytes = (bytes + (post.size() + 1) * 4 + 255) & ~(size_t)255;
      // every job's zero-started buffers, cleared by one launch before the branches fork
      for (size_t t = 0; t < m; t++) job_fill_ranges(jobs[t], fills);
      off_fills = bytes;
      bytes = (bytes + fills.size() * sizeof(FillRange) + 255) & ~(size_t)255;
      // the read-back of every job (counters, row accumulators, histogram replicas summed, the first spec[t] links)
      // staged in one device buffer by one launch at the end of the graph, then one transfer; the buffer begins with
      // the two wall-clock stamps around the merged pair kernels
      stage_bytes = 256;
      for (size_t t = 0; t < m; t++) {
        const Job &J = jobs[t];
        spec[t] = std::min<size_t>(J.B->link_spec, (size_t)J.link_cap * NSTRIPE);
        soff[t] = stage_bytes;
        stage_bytes += COUNTER_STRIDE + fast_layout(J, spec[t]).bytes;
      }
      if (ensure(c, c->stage, stage_bytes)) return fail_all(TN93G_E_NOMEM);
      for (size_t t = 0; t < m; t++) {
        const Job &J = jobs[t];
        char *d = (char *)c->stage.p + soff[t];
        const FastLayout F = fast_layout(J, spec[t]);
        copies.push_back(CopyRange{(const uint2 *)J.B->counters.p, (uint2 *)d, (int64_t)(STRIPE_BYTES / 8)});
        copies.push_back(CopyRange{(const uint2 *)J.B->rowout.p, (uint2 *)(d + COUNTER_STRIDE + F.off_ra),
                                   (int64_t)(J.nr1 * sizeof(RowAcc) / 8)});
        copies.push_back(CopyRange{(const uint2 *)J.B->hist.p, (uint2 *)(d + COUNTER_STRIDE + F.off_h), 600});
        if (spec[t])
          copies.push_back(CopyRange{(const uint2 *)J.B->links.p, (uint2 *)(d + COUNTER_STRIDE + F.off_lk),
                                     (int64_t)(spec[t] * sizeof(LinkRec) / 8)});
      }
      off_copies = bytes;
      bytes = (bytes + copies.size() * sizeof(CopyRange) + 255) & ~(size_t)255;
      if (bytes && ensure(c, c->multi, bytes)) return fail_all(TN93G_E_NOMEM);
      std::vector<unsigned char> host(bytes, 0);
      memcpy(host.data() + off_fills, fills.data(), fills.size() * sizeof(FillRange));
      memcpy(host.data() + off_copies, copies.data(), copies.size() * sizeof(CopyRange));
      for (const Group &G : groups) {
        RunArgs *ra = (RunArgs *)(host.data() + G.off_args);
        int64_t *to = (int64_t *)(host.data() + G.off_toff);
        int64_t acc = 0;
        for (size_t q = 0; q < G.jobs.size(); q++) {
          ra[q] = jobs[G.jobs[q]].A;
          to[q] = acc;
          acc += jobs[G.jobs[q]].gx * jobs[G.jobs[q]].gy;
        }
        to[G.jobs.size()] = acc;
      }
      {
        RunArgs *ra = (RunArgs *)(host.data() + off_pargs);
        int64_t *tb = (int64_t *)(host.data() + off_tb), *ob = (int64_t *)(host.data() + off_ob);
        for (size_t q = 0; q < post.size(); q++) {
          const Job &J = jobs[post[q]];
          ra[q] = J.A;
          tb[q] = tail_total;
          ob[q] = order_total;
          tail_total += tail_blocks(J);
          order_total += order_blocks(J);
        }
        tb[post.size()] = tail_total;
        ob[post.size()] = order_total;
        int32_t *ri = (int32_t *)(host.data() + off_rid);
        for (int d = 0; d < 2; d++)
          for (int32_t q : rids[d]) *ri++ = q;
      }
      if (bytes) HIPCHK(c, hipMemcpyAsync(c->multi.p, host.data(), bytes, hipMemcpyHostToDevice, c->stream));
    }
    // the batch's work as one graph; a batch that enqueues exactly the same launches relaunches the last graph
    std::vector<unsigned char> key;
    {
      const void *mp[2] = {c->multi.p, c->stage.p};  // both are baked into the graph's launches
      const unsigned char *a = (const unsigned char *)mp;
      key.insert(key.end(), a, a + sizeof mp);
    }
    for (const Job &J : jobs) {
      const unsigned char *a = (const unsigned char *)&J.A;
      key.insert(key.end(), a, a + sizeof(RunArgs));
      const int64_t dims[6] = {J.gx, J.gy, J.ntiles, J.nrows, (int64_t)J.nr1, (int64_t)J.fast * 4 + J.P.backend};
      const unsigned char *b = (const unsigned char *)dims;
      key.insert(key.end(), b, b + sizeof dims);
    }
    // TN93G_NO_GRAPH=1 issues the same launches directly on the streams instead of through a captured graph
    // (profilers that collect hardware counters per kernel do not follow graph launches)
    const bool eager = getenv("TN93G_NO_GRAPH") != nullptr;
    if (eager || !c->gexec || key != c->gkey) {
      if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
      c->gkey.clear();
      hipGraph_t g = nullptr;
      // the jobs are independent: spread over NBRANCH forked streams (largest first, each to the least loaded
      // branch) so that the graph has parallel branches and small jobs overlap on the device
      std::vector<size_t> order(m);
      for (size_t t = 0; t < m; t++) order[t] = t;
      auto cost = [&](const Job &J) { return (double)std::max<int64_t>(J.npairs, 1) * (double)(J.A.W + 8); };
      std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost(jobs[x]) > cost(jobs[y]); });
      double load[NBRANCH] = {};
      std::vector<int> branch(m, 0);
      for (size_t t : order) {
        const int b = (int)(std::min_element(load, load + NBRANCH) - load);
        load[b] += cost(jobs[t]);
        branch[t] = b;
      }
      auto fork = [&]() -> int {
        HIPCHK(c, hipEventRecord(c->fork, c->stream));
        for (int b = 0; b < NBRANCH; b++) HIPCHK(c, hipStreamWaitEvent(c->side[b], c->fork, 0));
        return 0;
      };
      auto join = [&]() -> int {
        for (int b = 0; b < NBRANCH; b++) {
          HIPCHK(c, hipEventRecord(c->join[b], c->side[b]));
          HIPCHK(c, hipStreamWaitEvent(c->stream, c->join[b], 0));
        }
        return 0;
      };
      if (eager)
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
      else
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      // 0: every job's fills; 1: each job's tables on its branch; 2: the merged pair kernels on the main stream, the
      // other jobs' whole sequence on their branches; 3: the merged jobs' replays, tails and link kernels
      hipLaunchKernelGGL(k_fill_ranges, dim3(64, (unsigned)std::min<size_t>(fills.size(), 65535)), dim3(256), 0,
                         c->stream, (const FillRange *)((char *)c->multi.p + off_fills), (int)fills.size());
      int rc = fork();
      for (size_t t : order) {
        if (rc) break;
        if (merged[t])
          job_enqueue_pre(c, jobs[t], false, c->side[branch[t]], false);
        else
          job_enqueue(c, jobs[t], false, c->side[branch[t]], branch[t], false);
      }
      if (!rc && !groups.empty()) {
        rc = join();
        unsigned long long *stamps = (unsigned long long *)c->stage.p;
        if (!rc) hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, c->stream, stamps);
        for (const Group &G : groups)
          if (!rc)
            (void)launch_pairs_multi(G.v, G.rp, (unsigned)G.tiles, c->stream,
                                     (const RunArgs *)((char *)c->multi.p + G.off_args),
                                     (const int64_t *)((char *)c->multi.p + G.off_toff), (int)G.jobs.size());
        if (!rc) hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, c->stream, stamps + 1);
        // 3: the merged jobs' replays as one launch per replay form (plain, dense N) on two branches, beside one
        // tail launch over all of them on the main stream; then their link offsets, gather, ordering and histogram
        // sums as one launch each
        if (!rc) rc = fork();
        const RunArgs *pa = (const RunArgs *)((char *)c->multi.p + off_pargs);
        const int np = (int)post.size();
        const int32_t *rid = (const int32_t *)((char *)c->multi.p + off_rid);
        if (!rc && !rids[0].empty())
          hipLaunchKernelGGL(k_replay_multi<false>, dim3(2048), dim3(RB), 0, c->side[0], pa, rid, (int)rids[0].size());
        if (!rc && !rids[1].empty())
          hipLaunchKernelGGL(k_replay_multi<true>, dim3(3072), dim3(RB), 0, c->side[1], pa, rid + rids[0].size(),
                             (int)rids[1].size());
        if (!rc && tail_total)
          hipLaunchKernelGGL(k_tail_multi, dim3((unsigned)tail_total), dim3(256), 0, c->stream, pa,
                             (const int64_t *)((char *)c->multi.p + off_tb), np);
        if (!rc) rc = join();
        if (!rc) {
          hipLaunchKernelGGL(k_scan_multi, dim3(np), dim3(1024), 0, c->stream, pa);
          hipLaunchKernelGGL(k_link_gather_multi, dim3(NSTRIPE, np), dim3(256), 0, c->stream, pa);
          hipLaunchKernelGGL(k_link_order_multi, dim3((unsigned)order_total), dim3(256), 0, c->stream, pa,
                             (const int64_t *)((char *)c->multi.p + off_ob), np);
          hipLaunchKernelGGL(k_hist_sum_multi, dim3(3, np), dim3(256), 0, c->stream, pa);
        }
      }
      if (!rc) rc = join();
      if (!rc)
        hipLaunchKernelGGL(k_copy_ranges, dim3(64, (unsigned)std::min<size_t>(copies.size(), 65535)), dim3(256), 0,
                           c->stream, (const CopyRange *)((char *)c->multi.p + off_copies), (int)copies.size());
      if (rc) {
        hipGraph_t dropped = nullptr;
        if (!eager) (void)hipStreamEndCapture(c->stream, &dropped);
        if (dropped) (void)hipGraphDestroy(dropped);
        return fail_all(rc);
      }
      if (!eager) {
        const hipError_t e1 = hipStreamEndCapture(c->stream, &g);
        if (e1 != hipSuccess)
          return fail_all(c->fail(TN93G_E_HIP, fmt("hipStreamEndCapture: %s", hipGetErrorString(e1))));
        const hipError_t e2 = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
          c->gexec = nullptr;
          return fail_all(c->fail(TN93G_E_HIP, fmt("hipGraphInstantiate: %s", hipGetErrorString(e2))));
        }
        c->gkey = key;
      }
    }
    if (!eager) {
      HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
      HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    // one transfer of the staged read-back of every job
    if (ensure_pinned(c, c->pin, std::max<size_t>(stage_bytes, 1))) return fail_all(TN93G_E_NOMEM);
    HIPCHK(c, hipMemcpyAsync(c->pin.p, c->stage.p, stage_bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[2]);
    double pair_ms = 0;  // the merged pair kernels, from the stamps around them
    if (!groups.empty() && c->wall_khz > 0) {
      const unsigned long long *stamps = (const unsigned long long *)c->pin.p;
      pair_ms = stamps[1] >= stamps[0] ? (double)(stamps[1] - stamps[0]) / (double)c->wall_khz : 0.0;
    }
    std::vector<char> redo(m, 0);
    // jobs with more links than staged: all their links again, each into a block of its own, in one more round
    std::vector<std::vector<LinkRec>> more(m);
    bool any_more = false;
    for (size_t t = 0; t < m; t++) {
      Job &J = jobs[t];
      job_parse_counters(J, (const char *)c->pin.p + soff[t]);
      if ((J.n_undecided && !J.A.exact) || (J.fast && J.max_region > J.link_cap)) {
        if (J.fast && J.max_region > J.link_cap) J.B->link_cap_hint = (int64_t)J.max_region;
        redo[t] = 1;
        continue;
      }
      J.B->link_spec = std::max<size_t>(4096, J.n_linkrec + J.n_linkrec / 4);
      if (J.n_linkrec > spec[t]) {
        more[t].resize(J.n_linkrec);
        HIPCHK(c, hipMemcpyAsync(more[t].data(), J.B->links.p, J.n_linkrec * sizeof(LinkRec), hipMemcpyDeviceToHost,
                                 c->stream));
        any_more = true;
      }
    }
    if (any_more) HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t t = 0; t < m; t++) {
      if (redo[t]) continue;
      tn93g_result *R = (tn93g_result *)calloc(1, sizeof(tn93g_result));
      if (!R) return fail_all(TN93G_E_NOMEM);
      out[idx[t]] = R;
      const char *pj = (const char *)c->pin.p + soff[t] + COUNTER_STRIDE;
      const LinkRec *lk = more[t].empty() ? (const LinkRec *)(pj + fast_layout(jobs[t], 0).off_lk) : more[t].data();
      int rc = job_fast_result(jobs[t], pj, lk, R);
      if (rc) return fail_all(rc);
      R->gpu_ms = ms;  // device time of the whole batch
      R->pair_kernel_ms = merged[t] ? pair_ms : 0.0;  // the merged pair launches of all merged jobs
    }
    // jobs whose outputs could not be certified, or whose links overflowed a region, run again alone
    for (size_t t = 0; t < m; t++)
      if (redo[t]) {
        int rc = tn93g_run(c, sets[idx[t]], &params[idx[t]], &out[idx[t]]);
        if (rc) return fail_all(rc);
      }
  }
  for (int k : alone) {
    int rc = tn93g_run(c, sets[k], &params[k], &out[k]);
    if (rc) return fail_all(rc);
  }
  return 0;
}

}  // extern "C"


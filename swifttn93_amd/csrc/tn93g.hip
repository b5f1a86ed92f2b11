// tn93g.hip -- MI355X (gfx950) all-pairs TN93 engine behind the C ABI of include/tn93g.h.
//
// Data layout in HBM (one uploaded set of n sequences, L sites; Ls = L rounded up to 16, npad = n rounded up to
// 64, plus 64):
//   codes   [n][Ls]    u8   ValidChars codes, row-major, padded with gaps        (replay kernel, j-side lookups)
//   codesT  [Ls][npad] u8   site-major copy                                     (i-side event lookups, coalesced)
//   planes  [n][W]     uint4 per 32 sites: clean, C|T bit, G|T bit, T bit       (clean/clean counts, VALU popcount)
//   onehot  [n][W64][4][64] u8 one-hot A,C,G,T per site, 64-site blocks      (clean/clean counts, int8 MFMA)
//   ev_off  [n+1] i32, ev [] u32 (site << 5 | code) ambiguity events (codes 4..16) in site order
// Per run: D [rows][cols] f64 (tn93's thisD of each pair), U [rows][cols] u16 (ulps below/above D that bound the
// value when a logarithm was uncertain), the histogram, per-row link counts, replay and host-exact lists.
//
// Kernels of a run (all on the context's stream):
//   k_pairs   clean/clean counts (popcount or MFMA) + ambiguity events + tests + FP64 tail + histogram + D
//   k_replay  site-order FP64 replay for pairs with order-dependent (non-dyadic) contributions or a -g fallback
//   k_patch   host-exact distances back into D (only after an uncertain logarithm could not be certified)
//   k_scan    per-row link offsets;  k_rows  per-row sums in column order, links in tn93's emission order
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tn93g.h"
#include "tn93g_math.h"

#define TN93G_VERSION "tn93g 0.1 (gfx950)"

namespace {

constexpr int NCODE = 18;
constexpr int GAPC = 17;
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
  int32_t icd;  // -1: none
  int32_t amb;
  int32_t pad;
};

// Event-table entry of the pair kernel: the effect scaled by 16 and aggregated into what the formula needs, packed
// as 16-bit fields: x = nf0 | nf1 << 16, y = nf2 | nf3 << 16, z = diag | AG << 16, w = CT | amb << 16 | nondyadic << 31.
// Exact whenever every value is a multiple of 1/16 (then all of tn93's float sums are exact and order-free).
static void event_entry(const SiteEffect &E, uint32_t out[4])
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
  out[0] = nf[0] | nf[1] << 16;
  out[1] = nf[2] | nf[3] << 16;
  out[2] = dg | ag << 16;
  out[3] = ct | (uint32_t)E.amb << 16 | (dyadic ? 0u : 1u << 31);
}

// ------------------------------------------------------------------------------------------ device structures
struct SetDev {
  int64_t n, L, Ls, npad, W, W64;
  uint8_t *codes, *codesT;
  uint4 *planes;
  uint8_t *onehot;
  int32_t *ev_off;
  uint32_t *ev;
  int64_t *weights;
  int64_t total_events;
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

struct RunArgs {
  // set
  const uint4 *planes;
  const uint8_t *onehot;
  const uint8_t *codes, *codesT;
  const int32_t *ev_off;
  const uint32_t *ev;
  const int64_t *weights;
  int64_t n, Ls, npad, W, W64;
  // pair space: rows [r0, r1) x cols [c0, c1), triangle => only j > i
  int64_t r0, r1, c0, c1, ncols, n1;
  int triangle, fst, gapmm, skip_events, avg_fallback, exact, all_slow;
  double rf, thr, mind, window;
  int64_t min_overlap;
  const uint4 *evtab;          // [18*18] packed event entries (main mode)
  const ReplayEnt *rtab;       // [2][18*18]: main mode, AVERAGE
  // outputs
  double *D;
  uint16_t *U;
  unsigned long long *hist;    // [3*200]
  int32_t *row_links;          // [nrows]
  unsigned long long *slow_list;
  unsigned int *slow_cnt;
  ExactRec *exact_list;
  unsigned int *exact_cnt;
  unsigned int exact_cap;
  unsigned int *stats;         // [0] uncertain pairs, [1] undecided pairs (decision not certified)
};

__device__ inline int fst_bin(const RunArgs &A, int64_t i, int64_t j)
{
  if (!A.fst) return 0;
  if (i < A.n1) return j >= A.n1 ? 2 : 0;
  return 1;
}

// Everything after the counts of one pair: tests, tail, histogram, link count, D/U (refkernel.cpp:732-815 and the
// phase C of swifttn93.cpp:2547-2603 up to the row sums, which k_rows does in column order).
__device__ void finish_pair(const RunArgs &A, int64_t i, int64_t j, const tn93g_tailin &tin,
                            unsigned long long *lhist)
{
  tn93g_tailout t = tn93g_tail(tin, A.window);
  const int64_t off = (i - A.r0) * A.ncols + (j - A.c0);
  if (A.exact && t.unc) {
    unsigned k = atomicAdd(A.exact_cnt, 1u);
    if (k < A.exact_cap) {
      A.exact_list[k].i = i;
      A.exact_list[k].j = j;
      A.exact_list[k].tin = tin;
    }
    return;
  }
  const int64_t w = A.weights[i] * A.weights[j];
  const int bin = fst_bin(A, i, j);
  bool decided = true;
  // histogram bin: floor(200 d) must agree over the enclosure (or clamp to 199 on the same side)
  double flo = floor(t.dlo * 200.0), fhi = floor(t.dhi * 200.0);
  if (flo != fhi && !(flo >= 200.0 && fhi >= 200.0) && !(flo < 0.0 && fhi < 0.0)) decided = false;
  atomicAdd(&lhist[bin * 200 + tn93g_hist_index(t.d)], (unsigned long long)w);
  double x = t.d <= 0. ? 0. : t.d;
  double xlo = t.dlo <= 0. ? 0. : t.dlo, xhi = t.dhi <= 0. ? 0. : t.dhi;
  bool in_lo = xlo >= A.mind && xlo <= A.thr, in_hi = xhi >= A.mind && xhi <= A.thr;
  bool cert_in = xlo >= A.mind && xhi <= A.thr;
  bool cert_out = xhi < A.mind || xlo > A.thr;
  (void)in_lo;
  (void)in_hi;
  if (!cert_in && !cert_out) decided = false;
  if (x >= A.mind && x <= A.thr) atomicAdd(&A.row_links[i - A.r0], 1);
  uint16_t u = 0;
  if (t.unc) {
    uint64_t b = tn93g_bits(x), bl = tn93g_bits(xlo), bh = tn93g_bits(xhi);
    uint64_t dn = b - bl, up = bh - b;
    if (dn > 255 || up > 255 || !(x == x)) decided = false;
    u = (uint16_t)((dn > 255 ? 255 : dn) | (up > 255 ? 255 : up) << 8);
    if (dn || up) atomicAdd(&A.stats[0], 1u);
  }
  if (!decided) atomicAdd(&A.stats[1], 1u);
  A.D[off] = x;
  A.U[off] = u;
}

// overlap too small: tn93's thisD is -1 (a link only if the caller's range reaches -1)
__device__ void skip_pair(const RunArgs &A, int64_t i, int64_t j)
{
  const int64_t off = (i - A.r0) * A.ncols + (j - A.c0);
  A.D[off] = -1.0;
  A.U[off] = 0;
  if (-1.0 >= A.mind && -1.0 <= A.thr) atomicAdd(&A.row_links[i - A.r0], 1);
}

// ------------------------------------------------------------------------------------------ pack kernels
__global__ void k_transpose(const uint8_t *__restrict__ codes, uint8_t *__restrict__ codesT, int64_t n, int64_t Ls,
                            int64_t npad)
{
  __shared__ uint8_t t[64][65];
  const int64_t s0 = (int64_t)blockIdx.y * 64, p0 = (int64_t)blockIdx.x * 64;
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
__global__ void k_planes(const uint8_t *__restrict__ codes, uint4 *__restrict__ planes, int64_t n, int64_t L,
                         int64_t Ls, int64_t W)
{
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * W) return;
  int64_t s = t / W, w = t % W;
  uint32_t c = 0, h = 0, g = 0, tt = 0;
  for (int k = 0; k < 32; k++) {
    int64_t p = w * 32 + k;
    if (p >= L) break;
    unsigned code = codes[s * Ls + p];
    if (code < 4) {
      c |= 1u << k;
      h |= (code & 1u) << k;
      g |= (code >> 1 & 1u) << k;
      tt |= (code == 3 ? 1u : 0u) << k;
    }
  }
  planes[t] = make_uint4(c, h, g, tt);
}

// one-hot bytes for the MFMA backend: per sequence, per 64-site block, 4 rows (A, C, G, T) of 64 bytes
__global__ void k_onehot(const uint8_t *__restrict__ codes, uint8_t *__restrict__ onehot, int64_t n, int64_t L,
                         int64_t Ls, int64_t W64)
{
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one thread per (sequence, site)
  int64_t total = n * W64 * 64;
  if (t >= total) return;
  int64_t s = t / (W64 * 64), p = t % (W64 * 64);
  unsigned code = p < L ? codes[s * Ls + p] : (unsigned)GAPC;
  int64_t blk = p / 64, k = p % 64;
  uint8_t *o = onehot + ((s * W64 + blk) * 4) * 64 + k;
  for (int b = 0; b < 4; b++) o[b * 64] = code == (unsigned)b ? 1 : 0;
}

__global__ void k_evcount(const uint8_t *__restrict__ codes, int32_t *__restrict__ cnt, int64_t L, int64_t Ls)
{
  const int64_t s = blockIdx.x;
  int c = 0;
  for (int64_t p = threadIdx.x; p < L; p += blockDim.x) {
    unsigned code = codes[s * Ls + p];
    c += (code >= 4 && code < (unsigned)GAPC) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if (threadIdx.x == 0) cnt[s] = c;
}

// events in site order: one wave per sequence
__global__ void k_evfill(const uint8_t *__restrict__ codes, const int32_t *__restrict__ ev_off,
                         uint32_t *__restrict__ ev, int64_t L, int64_t Ls)
{
  const int64_t s = blockIdx.x;
  const int lane = threadIdx.x;
  int base = ev_off[s];
  for (int64_t p0 = 0; p0 < L; p0 += 64) {
    int64_t p = p0 + lane;
    unsigned code = p < L ? codes[s * Ls + p] : (unsigned)GAPC;
    bool f = code >= 4 && code < (unsigned)GAPC;
    unsigned long long m = __ballot(f);
    if (f) {
      int k = __popcll(m & ((1ull << lane) - 1ull));
      ev[base + k] = (uint32_t)(p << 5) | code;
    }
    base += __popcll(m);
  }
}

// ------------------------------------------------------------------------------------------ pair kernel
// Tile: 32 rows x 64 columns, 256 threads = 4 waves; lane = column j, each wave owns 8 rows (uniform i).
constexpr int TI = 32, TJ = 64, KW = 16, RPW = 8;

struct Cnt {
  uint32_t T, DG, TS, CT, B0, B1, BT;
};

__device__ inline void count_word(const uint4 pi, const uint4 pj, Cnt &c)
{
  const uint32_t M = pi.x & pj.x;
  const uint32_t x = pi.y ^ pj.y, y = pi.z ^ pj.z;
  c.T += __popc(M);
  c.DG += __popc(M & ~(x | y));
  const uint32_t tr = M & y & ~x;
  c.TS += __popc(tr);
  c.CT += __popc(tr & pi.y);
  c.B0 += __popc(pj.x & pi.y) + __popc(pi.x & pj.y);
  c.B1 += __popc(pj.x & pi.z) + __popc(pi.x & pj.z);
  c.BT += __popc(pj.x & pi.w) + __popc(pi.x & pj.w);
}

// event accumulators: packed 16-bit fields plus 32-bit totals (flushed before a 16-bit field can overflow)
struct EvAcc {
  uint32_t px, py, pz, pw, fl;
};
struct Tot {
  uint32_t nf[4], dg, ag, ct, amb;
};

__device__ inline void ev_add(EvAcc &e, const uint4 t)
{
  e.px += t.x;
  e.py += t.y;
  e.pz += t.z;
  e.pw += t.w & 0x7fffffffu;
  e.fl |= t.w;
}
__device__ inline void ev_flush(EvAcc &e, Tot &o)
{
  o.nf[0] += e.px & 0xffffu;
  o.nf[1] += e.px >> 16;
  o.nf[2] += e.py & 0xffffu;
  o.nf[3] += e.py >> 16;
  o.dg += e.pz & 0xffffu;
  o.ag += e.pz >> 16;
  o.ct += e.pw & 0xffffu;
  o.amb += e.pw >> 16;
  e.px = e.py = e.pz = e.pw = 0;
}

template <int BACKEND>
__global__ __launch_bounds__(256) void k_pairs(RunArgs A)
{
  const int64_t i0 = A.r0 + (int64_t)blockIdx.y * TI, j0 = A.c0 + (int64_t)blockIdx.x * TJ;
  if (A.triangle && j0 + TJ - 1 <= i0) return;  // no pair j > i in this tile
  __shared__ uint4 sI[KW][TI];
  __shared__ uint4 sJ[KW][TJ];
  __shared__ uint4 sTab[NCODE * NCODE];
  __shared__ unsigned long long lhist[3 * 200];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int k = tid; k < NCODE * NCODE; k += 256) sTab[k] = A.evtab[k];
  for (int k = tid; k < 600; k += 256) lhist[k] = 0;
  const int64_t j = j0 + lane;
  const int64_t jl = j < A.n ? j : A.n - 1;
  const int64_t ib = i0 + wave * RPW;
  Cnt cnt[RPW];
#pragma unroll
  for (int a = 0; a < RPW; a++) cnt[a] = Cnt{0, 0, 0, 0, 0, 0, 0};
  const bool do_fast = !A.gapmm && !A.all_slow;
  if (do_fast) {
    for (int64_t w0 = 0; w0 < A.W; w0 += KW) {
      __syncthreads();
      for (int k = tid; k < KW * (TI + TJ); k += 256) {
        int s = k / KW, w = k % KW;  // consecutive threads: consecutive words of one sequence
        int64_t seq = s < TI ? i0 + s : j0 + (s - TI);
        uint4 v = make_uint4(0, 0, 0, 0);
        if (seq < A.n && w0 + w < A.W) v = A.planes[seq * A.W + w0 + w];
        if (s < TI)
          sI[w][s] = v;
        else
          sJ[w][s - TI] = v;
      }
      __syncthreads();
      const int kw = (int)min((int64_t)KW, A.W - w0);
      for (int k = 0; k < kw; k++) {
        const uint4 pj = sJ[k][lane];
#pragma unroll
        for (int a = 0; a < RPW; a++) count_word(sI[k][wave * RPW + a], pj, cnt[a]);
      }
    }
  }
  __syncthreads();
  // totals x16 from the clean/clean counts
  Tot tot[RPW];
  EvAcc acc[RPW];
#pragma unroll
  for (int a = 0; a < RPW; a++) {
    const Cnt &c = cnt[a];
    uint32_t nf3 = c.BT, nf1 = c.B0 - c.BT, nf2 = c.B1 - c.BT, nf0 = 2 * c.T - c.B0 - c.B1 + c.BT;
    tot[a].nf[0] = 16 * nf0;
    tot[a].nf[1] = 16 * nf1;
    tot[a].nf[2] = 16 * nf2;
    tot[a].nf[3] = 16 * nf3;
    tot[a].dg = 16 * c.DG;
    tot[a].ag = 16 * (c.TS - c.CT);
    tot[a].ct = 16 * c.CT;
    tot[a].amb = 0;
    acc[a] = EvAcc{0, 0, 0, 0, 0};
  }
  if (do_fast && !A.skip_events) {
    // events of each row i (uniform loop): the column's code at that site from the site-major copy
#pragma unroll
    for (int a = 0; a < RPW; a++) {
      const int64_t i = ib + a;
      if (i >= A.r1) break;
      const int e0 = A.ev_off[i], e1 = A.ev_off[i + 1];
      int since = 0;
      for (int e = e0; e < e1; e++) {
        const uint32_t ev = A.ev[e];
        const int64_t p = ev >> 5;
        const unsigned ca = ev & 31u;
        const unsigned cb = A.codesT[p * A.npad + jl];
        ev_add(acc[a], sTab[ca * NCODE + cb]);
        if (++since == 2047) {
          ev_flush(acc[a], tot[a]);
          since = 0;
        }
      }
      ev_flush(acc[a], tot[a]);
    }
    // events of the column j (per lane): the 8 rows' codes at that site in one 8-byte load; sites where the row
    // is itself ambiguous were counted above
    {
      const int f0 = A.ev_off[jl], f1 = A.ev_off[jl + 1];
      int since = 0;
      for (int e = f0; e < f1; e++) {
        const uint32_t ev = A.ev[e];
        const int64_t p = ev >> 5;
        const unsigned cb = ev & 31u;
        const uint2 ci = *reinterpret_cast<const uint2 *>(A.codesT + p * A.npad + ib);
#pragma unroll
        for (int a = 0; a < RPW; a++) {
          const unsigned ca = ((a < 4 ? ci.x : ci.y) >> (8 * (a & 3))) & 255u;
          if (ca < 4 || ca >= (unsigned)GAPC) ev_add(acc[a], sTab[ca * NCODE + cb]);
        }
        if (++since == 2047) {
#pragma unroll
          for (int a = 0; a < RPW; a++) ev_flush(acc[a], tot[a]);
          since = 0;
        }
      }
#pragma unroll
      for (int a = 0; a < RPW; a++) ev_flush(acc[a], tot[a]);
    }
  }
  // per pair: tests, replay routing, tail
#pragma unroll
  for (int a = 0; a < RPW; a++) {
    const int64_t i = ib + a;
    const bool valid = i < A.r1 && j < A.c1 && (!A.triangle || j > i);
    if (!valid) continue;
    if (!do_fast || (acc[a].fl >> 31)) {  // order-dependent: exact site-order replay
      unsigned k = atomicAdd(A.slow_cnt, 1u);
      A.slow_list[k] = (unsigned long long)i << 32 | (unsigned long long)j;
      continue;
    }
    const Tot &o = tot[a];
    tn93g_tailin tin;
    for (int k = 0; k < 4; k++) tin.nf[k] = (double)o.nf[k] * 0.0625;
    const double tng = (double)(o.nf[0] + o.nf[1] + o.nf[2] + o.nf[3]) * 0.03125;
    if (tng <= (double)A.min_overlap) {  // refkernel.cpp:732-735
      skip_pair(A, i, j);
      continue;
    }
    if (A.avg_fallback && tng * A.rf <= (double)o.amb) {  // refkernel.cpp:742-750: AVERAGE, via the replay
      unsigned k = atomicAdd(A.slow_cnt, 1u);
      A.slow_list[k] = (unsigned long long)i << 32 | (unsigned long long)j;
      continue;
    }
    tin.diag = (double)o.dg * 0.0625;
    tin.ag = (double)o.ag * 0.0625;
    tin.ct = (double)o.ct * 0.0625;
    finish_pair(A, i, j, tin, lhist);
  }
  __syncthreads();
  for (int k = tid; k < 600; k += 256)
    if (lhist[k]) atomicAdd(&A.hist[k], lhist[k]);
}

// ------------------------------------------------------------------------------------------ replay kernel
// computeTN93's scan (refkernel.cpp:384-729) for one pair per thread, in site order: clean/clean sites counted in
// integers, every other site applied through the handler table in FP64 in increasing site order, then tn93's merge.
constexpr int RB = 64;

__device__ bool replay_scan(const RunArgs &A, int64_t i, int64_t j, const ReplayEnt *__restrict__ tab,
                            double (*fc)[RB], int (*ic)[RB], int t, tn93g_tailin &tin, double &tng, uint32_t &amb)
{
  for (int k = 0; k < 16; k++) {
    fc[k][t] = 0.0;
    ic[k][t] = 0;
  }
  amb = 0;
  const uint4 *s1 = reinterpret_cast<const uint4 *>(A.codes + i * A.Ls);
  const uint4 *s2 = reinterpret_cast<const uint4 *>(A.codes + j * A.Ls);
  for (int64_t q = 0; q < A.Ls / 16; q++) {
    const uint4 a4 = s1[q], b4 = s2[q];
    const uint32_t aw[4] = {a4.x, a4.y, a4.z, a4.w}, bw[4] = {b4.x, b4.y, b4.z, b4.w};
    for (int wd = 0; wd < 4; wd++) {
      if (((aw[wd] | bw[wd]) & 0xFCFCFCFCu) == 0) {  // four clean/clean sites
        for (int k = 0; k < 4; k++) {
          unsigned c1 = aw[wd] >> (8 * k) & 255u, c2 = bw[wd] >> (8 * k) & 255u;
          ic[c1 * 4 + c2][t]++;
        }
        continue;
      }
      for (int k = 0; k < 4; k++) {
        unsigned c1 = aw[wd] >> (8 * k) & 255u, c2 = bw[wd] >> (8 * k) & 255u;
        if ((c1 | c2) < 4) {
          ic[c1 * 4 + c2][t]++;
          continue;
        }
        const ReplayEnt E = tab[c1 * NCODE + c2];
        uint32_t m = E.mask;
        while (m) {
          int c = __builtin_ctz(m);
          m &= m - 1;
          fc[c][t] += E.v;
        }
        if (E.icd >= 0) ic[5 * E.icd][t]++;
        amb += (uint32_t)E.amb;
      }
    }
  }
  // merge (refkernel.cpp:714-729)
  double nf[4] = {0., 0., 0., 0.};
  tng = 0.;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      double pc = (fc[4 * a + b][t] += (double)ic[4 * a + b][t]);
      tng += pc;
      nf[a] += pc;
      nf[b] += pc;
    }
  for (int k = 0; k < 4; k++) tin.nf[k] = nf[k];
  tin.diag = fc[0][t] + fc[5][t] + fc[10][t] + fc[15][t];
  tin.ag = fc[2][t] + fc[8][t];
  tin.ct = fc[7][t] + fc[13][t];
  return true;
}

__global__ __launch_bounds__(RB) void k_replay(RunArgs A)
{
  __shared__ double fc[16][RB];
  __shared__ int ic[16][RB];
  __shared__ unsigned long long lhist[3 * 200];
  const int t = threadIdx.x;
  for (int k = t; k < 600; k += RB) lhist[k] = 0;
  __syncthreads();
  const unsigned cnt = *A.slow_cnt;
  for (unsigned k = blockIdx.x * RB + t; k < cnt; k += gridDim.x * RB) {
    const unsigned long long pr = A.slow_list[k];
    const int64_t i = (int64_t)(pr >> 32), j = (int64_t)(pr & 0xffffffffull);
    tn93g_tailin tin;
    double tng;
    uint32_t amb;
    replay_scan(A, i, j, A.rtab, fc, ic, t, tin, tng, amb);
    if (tng <= (double)A.min_overlap) {
      skip_pair(A, i, j);
      continue;
    }
    if (A.avg_fallback && tng * A.rf <= (double)amb) {  // recompute in AVERAGE (refkernel.cpp:742-750)
      replay_scan(A, i, j, A.rtab + NCODE * NCODE, fc, ic, t, tin, tng, amb);
      if (tng <= (double)A.min_overlap) {
        skip_pair(A, i, j);
        continue;
      }
    }
    finish_pair(A, i, j, tin, lhist);
  }
  __syncthreads();
  for (int k = t; k < 600; k += RB)
    if (lhist[k]) atomicAdd(&A.hist[k], lhist[k]);
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

// exclusive scan of the per-row link counts (one block)
__global__ void k_scan(const int32_t *__restrict__ in, int64_t *__restrict__ out, int64_t n)
{
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t b = t * per, e = min(n, b + per);
  int64_t s = 0;
  for (int64_t k = b; k < e; k++) s += in[k];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int k = 0; k < 1024; k++) {
      int64_t v = part[k];
      part[k] = run;
      run += v;
    }
  }
  __syncthreads();
  int64_t run = part[t];
  for (int64_t k = b; k < e; k++) {
    out[k] = run;
    run += in[k];
  }
  if (t == 1023) out[n] = run;
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

__global__ void k_rows(RunArgs A, RowOut O)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nrows = A.r1 - A.r0;
  if (r >= nrows) return;
  const int64_t i = A.r0 + r;
  const int64_t lower = A.triangle ? i + 1 : A.c0;
  const int64_t wi = A.weights[i];
  double slo[3] = {0., 0., 0.}, shi[3] = {0., 0., 0.}, ws[3] = {0., 0., 0.}, mlo[3] = {0., 0., 0.},
         mhi[3] = {0., 0., 0.};
  if (A.triangle && wi > 1) {  // self_v_self (swifttn93.cpp:2406-2415)
    unsigned long long svs = (unsigned long long)((wi * (wi - 1)) >> 1);
    ws[(A.fst && i >= A.n1) ? 1 : 0] += (double)svs;
  }
  int64_t skipped = 0, linkw = 0, k = O.link_off[r];
  const int64_t kend = O.link_off[r + 1];
  const double *Drow = A.D + r * A.ncols - A.c0;
  const uint16_t *Urow = A.U + r * A.ncols - A.c0;
  for (int64_t j = lower; j < A.c1; j++) {
    const double x = Drow[j];
    const int bin = fst_bin(A, i, j);
    const int64_t w = wi * A.weights[j];
    const uint16_t u = Urow[j];
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
        double q = tl;
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
  for (int b = 0; b < 3; b++) {
    O.sum_lo[b * nrows + r] = slo[b];
    O.sum_hi[b * nrows + r] = shi[b];
    O.wsum[b * nrows + r] = ws[b];
    O.max_lo[b * nrows + r] = mlo[b];
    O.max_hi[b * nrows + r] = mhi[b];
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

struct tn93g_set {
  tn93g_ctx *ctx;
  SetDev d;
};

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};

struct tn93g_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4];
  std::string err;
  DevBuf D, U, hist, row_links, link_off, slow_list, counters, exact_list, rowout, links, evtab, rtab, patch;
  unsigned long long *pin_counters = nullptr;  // host pinned copy of the counters

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

static int ensure(tn93g_ctx *c, DevBuf &b, size_t bytes)
{
  if (b.n >= bytes && b.p) return 0;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.n = 0;
  size_t want = std::max(bytes, (size_t)256);
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) return c->fail(TN93G_E_NOMEM, fmt("hipMalloc(%zu): %s", want, hipGetErrorString(e)));
  b.n = want;
  return 0;
}

extern "C" {

const char *tn93g_version(void) { return TN93G_VERSION; }

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
  if (hipHostMalloc((void **)&c->pin_counters, 64 * sizeof(unsigned long long)) != hipSuccess) {
    delete c;
    return TN93G_E_NOMEM;
  }
  *out = c;
  return 0;
}

void tn93g_close(tn93g_ctx *c)
{
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (DevBuf *b : {&c->D, &c->U, &c->hist, &c->row_links, &c->link_off, &c->slow_list, &c->counters,
                    &c->exact_list, &c->rowout, &c->links, &c->evtab, &c->rtab, &c->patch})
    if (b->p) (void)hipFree(b->p);
  for (auto &e : c->ev) (void)hipEventDestroy(e);
  if (c->pin_counters) (void)hipHostFree(c->pin_counters);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *tn93g_error(const tn93g_ctx *c) { return c ? c->err.c_str() : "no context"; }

int tn93g_set_upload(tn93g_ctx *c, const uint8_t *codes, int64_t n, int64_t L, const int64_t *weights,
                     tn93g_set **out)
{
  if (!c || !out || (n > 0 && !codes) || n < 0 || L <= 0 || n >= (1ll << 31) || L >= (1ll << 26))
    return TN93G_E_ARG;
  for (int64_t k = 0; k < n * L; k++)
    if (codes[k] > GAPC) return c->fail(TN93G_E_ARG, "codes must be ValidChars indices 0..17");
  HIPCHK(c, hipSetDevice(c->device));
  tn93g_set *S = new tn93g_set();
  S->ctx = c;
  SetDev &d = S->d;
  memset(&d, 0, sizeof d);
  d.n = n;
  d.L = L;
  d.Ls = (L + 15) / 16 * 16;
  d.npad = (n + 63) / 64 * 64 + 64;
  d.W = (L + 31) / 32;
  d.W64 = (L + 63) / 64;
  const int64_t nn = std::max<int64_t>(n, 1);
  std::vector<uint8_t> padded((size_t)(nn * d.Ls), (uint8_t)GAPC);
  for (int64_t s = 0; s < n; s++) memcpy(&padded[(size_t)(s * d.Ls)], codes + s * L, (size_t)L);
  std::vector<int64_t> wts((size_t)nn, 1);
  if (weights)
    for (int64_t s = 0; s < n; s++) wts[(size_t)s] = weights[s];
  auto fail = [&](int code, const std::string &m) {
    tn93g_set_free(S);
    return c->fail(code, m);
  };
#define SCHK(x)                                                                              \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail(TN93G_E_HIP, fmt("%s: %s", #x, hipGetErrorString(e_))); \
  } while (0)
  SCHK(hipMalloc(&d.codes, padded.size()));
  SCHK(hipMalloc(&d.codesT, (size_t)(d.Ls * d.npad)));
  SCHK(hipMalloc(&d.planes, (size_t)(nn * d.W) * sizeof(uint4)));
  SCHK(hipMalloc(&d.onehot, (size_t)(nn * d.W64 * 256)));
  SCHK(hipMalloc(&d.ev_off, (size_t)(nn + 1) * sizeof(int32_t)));
  SCHK(hipMalloc(&d.weights, (size_t)nn * sizeof(int64_t)));
  SCHK(hipMemcpyAsync(d.codes, padded.data(), padded.size(), hipMemcpyHostToDevice, c->stream));
  SCHK(hipMemcpyAsync(d.weights, wts.data(), wts.size() * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  {
    dim3 g((unsigned)((d.Ls + 63) / 64), (unsigned)((d.npad + 63) / 64));
    hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, c->stream, d.codes, d.codesT, n, d.Ls, d.npad);
    int64_t tot = n * d.W;
    if (tot > 0)
      hipLaunchKernelGGL(k_planes, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, d.codes, d.planes,
                         n, L, d.Ls, d.W);
    int64_t toh = n * d.W64 * 64;
    if (toh > 0)
      hipLaunchKernelGGL(k_onehot, dim3((unsigned)((toh + 255) / 256)), dim3(256), 0, c->stream, d.codes, d.onehot,
                         n, L, d.Ls, d.W64);
  }
  int32_t *cnt_d = nullptr;
  SCHK(hipMalloc(&cnt_d, (size_t)nn * sizeof(int32_t)));
  if (n > 0) hipLaunchKernelGGL(k_evcount, dim3((unsigned)n), dim3(64), 0, c->stream, d.codes, cnt_d, L, d.Ls);
  std::vector<int32_t> cnt((size_t)nn, 0), off((size_t)nn + 1, 0);
  SCHK(hipMemcpyAsync(cnt.data(), cnt_d, (size_t)nn * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  SCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(cnt_d);
  int64_t total = 0;
  for (int64_t s = 0; s < n; s++) {
    off[(size_t)s] = (int32_t)total;
    total += cnt[(size_t)s];
    if (total >= (1ll << 31)) return fail(TN93G_E_ARG, "too many ambiguity events");
  }
  off[(size_t)n] = (int32_t)total;
  d.total_events = total;
  SCHK(hipMalloc(&d.ev, (size_t)std::max<int64_t>(total, 1) * sizeof(uint32_t)));
  SCHK(hipMemcpyAsync(d.ev_off, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  if (n > 0) hipLaunchKernelGGL(k_evfill, dim3((unsigned)n), dim3(64), 0, c->stream, d.codes, d.ev_off, d.ev, L, d.Ls);
  SCHK(hipGetLastError());
  SCHK(hipStreamSynchronize(c->stream));
#undef SCHK
  *out = S;
  return 0;
}

void tn93g_set_free(tn93g_set *S)
{
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  (void)hipStreamSynchronize(S->ctx->stream);
  void *ps[] = {S->d.codes, S->d.codesT, S->d.planes, S->d.onehot, S->d.ev_off, S->d.ev, S->d.weights};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  delete S;
}

void tn93g_result_free(tn93g_result *r)
{
  if (!r) return;
  free(r->link_row);
  free(r->link_col);
  free(r->link_dist);
  free(r->link_dist_lo);
  free(r->link_dist_hi);
  free(r);
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

int tn93g_run(tn93g_ctx *c, const tn93g_set *S, const tn93g_params *P, tn93g_result **out)
{
  if (!c || !S || !P || !out) return TN93G_E_ARG;
  *out = nullptr;
  const SetDev &d = S->d;
  if (P->mode < 0 || P->mode > 4 || P->kind < 0 || P->kind > 2) return c->fail(TN93G_E_ARG, "bad mode or kind");
  if (P->kind != TN93G_TRIANGLE && (P->n1 < 0 || P->n1 > d.n)) return c->fail(TN93G_E_ARG, "bad n1");
  HIPCHK(c, hipSetDevice(c->device));
  RunArgs A;
  memset(&A, 0, sizeof A);
  A.planes = d.planes;
  A.onehot = d.onehot;
  A.codes = d.codes;
  A.codesT = d.codesT;
  A.ev_off = d.ev_off;
  A.ev = d.ev;
  A.weights = d.weights;
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
  } else {
    A.r0 = 0;
    A.r1 = d.n;
    A.c0 = 0;
    A.c1 = d.n;
    A.triangle = 1;
  }
  A.fst = P->kind == TN93G_FST;
  A.n1 = P->kind == TN93G_TRIANGLE ? d.n : P->n1;
  A.ncols = A.c1 - A.c0;
  const int64_t nrows = A.r1 - A.r0;
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
  const int64_t npairs_box = nrows * A.ncols;
  int64_t npairs = A.triangle ? d.n * (d.n - 1) / 2 : npairs_box;
  if (npairs >= (1ll << 32)) return c->fail(TN93G_E_ARG, "more than 2^32 pairs in one run");
  // tables
  {
    std::vector<uint4> et(NCODE * NCODE);
    std::vector<ReplayEnt> rt(2 * NCODE * NCODE);
    for (int a = 0; a < NCODE; a++)
      for (int b = 0; b < NCODE; b++) {
        SiteEffect E = site_effect(a, b, P->mode, P->subset_mask);
        uint32_t q[4];
        event_entry(E, q);
        et[(size_t)(a * NCODE + b)] = make_uint4(q[0], q[1], q[2], q[3]);
        for (int t = 0; t < 2; t++) {
          SiteEffect F = t ? site_effect(a, b, TN93G_AVERAGE, 0) : E;
          ReplayEnt &r = rt[(size_t)(t * NCODE * NCODE + a * NCODE + b)];
          r.v = F.v;
          r.mask = F.mask;
          r.icd = F.icd;
          r.amb = F.amb;
          r.pad = 0;
        }
      }
    if (ensure(c, c->evtab, et.size() * sizeof(uint4)) || ensure(c, c->rtab, rt.size() * sizeof(ReplayEnt)))
      return TN93G_E_NOMEM;
    HIPCHK(c, hipMemcpyAsync(c->evtab.p, et.data(), et.size() * sizeof(uint4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->rtab.p, rt.data(), rt.size() * sizeof(ReplayEnt), hipMemcpyHostToDevice, c->stream));
  }
  A.evtab = (const uint4 *)c->evtab.p;
  A.rtab = (const ReplayEnt *)c->rtab.p;
  // buffers
  const size_t nbox = (size_t)std::max<int64_t>(npairs_box, 1);
  if (ensure(c, c->D, nbox * 8) || ensure(c, c->U, nbox * 2) || ensure(c, c->hist, 600 * 8) ||
      ensure(c, c->row_links, (size_t)std::max<int64_t>(nrows, 1) * 4) ||
      ensure(c, c->link_off, (size_t)(nrows + 1) * 8) ||
      ensure(c, c->slow_list, (size_t)std::max<int64_t>(npairs, 1) * 8) || ensure(c, c->counters, 64 * 8) ||
      ensure(c, c->rowout, (size_t)std::max<int64_t>(nrows, 1) * (15 * 8 + 16)))
    return TN93G_E_NOMEM;
  const unsigned exact_cap = A.exact ? (unsigned)std::min<int64_t>(npairs, 1ll << 31) : 1u;
  if (ensure(c, c->exact_list, (size_t)exact_cap * sizeof(ExactRec))) return TN93G_E_NOMEM;
  A.D = (double *)c->D.p;
  A.U = (uint16_t *)c->U.p;
  A.hist = (unsigned long long *)c->hist.p;
  A.row_links = (int32_t *)c->row_links.p;
  A.slow_list = (unsigned long long *)c->slow_list.p;
  unsigned *ctr = (unsigned *)c->counters.p;
  A.slow_cnt = ctr + 0;
  A.exact_cnt = ctr + 1;
  A.stats = ctr + 2;
  A.exact_list = (ExactRec *)c->exact_list.p;
  A.exact_cap = exact_cap;
  HIPCHK(c, hipMemsetAsync(c->hist.p, 0, 600 * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(c->counters.p, 0, 64 * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(c->row_links.p, 0, (size_t)std::max<int64_t>(nrows, 1) * 4, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  if (nrows > 0 && A.ncols > 0) {
    dim3 g((unsigned)((A.ncols + TJ - 1) / TJ), (unsigned)((nrows + TI - 1) / TI));
    hipLaunchKernelGGL(k_pairs<TN93G_BACKEND_POPC>, g, dim3(256), 0, c->stream, A);
  }
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  hipLaunchKernelGGL(k_replay, dim3(2048), dim3(RB), 0, c->stream, A);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->pin_counters, c->counters.p, 64, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const unsigned *hc = (const unsigned *)c->pin_counters;
  const unsigned n_slow = hc[0], n_exact = hc[1], n_unc = hc[2], n_undecided = hc[3];
  if (n_undecided && !A.exact) {
    // an output decision could not be certified from the enclosures: run again resolving logs on the host
    tn93g_params Q = *P;
    Q.flags |= TN93G_FLAG_EXACT;
    return tn93g_run(c, S, &Q, out);
  }
  if (A.exact && n_exact) {
    if (n_exact > exact_cap) return c->fail(TN93G_E_NOMEM, "host-exact list overflow");
    std::vector<ExactRec> ex(n_exact);
    HIPCHK(c, hipMemcpyAsync(ex.data(), c->exact_list.p, n_exact * sizeof(ExactRec), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<PatchRec> pr(n_exact);
    auto work = [&](size_t b, size_t e) {
      for (size_t k = b; k < e; k++) {
        double dd = host_tail(ex[k].tin);
        double x = dd <= 0. ? 0. : dd;
        pr[k].i = ex[k].i;
        pr[k].j = ex[k].j;
        pr[k].thisD = x;
        pr[k].idx = host_hist_index(dd);
        pr[k].link = x >= P->min_distance && x <= P->threshold;
      }
    };
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n_exact < 4096) nth = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++) th.emplace_back(work, (size_t)n_exact * t / nth, (size_t)n_exact * (t + 1) / nth);
    for (auto &t : th) t.join();
    if (ensure(c, c->patch, pr.size() * sizeof(PatchRec))) return TN93G_E_NOMEM;
    HIPCHK(c, hipMemcpyAsync(c->patch.p, pr.data(), pr.size() * sizeof(PatchRec), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_patch, dim3((unsigned)((n_exact + 255) / 256)), dim3(256), 0, c->stream, A,
                       (const PatchRec *)c->patch.p, (int64_t)n_exact);
  }
  // link offsets, then the rows
  int64_t *loff = (int64_t *)c->link_off.p;
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, (const int32_t *)c->row_links.p, loff, nrows);
  int64_t n_links = 0;
  HIPCHK(c, hipMemcpyAsync(&n_links, loff + nrows, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ensure(c, c->links, (size_t)std::max<int64_t>(n_links, 1) * 40)) return TN93G_E_NOMEM;
  RowOut O;
  {
    double *base = (double *)c->rowout.p;
    const size_t nr = (size_t)std::max<int64_t>(nrows, 1);
    O.sum_lo = base;
    O.sum_hi = base + 3 * nr;
    O.wsum = base + 6 * nr;
    O.max_lo = base + 9 * nr;
    O.max_hi = base + 12 * nr;
    O.skipped = (int64_t *)(base + 15 * nr);
    O.linkw = (int64_t *)(base + 16 * nr);
    const size_t nl = (size_t)std::max<int64_t>(n_links, 1);
    O.link_row = (int64_t *)c->links.p;
    O.link_col = O.link_row + nl;
    O.link_d = (double *)(O.link_col + nl);
    O.link_lo = O.link_d + nl;
    O.link_hi = O.link_lo + nl;
    O.link_off = loff;
  }
  if (nrows > 0) hipLaunchKernelGGL(k_rows, dim3((unsigned)((nrows + 63) / 64)), dim3(64), 0, c->stream, A, O);
  HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  HIPCHK(c, hipGetLastError());
  // results
  tn93g_result *R = (tn93g_result *)calloc(1, sizeof(tn93g_result));
  if (!R) return TN93G_E_NOMEM;
  const size_t nl = (size_t)n_links;
  R->n_links = n_links;
  R->link_row = (int64_t *)malloc(std::max<size_t>(nl, 1) * 8);
  R->link_col = (int64_t *)malloc(std::max<size_t>(nl, 1) * 8);
  R->link_dist = (double *)malloc(std::max<size_t>(nl, 1) * 8);
  R->link_dist_lo = (double *)malloc(std::max<size_t>(nl, 1) * 8);
  R->link_dist_hi = (double *)malloc(std::max<size_t>(nl, 1) * 8);
  std::vector<double> rows((size_t)std::max<int64_t>(nrows, 1) * 17);
  if (nl) {
    HIPCHK(c, hipMemcpyAsync(R->link_row, O.link_row, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_col, O.link_col, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_dist, O.link_d, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_dist_lo, O.link_lo, nl * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(R->link_dist_hi, O.link_hi, nl * 8, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(rows.data(), c->rowout.p, rows.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(R->hist, c->hist.p, 600 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms_all = 0, ms_pair = 0;
  (void)hipEventElapsedTime(&ms_all, c->ev[0], c->ev[2]);
  (void)hipEventElapsedTime(&ms_pair, c->ev[0], c->ev[1]);
  R->gpu_ms = ms_all;
  R->pair_kernel_ms = ms_pair;
  // row-order merge (swifttn93.cpp:2832-2848)
  const size_t nr = (size_t)std::max<int64_t>(nrows, 1);
  for (int b = 0; b < 3; b++) {
    R->mean_lo[b] = R->mean_hi[b] = R->weighted[b] = R->max_lo[b] = R->max_hi[b] = 0.;
  }
  for (int64_t r = 0; r < nrows; r++) {
    for (int b = 0; b < 3; b++) {
      const double mlo = rows[9 * nr + b * nr + (size_t)r], mhi = rows[12 * nr + b * nr + (size_t)r];
      if (mlo > R->max_lo[b]) R->max_lo[b] = mlo;
      if (mhi > R->max_hi[b]) R->max_hi[b] = mhi;
      R->mean_lo[b] += rows[b * nr + (size_t)r];
      R->mean_hi[b] += rows[3 * nr + b * nr + (size_t)r];
      R->weighted[b] += rows[6 * nr + b * nr + (size_t)r];
    }
    int64_t sk, lw;
    memcpy(&sk, &rows[15 * nr + (size_t)r], 8);
    memcpy(&lw, &rows[16 * nr + (size_t)r], 8);
    R->skipped += sk;
    R->found_links += lw;
  }
  R->exact = (n_unc == 0) ? 1 : 0;
  R->nrows = (int32_t)nrows;
  R->n_replayed = n_slow;
  R->n_uncertain = n_unc;
  R->n_host_exact = n_exact;
  // self_v_self histogram entries (swifttn93.cpp:2406-2415): the row kernel adds their weights
  if (A.triangle) {
    std::vector<int64_t> w((size_t)std::max<int64_t>(d.n, 1));
    HIPCHK(c, hipMemcpy(w.data(), d.weights, (size_t)d.n * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < d.n; i++)
      if (w[(size_t)i] > 1) R->hist[(A.fst && i >= A.n1) ? 1 : 0][0] += (unsigned long long)((w[(size_t)i] * (w[(size_t)i] - 1)) >> 1);
  }
  *out = R;
  return 0;
}

}  // extern "C"

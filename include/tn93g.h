/* tn93g.h -- C ABI of the MI355X all-pairs TN93 engine (libtn93g.so).
 *
 * The engine computes, on one gfx950 GPU, exactly what tn93 v1.1.0's standard all-pairs mode computes: for every
 * pair of aligned nucleotide sequences the TN93 distance (byte-identical to single-thread tn93 once printed with
 * %g), the 200-bin histogram(s), the links at or below the threshold, and the per-row sums that tn93 reduces in row
 * order.  Plain pointers and sizes only; no C++ or GPU types cross this boundary.
 *
 * It replaces the reference's inner boundary (paths relative to the reference repository):
 *   Engine::prepare(const SeqSet&, long L, long n1, int mode, double rf, long overlap)   swifttn93.cpp:570
 *   the per (row, 8-column block) kernels + phaseA + replay + tail8a/b + phaseC           swifttn93.cpp:2610-2806
 *   ref::computeTN93(s1, s2, L, matchMode, NULL, min_overlap, histogram, 200, 200,
 *                    count1, count2, d1, d2, 0.0)                                         refkernel.h:58-62
 * The process interface above it (tn93's command line, argparse.cpp:138-226) is the CLI tn93g_cli.cpp.
 *
 * Conventions: every function returns 0 on success and a negative TN93G_E* code on failure; the library never
 * exits.  One host thread drives a context.  Codes are the reference's ValidChars indices
 * ("ACGTURYSWKMBDHVN?-" -> 0..17, refkernel.cpp:9), one byte per site.
 */
#ifndef TN93G_H
#define TN93G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TN93G_OK 0
#define TN93G_E_ARG (-1)    /* bad argument */
#define TN93G_E_HIP (-2)    /* HIP runtime error (message in tn93g_error) */
#define TN93G_E_NOMEM (-3)  /* device or host allocation failed */
#define TN93G_E_NODEV (-4)  /* no usable gfx950 device */

/* ambiguity modes (refkernel.h:14-18) */
#define TN93G_RESOLVE 0
#define TN93G_AVERAGE 1
#define TN93G_SKIP 2
#define TN93G_GAPMM 3
#define TN93G_SUBSET 4

/* pair sets */
#define TN93G_TRIANGLE 0  /* single file: pairs i < j over the set (swifttn93.cpp:2392)            */
#define TN93G_CROSS 1     /* -s without -m: rows [0, n1) x columns [n1, n) (swifttn93.cpp:2321-2322) */
#define TN93G_FST 2       /* -s -m: triangle over both files, three histograms (swifttn93.cpp:2680-2686) */
#define TN93G_APPEND 3    /* pairs i < j with j in [n1, n2): sequences appended to a set against all earlier ones
                             and each other, exactly the pairs a triangle over [0, n2) adds to one over [0, n1)  */

/* backends of the clean/clean site counts */
#define TN93G_BACKEND_AUTO 0
#define TN93G_BACKEND_POPC 1 /* VALU popcount over bit planes */
#define TN93G_BACKEND_MFMA 2 /* int8 one-hot contraction on MFMA, int32 accumulation */
#define TN93G_BACKEND_GEMM 3 /* reported only (backend_used): the clean/clean counts of a long dense-N alignment as
                                int8 GEMMs (rocBLAS) on the matrix cores, chosen by the library for single runs */

/* run flags */
#define TN93G_FLAG_EXACT 1u      /* resolve every uncertain log on the host (libm) instead of bracketing it   */
#define TN93G_FLAG_ALL_SLOW 2u   /* route every pair through the site-order replay kernel (tests)            */
#define TN93G_FLAG_AUTOTUNE 4u   /* time the pair kernel's free choices (TN93G_BACKEND_AUTO, tile_rows 0) on the
                                    first job of each shape and keep the fastest for the context               */

typedef struct tn93g_ctx tn93g_ctx;
typedef struct tn93g_set tn93g_set;

/* Opens device `device` (HIP ordinal).  Replaces nothing in the reference (which is CPU-only); it owns the stream,
 * the device buffers reused across runs and the last error message. */
int tn93g_open(int device, tn93g_ctx **out);
void tn93g_close(tn93g_ctx *ctx);
/* Last error message of the context (never NULL). */
const char *tn93g_error(const tn93g_ctx *ctx);

/* Copies n sequences of L codes (row-major, n*L bytes, values 0..17) to the device and packs them (bit planes,
 * site-major codes, per-sequence ambiguity event lists).  Replaces Engine::prepare (swifttn93.cpp:570-1044).
 * The host keeps ownership of `codes`.  weights: n copy numbers (swifttn93.cpp:324-333), or NULL for all 1.
 * Equivalent to tn93g_set_create(ctx, L, n, out) followed by tn93g_set_append(*out, codes, n, weights). */
int tn93g_set_upload(tn93g_ctx *ctx, const uint8_t *codes, int64_t n, int64_t L, const int64_t *weights,
                     tn93g_set **out);
/* An empty set of sequences of L sites with room for `capacity` sequences, kept on the device: a reference panel
 * that query batches are appended to.  No counterpart in the reference, which reads every input once per run. */
int tn93g_set_create(tn93g_ctx *ctx, int64_t L, int64_t capacity, tn93g_set **out);
/* Appends n sequences (as in tn93g_set_upload) after the set's current ones.  Runs on the set from other threads
 * or contexts may be in flight: each run sees the set as it was when it started (TN93G_APPEND runs name their
 * range explicitly).  Fails with TN93G_E_ARG beyond the capacity. */
int tn93g_set_append(tn93g_set *set, const uint8_t *codes, int64_t n, const int64_t *weights);
/* Number of sequences in the set. */
int64_t tn93g_set_size(const tn93g_set *set);
void tn93g_set_free(tn93g_set *set);

typedef struct {
  int32_t mode;            /* TN93G_RESOLVE .. TN93G_SUBSET                                     */
  int32_t kind;            /* TN93G_TRIANGLE, TN93G_CROSS or TN93G_FST                          */
  int64_t n1;              /* sequences of the first file (CROSS, FST); first appended (APPEND) */
  uint32_t subset_mask;    /* SUBSET: bit c set when code c (4..16) is resolved (-a RY ...)     */
  int32_t backend;         /* TN93G_BACKEND_*                                                   */
  double resolve_fraction; /* -g (1.0 = off)                                                    */
  int64_t min_overlap;     /* -l, as tn93 casts it to long                                      */
  double threshold;        /* -t: links have min_distance <= d <= threshold                    */
  double min_distance;     /* -w (default -1e-10)                                               */
  uint32_t flags;          /* TN93G_FLAG_*                                                      */
  int64_t n2;              /* APPEND: end of the appended range (0: the set's size)             */
  int32_t tile_rows;       /* rows of the pair kernel's tiles of 64 columns: 8, 16 or 32; 0: the autotuner's
                              choice for the job's shape (TN93G_FLAG_AUTOTUNE), else 16.  The MFMA backend and
                              the exact path tile 16 rows.  The reference tiles 32 rows by blocks of 8 columns
                              (swifttn93.cpp:2348, 2440); the results never depend on the shape                 */
} tn93g_params;

/* Result of one run.  All arrays are owned by the result.  Everything tn93 prints is here:
 *   hist[b][k]           the histograms (b = 0 single set; FST: 0 file 1, 1 file 2, 2 between), u64 wrap-around;
 *   links                (row, col, dist) in tn93's emission order (row-major, columns ascending); col is the
 *                        global sequence index (the second file continues the numbering, as csvn prints it);
 *   mean, weighted, max  tn93's row-order reductions (swifttn93.cpp:2839-2848) of the per-row sums;
 *   found_links          sum of link weights; skipped = pairs with overlap <= min_overlap.
 * Floating values that depend on a GPU logarithm are exact when `exact` is 1.  Otherwise each comes as an
 * enclosure [lo, hi] that contains tn93's value (lo == hi when it is known exactly); link_dist lies in it.  The
 * histogram bins and the link decisions are certified on the device (an undecidable pair makes the run repeat
 * itself in exact mode).  Whether a link distance prints as tn93's does is checked by tn93g_result_certify; a
 * caller that finds lo and hi print differently calls tn93g_run again with TN93G_FLAG_EXACT. */
typedef struct {
  int32_t exact;
  int32_t nrows;
  uint64_t hist[3][200];
  int64_t n_links;
  int64_t *link_row, *link_col;
  double *link_dist, *link_dist_lo, *link_dist_hi;
  double mean_lo[3], mean_hi[3];
  double weighted[3];
  double max_lo[3], max_hi[3];
  int64_t found_links;
  int64_t skipped;
  /* diagnostics */
  int64_t n_replayed;   /* pairs that went through the site-order replay kernel        */
  int64_t n_uncertain;  /* pairs whose distance is an enclosure (lo < hi)               */
  int64_t n_host_exact; /* pairs whose logarithms were resolved on the host (EXACT)     */
  double gpu_ms;        /* device time of the run's kernels (HIP events)                 */
  double pair_kernel_ms;/* device time of the pair kernel (clean/clean counts + events); in a batch, the time of
                           the one pair-kernel launch shared by the jobs of the same kernel variant (all such launches
                           together, from the device clock), 0 for a job that ran its own                          */
  double tail_kernel_ms;/* device time of the FP64 tail kernel                            */
  int32_t backend_used;  /* the pair kernel's backend (TN93G_BACKEND_POPC or _MFMA) and the rows of its tiles: the */
  int32_t tile_rows_used;/* caller's or the autotuner's choice                                                    */
  int64_t folded_columns;/* constant columns (the same clean base in every sequence) that tn93g_set_upload left out
                            of the device arrays, their sites added to every pair's counts (swifttn93.cpp:591-660) */
} tn93g_result;

/* Device memory (bytes) that tn93g_run of this set and these parameters works in, beyond the set itself (for a run
 * computed in parts, its largest part's): what a caller admitting runs under an HBM budget reserves.  No counterpart
 * in the reference. */
int tn93g_run_bytes(tn93g_ctx *ctx, const tn93g_set *set, const tn93g_params *params, int64_t *bytes);

/* Computes all pairs of `set` described by `params` (set must hold n1 + n2 sequences for CROSS/FST).
 * Replaces the OpenMP tile loop of swifttn93's main (swifttn93.cpp:2339-2864).  *out is allocated by the library
 * and released with tn93g_result_free.  A run of 2^31 pairs or more (a triangle of 65,536 sequences and up), or of
 * more than 2^25 (row, 64-column tile) slots, is computed in parts of consecutive rows (at most 2^30 pairs and 2^24
 * slots each, about 25 GB of device memory) and returned as one result, the same as one run's; TN93G_FST runs are
 * not split and hold fewer than 2^32 pairs.  Like the reference, there is no other limit on the number of pairs. */
int tn93g_run(tn93g_ctx *ctx, const tn93g_set *set, const tn93g_params *params, tn93g_result **out);
void tn93g_result_free(tn93g_result *res);

/* Scratch of tn93g_run_noalloc: one device array of this many bytes per GPU, held from the library's loading until
 * the process ends. */
#define TN93G_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)

/* tn93g_run without any allocation: no counterpart in the reference.  For callers that cannot allocate device or
 * host memory during a run (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every work buffer of the run (chunk slots, replay list, row accumulators, link regions, histogram replicas,
 * counters, event tables) is carved from the device's TN93G_NOALLOC_SCRATCH_BYTES scratch array; read-backs pass
 * through fixed buffers on the caller's stack.  The caller passes the result: `res` is overwritten, except that its
 * link_row, link_col, link_dist, link_dist_lo and link_dist_hi must point to link_capacity entries each, which
 * receive the links.  The values are those tn93g_run gives (the same kernels; the mean's enclosure may differ in
 * its last bits, both contain tn93's value).
 * Fails with TN93G_E_ARG, writing no link, and tn93g_error names the reason, when: the run's buffers exceed the
 * scratch (a triangle of up to about 850 sequences of fewer than 2,048 columns fits); it has more links than
 * link_capacity or than the scratch's link regions hold; it is an exact-path run (TN93G_FLAG_EXACT, TN93G_FST); or an
 * output decision needs logarithms resolved on the host.  In each case tn93g_run computes it.  With
 * TN93G_BACKEND_AUTO the backend is the context's tuned choice for the shape, else popcount (no trial runs).
 * Calls share their device's scratch, which keeps the last call's contents (a run clears what it needs cleared):
 * calls that overlap on one GPU, from any thread or context, race and must be serialised by the caller. */
int tn93g_run_noalloc(tn93g_ctx *ctx, const tn93g_set *set, const tn93g_params *params, tn93g_result *res,
                      int64_t link_capacity);

/* 1 when every link distance of the result prints the same with "%.{precision}g" at both ends of its enclosure
 * (always for an exact result), 0 otherwise, or TN93G_E_ARG.  tn93 prints distances with %g (precision 6,
 * swifttn93.cpp:2575-2585 through fastfmt.h:75): a caller that prints them, or hands them on as tn93's values, and gets 0
 * runs again with TN93G_FLAG_EXACT.  The check the reference needs none of: it computes every logarithm with libm. */
int tn93g_result_certify(const tn93g_result *res, int precision);

/* Runs n independent jobs (sets[k] with params[k]; sets may repeat) as one batch: every job's kernels are enqueued
 * back to back on the context's stream, captured into one HIP graph and launched once (a later batch that enqueues
 * exactly the same work launches the same graph again), then all results are read back together.  out[k] is the
 * result tn93g_run(ctx, sets[k], &params[k], &out[k]) would give; jobs with TN93G_FLAG_EXACT or TN93G_FST, and jobs
 * whose outputs cannot be certified from the batch, are run alone after it.  gpu_ms of a batched result is the
 * device time of the whole batch.  Replaces the per-alignment loop a caller of the reference would run
 * (one swifttn93 process per alignment, swifttn93.cpp:2222). */
int tn93g_run_batch(tn93g_ctx *ctx, int n, const tn93g_set *const *sets, const tn93g_params *params,
                    tn93g_result **out);

/* Library version string. */
const char *tn93g_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TN93G_H */

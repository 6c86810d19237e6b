/*
 * vdhmm -- MI355X PairHMM engine, C ABI.
 *
 * Two groups of entry points:
 *
 * 1. Drop-ins with the exact names, argument order and return conventions of
 *    the reference's native PairHMM libraries (caller/native/caller), so that
 *    the reference's ctypes loaders can bind this library instead:
 *      pairhmm_pairs             batch_hmm.cpp:192   (ctypes: cpu_batch.py:14-17)
 *      trie_many                 trie_vector.cpp:284 (ctypes: trie_vector.py:11-13)
 *      read_haplotype            wide.cpp:9          (ctypes: accelerator.py:50-51)
 *      read_haplotype_batch      wide.cpp:185        (ctypes: whole_pairhmm.py:43-44)
 *      read_haplotype_batch_mode wide.cpp:349        (ctypes: whole_pairhmm.py:41-42)
 *      wide_capability           wide.cpp:342        (ctypes: whole_pairhmm.py:45-46)
 *      anchored_pairhmm          anchored.cpp:5      (ctypes: indel_refine.py:56-57)
 *    base/core/core.cpp:9,176 exports a second read_haplotype /
 *    read_haplotype_batch (mode-0 semantics); those are vd_core_read_haplotype
 *    and vd_core_read_haplotype_batch here, and libvdhmm_core.so re-exports
 *    them under the reference names.
 *    All return natural-log likelihoods in FP64: -inf for an empty read or
 *    haplotype; trie_many returns NaN where the likelihood is <= 1e-200
 *    (trie_vector.cpp:277), as the reference does.
 *
 * 2. The batched region engine (vd_forward*): many active regions in one
 *    call, CSR layout, explicit gap / continuation / per-quality-bin error
 *    parameters, FP32 bulk with FP64 rescue or all-FP64.
 *
 * The library keeps one HIP context per process and is not re-entrant across
 * host threads: calls on one process are serialised by an internal mutex.
 */
#ifndef VDHMM_H
#define VDHMM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- drop-ins ---------------------------------------------------------- */

/* batch_hmm.cpp:192.  Dense per-pair likelihoods, modes 0/1/2.  Returns 1 if
 * count<0 or mode not in {0,1,2}, 2 on a device error, else 0.  `lanes` is
 * accepted and ignored (it selects a CPU SIMD width in the reference and does
 * not change values).  Computed in FP64 on the GPU. */
int pairhmm_pairs(const char *const *reads, const unsigned char *const *quals, const int *rn,
                  const char *const *haps, const int *hn, int count, int mode, int lanes,
                  double *output);

/* trie_vector.cpp:284.  nr reads x nh haplotypes, out row-major nr x nh.
 * *cells receives the reference's evaluated-cell counter (AVX-512 lane
 * grouping).  Returns 1 if nr<0, nh<0 or mode not in {0,1,2}.  Precision:
 * FP32 + FP64 rescue unless VARDEPTH_CUDA_PRECISION=float64. */
int trie_many(const char *const *reads, const unsigned char *const *quals, const int *rn, int nr,
              const char *const *haps, const int *hn, int nh, int mode, int lanes, double *out,
              uint64_t *cells);

/* wide.cpp:9 -- semiglobal (mode 2 model) with explicit gap / ext. FP64. */
double read_haplotype(const char *read, const unsigned char *qual, int n, const char *hap, int m,
                      double gap, double ext);
/* wide.cpp:185 -- one read against count haplotypes. FP64. */
void read_haplotype_batch(const char *read, const unsigned char *qual, int n, const char *const *hap,
                          const int *lengths, int count, double gap, double ext, double *out);
/* wide.cpp:349 -- as above; `mode` is the reference's SIMD lane request
 * (0/2/4/8) and does not change values. */
void read_haplotype_batch_mode(const char *read, const unsigned char *qual, int n,
                               const char *const *haps, const int *lengths, int count, double gap,
                               double ext, double *out, int mode);
/* wide.cpp:342 -- widest lane request the batch entry honours (8). */
int wide_capability(void);
/* anchored.cpp:5 -- global/anchored (mode 1) with explicit gap / ext. FP64. */
double anchored_pairhmm(const char *read, const unsigned char *qual, int n, const char *hap, int m,
                        double gap, double extension);
/* base/core/core.cpp:9 / :176 -- semiglobal, per-row normalised (mode 0). FP64. */
double vd_core_read_haplotype(const char *read, const unsigned char *qual, int n, const char *hap,
                              int m, double gap, double ext);
void vd_core_read_haplotype_batch(const char *read, const unsigned char *qual, int n,
                                  const char *const *hap, const int *lengths, int count, double gap,
                                  double ext, double *out);

/* ---- batched region engine --------------------------------------------- */

enum { VD_MODE_SEMIGLOBAL = 0, VD_MODE_ANCHORED = 1, VD_MODE_SEMIGLOBAL_RESCALE = 2 };
enum { VD_PREC_MIXED = 0, VD_PREC_FP64 = 1 };

typedef struct vd_params {
  double gap;      /* gap open: M->I and M->D */
  double ext;      /* gap continuation: I->I and D->D */
  double err[64];  /* error probability of quality bin min(q, 63) */
} vd_params;

/* Region batch.  Read i has bases/quals [read_off[i], read_off[i+1]);
 * haplotype h has bases [hap_off[h], hap_off[h+1]); region k owns reads
 * [region_reads[k], region_reads[k+1]) and haplotypes
 * [region_haps[k], region_haps[k+1]) and writes the row-major R_k x H_k
 * block at out + out_off[k]. */
typedef struct vd_batch {
  const uint8_t *read_bytes;
  const uint8_t *qual_bytes;
  const int64_t *read_off;
  const uint8_t *hap_bytes;
  const int64_t *hap_off;
  const int64_t *region_reads;
  const int64_t *region_haps;
  const int64_t *out_off;
  int64_t nregions, nreads, nhaps;
} vd_batch;

typedef struct vd_stats {
  uint64_t pairs;        /* read x haplotype pairs in the batch */
  uint64_t cells;        /* sum of n*m (accelerator.py:155 dp_cells) */
  uint64_t fp32_units;   /* FP32 work units (4 pairs each) */
  uint64_t flagged;      /* FP32 pairs sent to the FP64 rescue */
  uint64_t fp64_pairs;   /* pairs computed in FP64 (rescued + long + forced) */
  uint64_t fp64_bad;     /* FP64 results that were not finite (should be 0) */
  uint64_t overflow;     /* FP64 list overflow (must be 0) */
  double device_ms;      /* GPU time of the batch (events) */
  double plan_ms;        /* GPU time of the planning kernels */
  double fwd32_ms;       /* GPU time of the FP32 anti-diagonal kernel */
  double fwd64_ms;       /* GPU time of the FP64 kernel (rescue / long / forced) */
  double total_ms;       /* wall time of the call */
  uint64_t fp32_cells_swept;  /* read rows x haplotype columns the FP32 kernels computed
                                 (below the pairs' n*m when haplotype prefixes are shared) */
} vd_stats;

/* Reference error table 10^(-min(q,60)/10) for the 64 bins, and gap / ext. */
void vd_params_default(vd_params *p, double gap, double ext);

/* Host pointers in, host out.  Returns 0, or 2 with vd_last_error() set. */
int vd_forward(const vd_batch *batch, const vd_params *params, int mode, int precision, double *out,
               vd_stats *stats);

/* Device pointers in and out (already resident in HBM), launched on `stream`
 * (a hipStream_t, NULL = default stream).  npairs = out_off[nregions];
 * max_hap_len bounds every haplotype length.  If stats is non-NULL the call
 * synchronises the stream and fills it; with stats NULL it only enqueues.
 *
 * Graph capture: a call on a capturing stream records its work into the
 * graph on the workspaces and parameter tables that the eager (uncaptured)
 * call made right before it left; it allocates and uploads nothing, and
 * fails (2, vd_last_error) when it would need to -- a larger batch, other
 * parameters, or buffers that an earlier capture holds.  The graph keeps
 * those buffers for the life of the process: later calls allocate their own,
 * so replays reproduce the captured call after any other work.  Replays of
 * one graph must not overlap each other.  The same holds for
 * vd_grad_device and the genotype *_device entry points. */
int vd_forward_device(const vd_batch *dev_batch, const vd_params *params, int mode, int precision,
                      int64_t npairs, int max_hap_len, double *dev_out, void *stream,
                      vd_stats *stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the forward of vd_forward_device with the same arguments
 * (no stats), every scratch buffer -- plan, FP64 work list, counters,
 * parameter tables, FP64 tile rows -- taken from one static device array of
 * 64 MiB per GPU, held by the library until the process ends.  The call
 * allocates and frees nothing and copies nothing from the host: it only
 * enqueues kernels on `stream`.  Limits: 13 bytes per read, 17 per pair and
 * 204 per region plus one FP64 tile workspace of 128 * (max_hap_len + 48)
 * bytes must fit in the 64 MiB (about 3.5 M pairs of 150-600 bp haplotypes);
 * a larger batch is refused (2, vd_last_error) before anything is enqueued.
 * The forward runs without haplotype prefix sharing and on one stream; its
 * results equal vd_forward_device's with vd_suffix_sharing off (bit for bit)
 * and agree to FP32 rounding with it on.  Every call on one GPU shares that
 * GPU's array: calls whose device work overlaps (other streams, no ordering
 * by the caller) race. */
int vd_forward_device_noalloc(const vd_batch *dev_batch, const vd_params *params, int mode, int precision,
                              int64_t npairs, int max_hap_len, double *dev_out, void *stream);

/* ---- gradients (no counterpart in the reference) ------------------------ */

/* Gradient of sum_p w_p ln L_p over every pair of the batch with respect to
 * the model parameters, computed by the adjoint kernels (FP32 with FP64 for
 * the pairs the FP32 pass flags).  grad[0] = d/dgap, grad[1] = d/dext,
 * grad[2 + q] = d/derr[q] (q = 0..63).  weights: one per output slot, NULL =
 * all 1.  out (optional): the ln likelihoods, as vd_forward(mixed) returns.
 * Mode 2 is differentiated as mode 0 (same model).  Reads of up to 256 rows
 * (256 bases, 255 anchored) run in the register-resident kernels; longer
 * reads run in the tiled FP64 adjoint (64-row tiles, each tile's boundary row
 * kept from the forward, the tiles' adjoints chained bottom to top).  The
 * error-table bins enter through the rescaled emissions with each bin's
 * normaliser held fixed; the likelihood does not depend on that choice, so
 * the derivatives at fixed normalisers are the gradient.
 * Host pointers; returns 0, or 2 with vd_last_error() set. */
int vd_grad(const vd_batch *batch, const vd_params *params, int mode, const double *weights, double *out,
            double *grad, vd_stats *stats);

/* Same with every buffer already in HBM, enqueued on `stream`; max_read_len
 * bounds every read length.  dev_grad receives the 66 doubles. */
int vd_grad_device(const vd_batch *dev_batch, const vd_params *params, int mode, int64_t npairs,
                   int max_hap_len, int max_read_len, const double *dev_weights, double *dev_out,
                   double *dev_grad, void *stream, vd_stats *stats);

/* ---- tuning (no counterpart in the reference) ---------------------------- */

/* Launch choices per rows-per-lane class r = 1..16 (index 0 unused), picked by
 * vardepth_amd/autotune.py from timed sweeps:
 *   fwd_unroll          anti-diagonals per iteration of the forward main loop (2 or 4)
 *   fwd_blocks_per_cu   persistent forward grid per CU (0 = the occupancy limit)
 *   grad_variant        adjoint register block: 0 = 2 steps (two waves per SIMD
 *                       where it fits), 1 = 4 steps
 *   grad_blocks_per_cu  persistent adjoint grid per CU (0 = the occupancy limit)
 *   fwd_share           0: pair units; 1: haplotype prefix sharing for the class
 *                       (read units of share_chunk haplotypes); 2: prefix sharing
 *                       and the shared-suffix continuation for every read unit;
 *                       3: the suffix kernel for the read units of regions with
 *                       a shared suffix, the prefix-only kernel (which holds
 *                       more waves in some classes) for the others.  Applies
 *                       while vd_prefix_sharing is on (the suffix part while
 *                       vd_suffix_sharing is on)
 *   share_chunk         haplotypes per read unit under prefix sharing (1..64)
 * Defaults: 2, 0, 0, 0, 2, 4, 1 (0 with VARDEPTH_CLASS_OVERLAP=0), 0.  Values
 * never change results beyond rounding order (the suffix continuation agrees
 * with full sweeps to FP32 rounding), except fwd_route64, which gives the
 * routed pairs their FP64 values. */
typedef struct vd_tuning {
  int fwd_unroll[17];
  int fwd_blocks_per_cu[17];
  int grad_variant[17];
  int grad_blocks_per_cu[17];
  int fwd_share[17];
  int share_chunk;
  int class_overlap;  /* 1: consecutive FP32 class kernels alternate between two streams
                         (vd_class_overlap); 0: one stream */
  int fwd_route64[17];  /* 1: with precision VD_PREC_MIXED, the reads of this class skip
                           the FP32 pass and go straight to the FP64 rescue (for
                           batches where FP32 would flag most of them) */
  int fwd_pack;         /* 1: every class runs pair units whose 4 pairs may come from
                           different regions (no prefix sharing): for batches of
                           regions with few pairs, where region-bound units would
                           leave most lanes idle; 0 (default): fwd_share decides */
} vd_tuning;

int vd_tuning_get(vd_tuning *t);
int vd_tuning_set(const vd_tuning *t);
/* Haplotype prefix sharing of the FP32 forward (default on; no counterpart
 * call in the reference, whose CPU trie_many shares prefixes the same way,
 * trie_vector.cpp:204-210,259-273): each read sweeps its region's haplotypes in byte
 * order and resumes every haplotype from the column it shares with the
 * previous one, with results identical to full sweeps.  on = 1 / 0 sets it,
 * on < 0 only queries; returns the previous setting.  The FP64 path and the
 * gradients always sweep in full. */
int vd_prefix_sharing(int on);
/* Shared-suffix continuation of the prefix-sharing forward (default on;
 * applies while vd_prefix_sharing is on; the reference's CPU trie_many does
 * the same reduction, trie_vector.cpp:230-257,274-275): when every haplotype
 * of a region ends in the same >= 16 bases, each read runs one adjoint
 * (backward) sweep over them, and every haplotype is swept only up to the
 * suffix, its terminal completed by a dot product with the adjoint.  Results
 * equal full sweeps to FP32 rounding (the adjoint and the FP64 dot product
 * reorder the sums).  on = 1 / 0 sets it, on < 0 only queries; returns the
 * previous setting. */
int vd_suffix_sharing(int on);
/* Launch option with no counterpart in the reference (its CPU path has no
 * streams): when on (the default; VARDEPTH_CLASS_OVERLAP=0 in the environment
 * at start-up turns it off), consecutive FP32 class kernels alternate between
 * the caller's stream and a second stream of the library, forked from and
 * joined back into the caller's stream within the call, so that one
 * persistent kernel's last units overlap the next kernel's first.  Results
 * are identical either way.  Calls that time each class (vd_class_timing)
 * keep one stream.  on = 1 / 0 sets it, on < 0 only queries; returns the
 * previous setting. */
int vd_class_overlap(int on);
/* When on, calls made with a stats pointer also time each class kernel;
 * vd_class_ms returns the last such times (17 each, index = class). */
void vd_class_timing(int on);
int vd_class_ms(double *fwd_ms, double *grad_ms);

/* ---- diploid genotype likelihood ---------------------------------------- */

/* Replaces mixture_general (caller/native/caller/mixture_general.cpp:9, bound
 * by mixture_cpu.py:18-21): same arguments and output.  lh is rows x n
 * row-major (finite, >= 0); prior < 0: no mixture, 0: allele-balance mixture,
 * > 0: balance and artifact mixture (requires prior < 1); triangular != 0:
 * n(n+1)/2 outputs in the order b = 0..n-1, a = 0..b, else n*n a-major.  Host
 * pointers.  Returns 0, or 2 with vd_last_error() set (the reference returns
 * void; mixture_cpu.py:25 rejects the same inputs before the call). */
int vd_mixture_general(int rows, int n, const double *lh, double prior, int triangular, double *out);

/* Replaces mixture_two (mixture_two.cpp:9, bound by mixture_cpu.py:8-11):
 * rows x 2 matrix, out[2a+b] for the ordered genotypes.  Host pointers. */
int vd_mixture_two(int rows, const double *lh, double prior, double *out);

/* Replaces molecule_rows.normalized_rows (molecule_rows.py:6-14): values =
 * (1-e) exp(ll - rowmax) + e mean, e = 10^(-min(mq,60)/10), certainty =
 * ptp(log(max(1e-300, values))) per row.  Host pointers, rows x n. */
int vd_normalized_rows(int rows, int n, const double *ll, const unsigned char *mq, double *values,
                       double *certainty);

/* A batch of genotype regions in device memory.  Region k: rows
 * row_off[k]..row_off[k+1] (global row indices into mq / the per-row
 * outputs), haplotype columns col_off[k]..col_off[k+1], a row-major
 * rows x columns block at mat_off[k] of the input matrix, and its pair
 * outputs at gl_off[k].  mat_size / gl_size bound every block; a region
 * whose block falls outside them is skipped and counted.  A vd_batch of the
 * forward maps onto it as row_off = region_reads, col_off = region_haps,
 * mat_off = out_off. */
typedef struct vd_geno_batch {
  const int64_t *row_off;
  const int64_t *col_off;
  const int64_t *mat_off;
  const int64_t *gl_off;
  int64_t nregions, nrows, mat_size, gl_size;
} vd_geno_batch;

typedef struct vd_geno_stats {
  int64_t units;        /* mixture wave units (64 pairs each) */
  int64_t kept_rows;    /* rows that passed the certainty filter */
  int64_t invalid;      /* regions with a non-finite or negative matrix entry: outputs NaN */
  int64_t skipped;      /* regions outside the batch bounds: outputs untouched */
  double molecule_ms, mixture_ms, total_ms;
} vd_geno_stats;

/* The diploid likelihood of every region of a batch of matrices (device
 * pointers; the batched form of vd_mixture_general with triangular order or
 * square).  keep (optional, per row): rows with keep == 0 are left out.
 * stream: a hipStream_t, NULL = the default stream, as for vd_forward_device
 * (so work the caller enqueued on the same stream, such as the forward that
 * wrote the likelihoods, is complete before it runs); the call only enqueues.
 * stats (optional) synchronises the stream and fills the counters and times. */
int vd_mixture_regions_device(const vd_geno_batch *batch, const double *lh, const signed char *keep,
                              double prior, int triangular, double *gl, void *stream,
                              vd_geno_stats *stats);

/* The single-end read-evidence chain of whole_read.py:30-49 on the device:
 * for each region of a log-likelihood batch ll (the forward's output layout),
 * normalized rows with per-row mapping quality mq, rows with certainty below
 * min_certainty dropped (whole_read.py:37 uses ln 3; NaN is kept), molecule
 * rows exp(log(max(1e-300, v)) - max), then the triangular diploid likelihood
 * with the given prior (whole_read.py:49 uses 1e-3) into gl.  values
 * (rows x columns, like ll) and certainty (per row) are optional outputs.
 * stream and stats as for vd_mixture_regions_device. */
int vd_genotype_device(const vd_geno_batch *batch, const double *ll, const unsigned char *mq,
                       double min_certainty, double prior, double *gl, double *values, double *certainty,
                       void *stream, vd_geno_stats *stats);

/* Near-tie guard of partial_anchor.py:119-133 on a genotype batch (device
 * pointers): per region, gaps[2k] = largest - second largest of gl, gaps[2k+1]
 * the same of gl + log(prior), and flags[k] = 1 when the smaller gap is
 * <= margin_scale * max(1, molecules), molecules = rows whose certainty (as
 * written by vd_genotype_device) is not below min_certainty.  The prior:
 * with allele_freq (one per haplotype column of the batch, indexed like
 * col_off) that of partial_anchor.py:56-57 -- each frequency floored at
 * .001 and normalised over the region, prior(a, b) = f_a f_b (1 if a == b
 * else 2) normalised; with allele_freq NULL the uniform haplotype prior of
 * whole_read.py:51 (1 for a == b, 2 otherwise, over n^2).
 * partial_anchor.py uses margin_scale 1e-4 and recomputes a flagged locus in
 * FP64.  Regions with fewer than two pairs are not flagged.  nflagged
 * (optional) synchronises and returns the count. */
int vd_genotype_guard(const vd_geno_batch *batch, const double *gl, const double *certainty,
                      const double *allele_freq, double min_certainty, double margin_scale, signed char *flags,
                      double *gaps, void *stream, int64_t *nflagged);

/* ---- baseline ----------------------------------------------------------- */

/* The plain GPU forward: the first max_pairs pairs of the batch (region
 * order), one kernel launch per pair, one workgroup with one thread per read
 * row, FP32 without rescue.  A measurement baseline for the batched engine,
 * not a production path.  *gpu_ms: device time from the first launch to the
 * last completion (inputs already staged).  out[i] is NaN for pairs not run.
 * Reads of at most 256 rows. */
int vd_forward_plain(const vd_batch *batch, const vd_params *params, int mode, int64_t max_pairs,
                     double *out, double *gpu_ms);

/* The reference's own GPU design, batched, as a second measurement
 * baseline (no counterpart call in the reference; it restates the strategy
 * of pairhmm_cuda.cu as dispatched by accelerator.py:80-159 in its "mixed"
 * precision): one thread per pair sweeping the read rows, the pair's DP row
 * in global memory with columns interleaved across the pairs of a launch,
 * FP32 values with an FP64 log scale, a non-anchored row rescaled by 2^64
 * below 2^-40, an anchored row normalised by its maximum, and NaN for a
 * terminal below 2^-40 (where the reference falls back to FP64).  Pairs run
 * in (read, haplotype) length buckets of 32 bases, each in launches whose
 * workspace stays within budget_mb MiB.  out: every pair of the batch
 * (ln, NaN, or -inf for an empty read or haplotype).  *gpu_ms: device time
 * of the launches (inputs staged before); *launches: their count. */
int vd_forward_refdesign(const vd_batch *batch, const vd_params *params, int mode, int64_t budget_mb,
                         double *out, double *gpu_ms, int64_t *launches);

/* Longest haplotype (bases) a batch may hold: the forward (gradients = 0)
 * and gradient (gradients = 1) kernels keep each work unit's haplotype codes
 * in one workgroup's LDS (160 KiB on gfx950: 20,104 bases for the forward,
 * 19,112 for gradients, whose workgroups also stage a checkpoint block).  Calls with
 * a longer haplotype fail with vd_last_error() set before touching the device.
 * No counterpart in the reference (its CPU kernels have no such bound). */
int vd_max_hap_len(int gradients);

/* Message of the last failed call on this process. */
const char *vd_last_error(void);
/* Device index used by the library (VARDEPTH_CUDA_DEVICE, default 0). */
int vd_device(void);

#ifdef __cplusplus
}
#endif
#endif

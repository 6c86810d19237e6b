/*
 * CPU FP64 oracle for the diploid genotype likelihood over haplotype pairs.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP genotype kernels and the
 * CPU baseline of their benchmark.  Nothing in the product library links,
 * loads or calls it.
 *
 * Each function restates the cited reference lines (caller/native/caller/)
 * in plain C, IEEE double, contraction disabled, same operation order.  The
 * reference takes logarithms through libmvec's vector log on AVX-512 hosts
 * (vector_log.hpp:8-14) and through the scalar libm log elsewhere; this
 * restatement uses the scalar libm log throughout, so results agree with a
 * reference build to the last bits of each log term, not bit for bit.
 */
#ifndef VD_ORACLE_GENOTYPE_H
#define VD_ORACLE_GENOTYPE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* molecule_rows.py:4-14.  values[r][h] = (1-e) exp(ll - rowmax) + e mean,
 * e = 10^(-min(mq,60)/10); certainty[r] = ptp(log(max(1e-300, values[r]))).
 * NaN propagates as in numpy. */
void vdo_normalized_rows(int rows, int n, const double *ll, const unsigned char *mq,
                         double *values, double *certainty);

/* mixture_general.cpp:9-40: rows x n matrix lh (row-major, finite, >= 0),
 * prior < 0: no mixture (mean of the pair), prior == 0: allele-balance
 * mixture only, prior > 0: balance and artifact mixture.  triangular != 0:
 * pairs (a,b) for b in 0..n-1, a in 0..b; otherwise a-major n x n. */
void vdo_mixture_general(int rows, int n, const double *lh, double prior, int triangular,
                         double *out);

/* mixture_two.cpp:9-29: rows x 2 matrix, ordered genotypes g = 2a+b. */
void vdo_mixture_two(int rows, const double *lh, double prior, double *out4);

/* The single-end evidence chain of whole_read.py:30-49 over a region batch:
 * per region k (rows row_off[k]..row_off[k+1], ncol[k] haplotypes, the
 * row-major log-likelihood block at ll + mat_off[k]), normalized rows with
 * per-row mapping quality mq[row], drop rows whose certainty is below
 * min_certainty (whole_read.py:36-37; NaN certainty is kept, as there),
 * molecule rows exp(log(max(1e-300, v)) - max) (whole_read.py:45-46), then
 * the triangular diploid likelihood with the given prior (triangular.py:6-11)
 * into gl + gl_off[k].  A region whose kept rows are not all finite and
 * non-negative gets NaN (the reference raises).  Regions are handed to
 * `threads` workers largest-first.  Returns the number of such regions. */
int64_t vdo_genotype_regions(int64_t nregions, const int64_t *row_off, const int32_t *ncol,
                             const int64_t *mat_off, const double *ll, const unsigned char *mq,
                             double min_certainty, double prior, const int64_t *gl_off,
                             int threads, double *gl);

/* partial_anchor.py:119-127 near-tie guard with the allele-frequency prior of
 * partial_anchor.py:56-57 (af per haplotype column) or, with af NULL, the
 * uniform haplotype prior of whole_read.py:51.  Returns the flagged count. */
int64_t vdo_guard(int64_t nregions, const int64_t *row_off, const int32_t *ncol, const int64_t *col_off,
                  const int64_t *gl_off, const double *gl, const double *cert, const double *af,
                  double min_certainty, double margin_scale, signed char *flags, double *gaps);

#ifdef __cplusplus
}
#endif
#endif

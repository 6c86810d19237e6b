/*
 * CPU FP64 oracle for the PairHMM read x haplotype likelihood.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU algorithm, used as the checker for the HIP engine (tests, smoke test and
 * the CPU-baseline leg of the benchmark).  Nothing in the product library
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  All arithmetic
 * is IEEE double with contraction disabled (-ffp-contract=off), in the same
 * operation order as the reference, so that with gap=.001, ext=.1 and the
 * reference error table the results equal the reference's bit for bit
 * (checked against golden vectors made from the reference compiled from
 * source).
 */
#ifndef VD_ORACLE_H
#define VD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* e[q] = 10^(-min(q,60)/10), q in [0,255]
 * (caller/native/caller/batch_hmm.cpp:14, trie_vector.cpp:13, accelerator.py:23). */
void vdo_error_table(double *out256);

/* One pair of the dense row-major recurrence (batch_hmm.cpp:31-70 for one
 * SIMD lane).  mode 0: semiglobal, divide every row by its max;
 * mode 1: anchored/global, divide every row by its max;
 * mode 2: semiglobal, multiply by 2^512 when the row max < 2^-450.
 * strict=0: a row max <= 0 is replaced by 1 (batch_hmm.cpp:175);
 * strict=1: a row max <= 0 returns -inf (base/core/core.cpp:24, wide.cpp:24,
 * anchored.cpp:20).  err256 is indexed by the raw quality byte.
 * n<=0 or m<=0 returns -inf. */
double vdo_dense(const char *read, const unsigned char *qual, int n,
                 const char *hap, int m, double gap, double ext,
                 const double *err256, int mode, int strict);

/* Array form of batch_hmm.cpp:192 pairhmm_pairs with explicit parameters.
 * Returns 1 if count<0 or mode not in {0,1,2}, else 0. */
int vdo_pairhmm_pairs(const char *const *reads, const unsigned char *const *quals,
                      const int *rn, const char *const *haps, const int *hn,
                      int count, int mode, double gap, double ext,
                      const double *err256, double *output);

/* trie_vector.cpp:284 trie_many restated per read (the SIMD lanes of
 * compute_8 are independent), including the shared-suffix backward
 * continuation (trie_vector.cpp:198-257) and the cells counter for a
 * lane width of 8/4/2 chosen as the reference chooses it on an AVX-512
 * host.  out is row-major nr x nh: ln likelihood, NaN when the likelihood
 * is <= 1e-200, -inf for an empty read or haplotype. */
int vdo_trie_many(const char *const *reads, const unsigned char *const *quals,
                  const int *rn, int nr, const char *const *haps, const int *hn,
                  int nh, int mode, int lanes, double gap, double ext,
                  const double *err256, double *out, uint64_t *cells);

/* Region batch on the CPU: for every region, trie_many over its reads (sorted
 * by length) and its haplotypes, then the dense mode for NaN entries --
 * the reference CPU production chain trie_vector.py:16-42 + cpu_batch.py:44-62.
 * Regions are handed to `threads` workers largest-first.  Output layout:
 * region k writes a row-major R_k x H_k block at out + out_off[k]. */
int vdo_regions_cpu(int nregions, const int64_t *read_off, const int64_t *hap_off,
                    const int64_t *out_off, const int64_t *read_base_off,
                    const int64_t *hap_base_off, const char *read_bytes,
                    const unsigned char *qual_bytes, const char *hap_bytes,
                    int mode, double gap, double ext, const double *err256,
                    int threads, double *out, uint64_t *dense_pairs);

#ifdef __cplusplus
}
#endif
#endif

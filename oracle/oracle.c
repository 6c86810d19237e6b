/*
 * CPU FP64 oracle for the PairHMM read x haplotype likelihood.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU
 * baseline of the benchmark.  Not part of the product library.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).  The operation
 * order of every floating-point expression follows the cited reference
 * lines exactly; do not "simplify" it.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

void vdo_error_table(double *out256)
{
    /* batch_hmm.cpp:14: pow(10., -min(i,60)/10.) */
    for (int i = 0; i < 256; ++i) {
        int q = i < 60 ? i : 60;
        out256[i] = pow(10., -q / 10.);
    }
}

/* batch_hmm.cpp:31-70 (one lane); strict variants: core.cpp:9-30,
 * wide.cpp:9-33, anchored.cpp:5-25. */
double vdo_dense(const char *read, const unsigned char *qual, int n,
                 const char *hap, int m, double gap, double ext,
                 const double *err256, int mode, int strict)
{
    if (n <= 0 || m <= 0) return -INFINITY;
    size_t w = (size_t)m + 1;
    double *buf = (double *)malloc(sizeof(double) * 6 * w);
    double *pm = buf, *pi = buf + w, *pd = buf + 2 * w;
    double *cm = buf + 3 * w, *ci = buf + 4 * w, *cd = buf + 5 * w;
    const double quarter = .25, g = gap, x = ext;
    const double match = 1 - 2 * gap, exit_gap = 1 - ext;
    /* batch_hmm.cpp:31-39 */
    for (int j = 0; j <= m; ++j) {
        double deletion = j ? gap * pow(ext, j - 1) : 0.;
        pm[j] = mode == 1 ? (j == 0 ? 1. : 0.) : 1. / (m + 1);
        pi[j] = 0.;
        pd[j] = mode == 1 ? deletion : 0.;
    }
    double logs = 0.;
    for (int i = 1; i <= n; ++i) {
        double e = err256[qual[i - 1]];
        /* batch_hmm.cpp:44-45 */
        cm[0] = 0.;
        cd[0] = 0.;
        ci[0] = quarter * (g * pm[0] + x * pi[0]);
        double scale = cm[0] + ci[0] + cd[0];
        double good = 1. - e, bad = e / 3.;
        unsigned char b = (unsigned char)read[i - 1];
        /* batch_hmm.cpp:46-52 */
        for (int j = 1; j <= m; ++j) {
            double emit = b == (unsigned char)hap[j - 1] ? good : bad;
            cm[j] = emit * (match * pm[j - 1] + exit_gap * (pi[j - 1] + pd[j - 1]));
            ci[j] = quarter * (g * pm[j] + x * pi[j]);
            cd[j] = g * cm[j - 1] + x * cd[j - 1];
            double sum = cm[j] + ci[j] + cd[j];
            if (sum > scale) scale = sum;
        }
        if (strict && !(scale > 0)) { free(buf); return -INFINITY; }
        if (mode == 2) {
            /* batch_hmm.cpp:53-56 */
            if (scale < 0x1p-450) {
                for (int j = 0; j <= m; ++j) { cm[j] *= 0x1p512; ci[j] *= 0x1p512; cd[j] *= 0x1p512; }
                logs -= 512 * log(2.);
            }
        } else {
            /* batch_hmm.cpp:57-60 */
            if (!(scale > 0)) scale = 1.;
            logs += log(scale);
            for (int j = 0; j <= m; ++j) { cm[j] /= scale; ci[j] /= scale; cd[j] /= scale; }
        }
        double *t;
        t = pm; pm = cm; cm = t;
        t = pi; pi = ci; ci = t;
        t = pd; pd = cd; cd = t;
    }
    /* batch_hmm.cpp:64-67 */
    double v;
    if (mode != 1) {
        double sum = 0.;
        for (int j = 0; j <= m; ++j) sum += pm[j] + pi[j];
        v = sum;
    } else {
        v = pm[m] + pi[m] + pd[m];
    }
    double r = log(1e-300 < v ? v : 1e-300) + logs;
    free(buf);
    return r;
}

int vdo_pairhmm_pairs(const char *const *reads, const unsigned char *const *quals,
                      const int *rn, const char *const *haps, const int *hn,
                      int count, int mode, double gap, double ext,
                      const double *err256, double *output)
{
    /* batch_hmm.cpp:192-200 */
    if (count < 0 || mode < 0 || mode > 2) return 1;
    for (int p = 0; p < count; ++p)
        output[p] = vdo_dense(reads[p], quals[p], rn[p], haps[p], hn[p], gap, ext,
                              err256, mode, 0);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* trie_vector.cpp:194-280, one read (lane) at a time.                      */

typedef struct {
    int nh, suffix, depth;
    int *shared;
} TriePlan;

static void trie_plan(TriePlan *tp, const char *const *haps, const int *hn, int nh)
{
    /* trie_vector.cpp:198-205 */
    int suffix = 0;
    if (nh > 1) {
        int minimum = hn[0];
        for (int h = 1; h < nh; ++h) if (hn[h] < minimum) minimum = hn[h];
        while (suffix < minimum) {
            char c = haps[0][hn[0] - suffix - 1];
            int same = 1;
            for (int h = 1; h < nh; ++h)
                if (haps[h][hn[h] - suffix - 1] != c) { same = 0; break; }
            if (!same) break;
            ++suffix;
        }
        if (suffix < 16) suffix = 0;
    }
    int depth = 0;
    int *shared = (int *)calloc(nh > 0 ? nh : 1, sizeof(int));
    for (int h = 0; h < nh; ++h) {
        if (hn[h] - suffix > depth) depth = hn[h] - suffix;
        if (h) {
            int lim = (hn[h] < hn[h - 1] ? hn[h] : hn[h - 1]) - suffix;
            while (shared[h] < lim && haps[h][shared[h]] == haps[h - 1][shared[h]]) ++shared[h];
        }
    }
    tp->nh = nh; tp->suffix = suffix; tp->depth = depth; tp->shared = shared;
}

static void trie_read(const char *read, const unsigned char *qual, int n,
                      const char *const *haps, const int *hn, const TriePlan *tp,
                      int mode, double gap, double ext, const double *err256,
                      double *out_row)
{
    const int nh = tp->nh, suffix = tp->suffix;
    const double zero = 0., one = 1., quarter = .25, g = gap, x = ext;
    const double match = 1 - 2 * gap, exitg = 1 - ext;
    size_t stride = (size_t)n + 1;
    double *err = (double *)malloc(sizeof(double) * stride);
    double *mis = (double *)malloc(sizeof(double) * stride);
    unsigned *base = (unsigned *)calloc(stride, sizeof(unsigned));
    /* column 0: trie_vector.cpp:220-226 */
    double *col0 = (double *)calloc(3 * stride, sizeof(double));
    col0[0] = one;
    for (int i = 1; i <= n; ++i) {
        col0[stride + i] = quarter * (g * col0[i - 1] + x * col0[stride + i - 1]);
        double e = err256[qual[i - 1]];
        base[i] = (unsigned char)read[i - 1];
        mis[i] = e / 3.;
        err[i] = one - e;
    }
    double sum0 = col0[n] + col0[stride + n];
    double free_starts = zero;
    double *coef = NULL;
    if (suffix) {
        /* trie_vector.cpp:230-257 */
        coef = (double *)calloc(3 * stride, sizeof(double));
        double *right = (double *)calloc(3 * stride, sizeof(double));
        double *current = (double *)calloc(3 * stride, sizeof(double));
        for (int j = suffix - 1; j >= 0; --j) {
            unsigned hc = 0;
            if (j + 1 < suffix) hc = (unsigned char)haps[0][hn[0] - suffix + j + 1];
            for (int i = n; i >= 0; --i) {
                int terminal = (mode != 1 || j == suffix - 1) ? (n == i) : 0;
                int active = mode == 1 ? (n >= i) : (n > i);
                double diagonal = zero, insert = zero;
                if (i < n) {
                    double emit = base[i + 1] == hc ? err[i + 1] : mis[i + 1];
                    if (j + 1 < suffix) diagonal = emit * right[i + 1];
                    insert = quarter * current[stride + i + 1];
                }
                double deletion = (i || mode == 1) ? right[2 * stride + i] : zero;
                current[i] = terminal ? one
                           : (active ? match * diagonal + g * insert + g * deletion : zero);
                current[stride + i] = terminal ? one
                           : (active ? exitg * diagonal + x * insert : zero);
                if (mode == 1)
                    current[2 * stride + i] = terminal ? one
                        : (active ? exitg * diagonal + x * right[2 * stride + i] : zero);
                else
                    current[2 * stride + i] =
                        active ? exitg * diagonal + x * right[2 * stride + i] : zero;
            }
            if (mode != 1) free_starts += current[0];
            double *t = current; current = right; right = t;
        }
        unsigned hc = (unsigned char)haps[0][hn[0] - suffix];
        for (int i = 0; i <= n; ++i) {
            int active = mode == 1 ? (n >= i) : (n > i);
            double diagonal = zero;
            if (i < n) {
                double emit = base[i + 1] == hc ? err[i + 1] : mis[i + 1];
                diagonal = emit * right[i + 1];
            }
            coef[i] = active ? match * diagonal + ((i || mode == 1) ? g * right[2 * stride + i] : zero)
                             : zero;
            coef[stride + i] = active ? exitg * diagonal : zero;
            coef[2 * stride + i] = active ? exitg * diagonal + x * right[2 * stride + i] : zero;
        }
        free(right);
        free(current);
    }
    /* trie_vector.cpp:259-278, every haplotype recomputed from column 0:
     * a shared prefix column has identical values, so this equals the
     * trie's reuse. */
    double *left = (double *)malloc(sizeof(double) * 3 * stride);
    double *now = (double *)malloc(sizeof(double) * 3 * stride);
    for (int h = 0; h < nh; ++h) {
        int length = hn[h] - suffix;
        memcpy(left, col0, sizeof(double) * 3 * stride);
        double sum = sum0;
        for (int j = 1; j <= length; ++j) {
            double *lm = left, *li = left + stride, *ld = left + 2 * stride;
            double *m = now, *ins = now + stride, *del = now + 2 * stride;
            m[0] = mode == 1 ? zero : one;
            ins[0] = zero;
            del[0] = zero + (mode == 1 ? g * pow(ext, j - 1) : 0.);
            unsigned hapcode = (unsigned char)haps[h][j - 1];
            for (int i = 1; i <= n; ++i) {
                double emit = base[i] == hapcode ? err[i] : mis[i];
                m[i] = emit * (match * lm[i - 1] + exitg * (li[i - 1] + ld[i - 1]));
                ins[i] = quarter * (g * m[i - 1] + x * ins[i - 1]);
                del[i] = g * lm[i] + x * ld[i];
            }
            if (mode != 1) sum = sum + (m[n] + ins[n]);
            double *t = left; left = now; now = t;
        }
        double *last = left;
        double probability = mode == 1 ? zero : sum + free_starts;
        if (suffix)
            for (int i = 0; i <= n; ++i)
                probability += last[i] * coef[i] + last[stride + i] * coef[stride + i]
                             + last[2 * stride + i] * coef[2 * stride + i];
        double value;
        if (mode == 1)
            value = suffix ? probability : last[n] + last[stride + n] + last[2 * stride + n];
        else
            value = probability / (hn[h] + 1);
        out_row[h] = (n == 0 || hn[h] == 0) ? -INFINITY : (value > 1e-200 ? log(value) : NAN);
    }
    free(left); free(now); free(col0); free(err); free(mis); free(base); free(coef);
}

int vdo_trie_many(const char *const *reads, const unsigned char *const *quals,
                  const int *rn, int nr, const char *const *haps, const int *hn,
                  int nh, int mode, int lanes, double gap, double ext,
                  const double *err256, double *out, uint64_t *cells)
{
    /* trie_vector.cpp:284-292 */
    if (nr < 0 || nh < 0 || mode < 0 || mode > 2) return 1;
    *cells = 0;
    TriePlan tp;
    trie_plan(&tp, haps, hn, nh);
    /* cells counter (trie_vector.cpp:271): per SIMD group, group-max n times
     * group size per evaluated column; lane width as dispatched at :288-291
     * on an AVX-512 host. */
    int L = (lanes == 0 || lanes == 8) ? 8 : (lanes == 4 ? 4 : 2);
    uint64_t columns = 0;
    for (int h = 0; h < nh; ++h) {
        int length = hn[h] - tp.suffix;
        if (length > tp.shared[h]) columns += (uint64_t)(length - tp.shared[h]);
    }
    for (int first = 0; first < nr; first += L) {
        int take = nr - first < L ? nr - first : L, n = 0;
        for (int a = 0; a < take; ++a) if (rn[first + a] > n) n = rn[first + a];
        *cells += (uint64_t)n * (uint64_t)take * columns;
    }
    for (int a = 0; a < nr; ++a)
        trie_read(reads[a], quals[a], rn[a], haps, hn, &tp, mode, gap, ext, err256,
                  out + (size_t)a * nh);
    free(tp.shared);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Region batch on host threads (trie + dense fallback).                    */

typedef struct {
    int nregions;
    const int64_t *read_off, *hap_off, *out_off, *read_base_off, *hap_base_off;
    const char *read_bytes;
    const unsigned char *qual_bytes;
    const char *hap_bytes;
    int mode;
    double gap, ext;
    const double *err;
    double *out;
    int *order;
    int next;
    uint64_t dense;
    pthread_mutex_t lock;
} RegionJob;

static int hap_cmp_ctx_len(const char *a, int la, const char *b, int lb)
{
    int l = la < lb ? la : lb;
    int c = memcmp(a, b, (size_t)l);
    if (c) return c;
    return la < lb ? -1 : (la > lb ? 1 : 0);
}

static void region_one(RegionJob *J, int k)
{
    int64_t r0 = J->read_off[k], r1 = J->read_off[k + 1];
    int64_t h0 = J->hap_off[k], h1 = J->hap_off[k + 1];
    int R = (int)(r1 - r0), H = (int)(h1 - h0);
    if (R <= 0 || H <= 0) return;
    /* trie_vector.py:27,30: haplotypes sorted, reads sorted by length */
    int *hord = (int *)malloc(sizeof(int) * H);
    int *rord = (int *)malloc(sizeof(int) * R);
    for (int i = 0; i < H; ++i) hord[i] = i;
    for (int i = 0; i < R; ++i) rord[i] = i;
    for (int i = 1; i < H; ++i) {  /* insertion sort: H is small */
        int v = hord[i], j = i - 1;
        const char *bv = J->hap_bytes + J->hap_base_off[h0 + v];
        int lv = (int)(J->hap_base_off[h0 + v + 1] - J->hap_base_off[h0 + v]);
        while (j >= 0) {
            int u = hord[j];
            const char *bu = J->hap_bytes + J->hap_base_off[h0 + u];
            int lu = (int)(J->hap_base_off[h0 + u + 1] - J->hap_base_off[h0 + u]);
            if (hap_cmp_ctx_len(bu, lu, bv, lv) <= 0) break;
            hord[j + 1] = u; --j;
        }
        hord[j + 1] = v;
    }
    for (int i = 1; i < R; ++i) {  /* stable insertion sort by length */
        int v = rord[i], j = i - 1;
        int64_t lv = J->read_base_off[r0 + v + 1] - J->read_base_off[r0 + v];
        while (j >= 0) {
            int u = rord[j];
            int64_t lu = J->read_base_off[r0 + u + 1] - J->read_base_off[r0 + u];
            if (lu <= lv) break;
            rord[j + 1] = u; --j;
        }
        rord[j + 1] = v;
    }
    const char **rp = (const char **)malloc(sizeof(char *) * R);
    const unsigned char **qp = (const unsigned char **)malloc(sizeof(char *) * R);
    int *rn = (int *)malloc(sizeof(int) * R);
    const char **hp = (const char **)malloc(sizeof(char *) * H);
    int *hn = (int *)malloc(sizeof(int) * H);
    for (int i = 0; i < R; ++i) {
        int64_t r = r0 + rord[i];
        rp[i] = J->read_bytes + J->read_base_off[r];
        qp[i] = J->qual_bytes + J->read_base_off[r];
        rn[i] = (int)(J->read_base_off[r + 1] - J->read_base_off[r]);
    }
    for (int i = 0; i < H; ++i) {
        int64_t h = h0 + hord[i];
        hp[i] = J->hap_bytes + J->hap_base_off[h];
        hn[i] = (int)(J->hap_base_off[h + 1] - J->hap_base_off[h]);
    }
    double *tmp = (double *)malloc(sizeof(double) * (size_t)R * H);
    uint64_t cells = 0, dense = 0;
    vdo_trie_many(rp, qp, rn, R, hp, hn, H, J->mode, 0, J->gap, J->ext, J->err, tmp, &cells);
    double *o = J->out + J->out_off[k];
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < H; ++j) {
            double v = tmp[(size_t)i * H + j];
            if (isnan(v)) {  /* trie_vector.py:40-41 -> cpu_batch.py:44 */
                v = vdo_dense(rp[i], qp[i], rn[i], hp[j], hn[j], J->gap, J->ext, J->err,
                              J->mode, 0);
                ++dense;
            }
            o[(size_t)rord[i] * H + hord[j]] = v;
        }
    pthread_mutex_lock(&J->lock);
    J->dense += dense;
    pthread_mutex_unlock(&J->lock);
    free(tmp); free(rp); free(qp); free(rn); free(hp); free(hn); free(hord); free(rord);
}

static void *region_worker(void *arg)
{
    RegionJob *J = (RegionJob *)arg;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int i = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (i >= J->nregions) break;
        region_one(J, J->order[i]);
    }
    return NULL;
}

static const int64_t *g_cost;
static int cost_desc(const void *a, const void *b)
{
    int64_t ca = g_cost[*(const int *)a], cb = g_cost[*(const int *)b];
    return ca > cb ? -1 : (ca < cb ? 1 : (*(const int *)a - *(const int *)b));
}

int vdo_regions_cpu(int nregions, const int64_t *read_off, const int64_t *hap_off,
                    const int64_t *out_off, const int64_t *read_base_off,
                    const int64_t *hap_base_off, const char *read_bytes,
                    const unsigned char *qual_bytes, const char *hap_bytes,
                    int mode, double gap, double ext, const double *err256,
                    int threads, double *out, uint64_t *dense_pairs)
{
    if (nregions < 0 || mode < 0 || mode > 2) return 1;
    RegionJob J;
    memset(&J, 0, sizeof(J));
    J.nregions = nregions; J.read_off = read_off; J.hap_off = hap_off; J.out_off = out_off;
    J.read_base_off = read_base_off; J.hap_base_off = hap_base_off;
    J.read_bytes = read_bytes; J.qual_bytes = qual_bytes; J.hap_bytes = hap_bytes;
    J.mode = mode; J.gap = gap; J.ext = ext; J.err = err256; J.out = out;
    pthread_mutex_init(&J.lock, NULL);
    /* largest-first hand-out (evidence_workers.py:64-79 orders by locality;
     * largest-first bounds the tail on a fixed pool) */
    int64_t *cost = (int64_t *)malloc(sizeof(int64_t) * (nregions > 0 ? nregions : 1));
    J.order = (int *)malloc(sizeof(int) * (nregions > 0 ? nregions : 1));
    for (int k = 0; k < nregions; ++k) {
        int64_t rb = read_base_off[read_off[k + 1]] - read_base_off[read_off[k]];
        int64_t hb = hap_base_off[hap_off[k + 1]] - hap_base_off[hap_off[k]];
        cost[k] = rb * hb;
        J.order[k] = k;
    }
    g_cost = cost;
    qsort(J.order, (size_t)nregions, sizeof(int), cost_desc);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, region_worker, &J);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    if (dense_pairs) *dense_pairs = J.dense;
    pthread_mutex_destroy(&J.lock);
    free(th); free(cost); free(J.order);
    return 0;
}

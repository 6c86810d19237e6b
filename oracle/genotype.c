/*
 * CPU FP64 oracle for the diploid genotype likelihood (see genotype.h).
 *
 * TEST INFRASTRUCTURE ONLY.  Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off); keep the operation order of every expression.
 */
#include "genotype.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* triangular.py / artifact.py:12-15 (BALANCE, BALANCE_WEIGHT, CONTAMINATION,
 * CONTAMINATION_WEIGHT); mixture_general.cpp:10-11, mixture_two.cpp:10-11. */
static const double BAL[7] = {.2, .3, .4, .5, .6, .7, .8};
static const double BW[7] = {.025, .1, .225, .3, .225, .1, .025};
static const double THETA[4] = {.02, .1, .2, .3};
static const double TW[4] = {.4, .35, .2, .05};

/* mixture_general.cpp:8, mixture_two.cpp:8 */
static double log_add(double x, double y)
{
    if (x == y) return x + log(2.);
    return (x > y ? x : y) + log1p(exp(-fabs(x - y)));
}

/* std::max(1e-300, v): (1e-300 < v) ? v : 1e-300 (NaN -> 1e-300) */
static double floor300(double v) { return (1e-300 < v) ? v : 1e-300; }

/* np.maximum(1e-300, v): NaN propagates */
static double np_floor300(double v) { return isnan(v) ? v : (v > 1e-300 ? v : 1e-300); }

/* numpy's pairwise summation of a contiguous float64 run (its add reduction
 * inner loop, block 128, 8 partial sums); only the rounding of the row mean
 * depends on it. */
static double pairwise_sum(const double *a, int n)
{
    if (n < 8) {
        double res = -0.0;
        for (int i = 0; i < n; ++i) res += a[i];
        return res;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        int i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    }
    int n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

/* np.add.reduce over a row: the first element, then the rest */
static double row_sum(const double *a, int n)
{
    if (n <= 0) return 0.;
    return a[0] + pairwise_sum(a + 1, n - 1);
}

/* numpy max/min reductions propagate NaN */
static double np_max(const double *a, int n)
{
    double m = a[0];
    for (int i = 1; i < n; ++i) {
        if (isnan(m)) break;
        if (isnan(a[i]) || a[i] > m) m = a[i];
    }
    return m;
}

static double np_min(const double *a, int n)
{
    double m = a[0];
    for (int i = 1; i < n; ++i) {
        if (isnan(m)) break;
        if (isnan(a[i]) || a[i] < m) m = a[i];
    }
    return m;
}

/* molecule_rows.py:4 _MQ_ERROR[mq] = 10**(-min(mq,60)/10) */
static double mq_error(int mq)
{
    int q = mq < 60 ? mq : 60;
    return pow(10., -q / 10.);
}

/* molecule_rows.py:6-14 for one row; tmp holds n doubles */
static void normalized_row(int n, const double *ll, int mq, double *v, double *tmp, double *cert)
{
    double mx = np_max(ll, n);
    for (int h = 0; h < n; ++h) v[h] = exp(ll[h] - mx);          /* :10 */
    double mean = row_sum(v, n) / n;                               /* :12 values.mean */
    double e = mq_error(mq);                                       /* :11 */
    for (int h = 0; h < n; ++h) v[h] = (1 - e) * v[h] + e * mean;  /* :12 */
    for (int h = 0; h < n; ++h) tmp[h] = log(np_floor300(v[h]));  /* :13 */
    *cert = np_max(tmp, n) - np_min(tmp, n);                       /* np.ptp */
}

void vdo_normalized_rows(int rows, int n, const double *ll, const unsigned char *mq,
                         double *values, double *certainty)
{
    if (rows <= 0 || n <= 0) return;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
    for (int r = 0; r < rows; ++r)
        normalized_row(n, ll + (size_t)r * n, mq[r], values + (size_t)r * n, tmp, certainty + r);
    free(tmp);
}

/* mixture_general.cpp:9-40.  The reference caches the logged values of
 * repeated rows (:17-32); a cached row contributes the same values as a
 * recomputed one, so the restatement recomputes every row. */
void vdo_mixture_general(int rows, int n, const double *lh, double prior, int triangular,
                         double *out)
{
    int pairs = triangular ? n * (n + 1) / 2 : n * n;
    int *aa = (int *)malloc(sizeof(int) * (size_t)(pairs > 0 ? pairs : 1));
    int *bb = (int *)malloc(sizeof(int) * (size_t)(pairs > 0 ? pairs : 1));
    int g = 0;
    if (triangular) {                                   /* :13 */
        for (int b = 0; b < n; ++b)
            for (int a = 0; a <= b; ++a) { aa[g] = a; bb[g] = b; ++g; }
    } else {                                            /* :14 */
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) { aa[g] = a; bb[g] = b; ++g; }
    }
    int mixes = prior < 0 ? 1 : 7 + (prior > 0 ? 4 * n : 0);  /* :15 */
    size_t total = (size_t)pairs * mixes;
    double *totals = (double *)calloc(total > 0 ? total : 1, sizeof(double));
    double *scratch = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
    for (int row = 0; row < rows; ++row) {             /* :19-34 */
        const double *v = lh + (size_t)row * n;
        for (g = 0; g < pairs; ++g) {
            double a = v[aa[g]], b = v[bb[g]];
            double *dest = scratch + (size_t)g * mixes;
            if (prior < 0) {
                dest[0] = floor300((a + b) / 2);
            } else {
                for (int j = 0; j < 7; ++j) dest[j] = floor300(BAL[j] * a + (1 - BAL[j]) * b);
                if (prior > 0)
                    for (int j = 0; j < 4; ++j)
                        for (int c = 0; c < n; ++c)
                            dest[7 + j * n + c] = floor300((1 - THETA[j]) * (a + b) / 2 + THETA[j] * v[c]);
            }
        }
        for (size_t j = 0; j < total; ++j) scratch[j] = log(scratch[j]);  /* vector_log */
        for (size_t j = 0; j < total; ++j) totals[j] += scratch[j];
    }
    for (g = 0; g < pairs; ++g) {                      /* :35-40 */
        const double *v = totals + (size_t)g * mixes;
        if (prior < 0) { out[g] = v[0]; continue; }
        double clean = v[0] + log(BW[0]);
        for (int j = 1; j < 7; ++j) clean = log_add(clean, v[j] + log(BW[j]));
        if (prior > 0) {
            double art = v[7] + log(TW[0] / n);
            for (int j = 1; j < 4 * n; ++j) art = log_add(art, v[7 + j] + log(TW[j / n] / n));
            out[g] = log_add(log1p(-prior) + clean, log(prior) + art);
        } else {
            out[g] = clean;
        }
    }
    free(aa); free(bb); free(totals); free(scratch);
}

/* mixture_two.cpp:9-29 (its row cache, :13-22, is result-neutral too) */
void vdo_mixture_two(int rows, const double *lh, double prior, double *out4)
{
    double totals[4][15];
    memset(totals, 0, sizeof(totals));
    for (int row = 0; row < rows; ++row) {
        double v[2] = {lh[2 * row], lh[2 * row + 1]};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                int g = a * 2 + b;
                for (int j = 0; j < 7; ++j)
                    totals[g][j] += log(floor300(BAL[j] * v[a] + (1 - BAL[j]) * v[b]));
                for (int j = 0; j < 4; ++j)
                    for (int c = 0; c < 2; ++c)
                        totals[g][7 + 2 * j + c] +=
                            log(floor300((1 - THETA[j]) * (v[a] + v[b]) / 2 + THETA[j] * v[c]));
            }
    }
    for (int g = 0; g < 4; ++g) {
        double clean = totals[g][0] + log(BW[0]);
        for (int j = 1; j < 7; ++j) clean = log_add(clean, totals[g][j] + log(BW[j]));
        if (prior > 0) {
            double art = totals[g][7] + log(TW[0] / 2);
            for (int j = 1; j < 8; ++j) art = log_add(art, totals[g][7 + j] + log(TW[j / 2] / 2));
            out4[g] = log_add(log1p(-prior) + clean, log(prior) + art);
        } else {
            out4[g] = clean;
        }
    }
}

/* mixture_cpu.py:14,25: finite, non-negative entries */
static int valid_matrix(const double *m, size_t count)
{
    for (size_t i = 0; i < count; ++i)
        if (!isfinite(m[i]) || m[i] < 0) return 0;
    return 1;
}

/* triangular.py:6-11 */
static void diploid(int rows, int n, const double *lh, double prior, double *out)
{
    if (n == 2 && prior >= 0) {
        double o4[4];
        vdo_mixture_two(rows, lh, prior, o4);
        out[0] = o4[0]; out[1] = o4[1]; out[2] = o4[3];   /* ravel()[[0,1,3]] */
    } else {
        vdo_mixture_general(rows, n, lh, prior, 1, out);
    }
}

typedef struct {
    int64_t nregions;
    const int64_t *row_off, *mat_off, *gl_off;
    const int32_t *ncol;
    const double *ll;
    const unsigned char *mq;
    double min_certainty, prior;
    double *gl;
    int *order;
    int64_t next, invalid;
    pthread_mutex_t lock;
} GenoJob;

static void geno_one(GenoJob *J, int64_t k)
{
    int n = J->ncol[k];
    int64_t r0 = J->row_off[k], rows = J->row_off[k + 1] - r0;
    int pairs = n * (n + 1) / 2;
    double *out = J->gl + J->gl_off[k];
    if (n <= 0) return;
    double *mat = (double *)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1) * n);
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
    int kept = 0;
    for (int64_t r = 0; r < rows; ++r) {
        const double *row = J->ll + J->mat_off[k] + r * n;
        double cert;
        normalized_row(n, row, J->mq[r0 + r], v, tmp, &cert);
        if (cert < J->min_certainty) continue;            /* whole_read.py:37 */
        /* whole_read.py:45-46 for a single accepted alignment:
         * logl = 0 + log(max(1e-300, v)); molecule = exp(logl - logl.max()) */
        for (int h = 0; h < n; ++h) tmp[h] = 0. + log(np_floor300(v[h]));
        double mx = np_max(tmp, n);
        for (int h = 0; h < n; ++h) mat[(size_t)kept * n + h] = exp(tmp[h] - mx);
        ++kept;
    }
    if (!valid_matrix(mat, (size_t)kept * n)) {
        for (int g = 0; g < pairs; ++g) out[g] = NAN;
        pthread_mutex_lock(&J->lock);
        J->invalid++;
        pthread_mutex_unlock(&J->lock);
    } else {
        diploid(kept, n, mat, J->prior, out);
    }
    free(mat); free(v); free(tmp);
}

static void *geno_worker(void *arg)
{
    GenoJob *J = (GenoJob *)arg;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int64_t i = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (i >= J->nregions) break;
        geno_one(J, J->order[i]);
    }
    return NULL;
}

static const int64_t *g_geno_cost;
static int geno_cost_desc(const void *a, const void *b)
{
    int64_t ca = g_geno_cost[*(const int *)a], cb = g_geno_cost[*(const int *)b];
    return ca < cb ? 1 : (ca > cb ? -1 : 0);
}

int64_t vdo_genotype_regions(int64_t nregions, const int64_t *row_off, const int32_t *ncol,
                             const int64_t *mat_off, const double *ll, const unsigned char *mq,
                             double min_certainty, double prior, const int64_t *gl_off,
                             int threads, double *gl)
{
    GenoJob J;
    memset(&J, 0, sizeof(J));
    J.nregions = nregions; J.row_off = row_off; J.ncol = ncol; J.mat_off = mat_off;
    J.gl_off = gl_off; J.ll = ll; J.mq = mq; J.min_certainty = min_certainty;
    J.prior = prior; J.gl = gl;
    pthread_mutex_init(&J.lock, NULL);
    size_t cnt = (size_t)(nregions > 0 ? nregions : 1);
    int64_t *cost = (int64_t *)malloc(sizeof(int64_t) * cnt);
    J.order = (int *)malloc(sizeof(int) * cnt);
    for (int64_t k = 0; k < nregions; ++k) {
        int64_t n = ncol[k];
        cost[k] = (row_off[k + 1] - row_off[k]) * n * n * (7 + 4 * n);
        J.order[k] = (int)k;
    }
    g_geno_cost = cost;
    qsort(J.order, (size_t)nregions, sizeof(int), geno_cost_desc);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, geno_worker, &J);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&J.lock);
    free(th); free(cost); free(J.order);
    return J.invalid;
}

/* Near-tie guard of partial_anchor.py:119-127 over a genotype batch.
 * Region k: n = ncol[k] haplotypes, pair outputs gl[gl_off[k] ..] in the
 * order b = 0..n-1, a = 0..b (triangular.py:14), rows row_off[k] ..
 * row_off[k+1] with certainty cert[row].  The haplotype prior is
 *   af (per global haplotype column col_off[k] + c, or NULL): a_c = max(.001, af),
 *   normalised; prior(a, b) = a_a a_b (1 if a == b else 2), normalised
 *   (partial_anchor.py:56-57); with af NULL the uniform haplotype frequencies
 *   1/n, i.e. prior 1/n^2 (hom) and 2/n^2 (het) (whole_read.py:51).
 * gaps[2k] = largest - second largest gl, gaps[2k+1] the same of gl + log(prior);
 * flags[k] = 1 when the smaller gap <= margin_scale * max(1, molecules), molecules
 * = rows whose certainty is not below min_certainty.  Regions with fewer than
 * two pairs get flag 0 and infinite gaps.  Sums run in index order. */
int64_t vdo_guard(int64_t nregions, const int64_t *row_off, const int32_t *ncol, const int64_t *col_off,
                  const int64_t *gl_off, const double *gl, const double *cert, const double *af,
                  double min_certainty, double margin_scale, signed char *flags, double *gaps)
{
    int64_t nflag = 0;
    double *w = NULL, *pr = NULL;
    size_t cap = 0;
    for (int64_t k = 0; k < nregions; ++k) {
        const int64_t n = ncol[k], pairs = n * (n + 1) / 2;
        flags[k] = 0;
        gaps[2 * k] = gaps[2 * k + 1] = INFINITY;
        if (pairs < 2) continue;
        if ((size_t)pairs > cap) {
            cap = (size_t)pairs;
            w = (double *)realloc(w, sizeof(double) * cap);
            pr = (double *)realloc(pr, sizeof(double) * cap);
        }
        if (af) {
            double S = 0;
            for (int64_t c = 0; c < n; ++c) S += fmax(.001, af[col_off[k] + c]);
            int64_t q = 0;
            double P = 0;
            for (int64_t b = 0; b < n; ++b)
                for (int64_t a = 0; a <= b; ++a, ++q) {
                    const double fa = fmax(.001, af[col_off[k] + a]) / S, fb = fmax(.001, af[col_off[k] + b]) / S;
                    pr[q] = fa * fb * (a == b ? 1.0 : 2.0);
                    P += pr[q];
                }
            for (q = 0; q < pairs; ++q) pr[q] = log(pr[q] / P);
        } else {
            int64_t q = 0;
            for (int64_t b = 0; b < n; ++b)
                for (int64_t a = 0; a <= b; ++a, ++q) pr[q] = log((a == b ? 1.0 : 2.0) / (double)(n * n));
        }
        double a1 = -INFINITY, a2 = -INFINITY, b1 = -INFINITY, b2 = -INFINITY;
        for (int64_t q = 0; q < pairs; ++q) {
            const double v = gl[gl_off[k] + q], x = v + pr[q];
            if (v > a1) { a2 = a1; a1 = v; } else if (v > a2) { a2 = v; }
            if (x > b1) { b2 = b1; b1 = x; } else if (x > b2) { b2 = x; }
        }
        int64_t mol = 0;
        for (int64_t r = row_off[k]; r < row_off[k + 1]; ++r) mol += !(cert[r] < min_certainty);
        gaps[2 * k] = a1 - a2;
        gaps[2 * k + 1] = b1 - b2;
        const double margin = margin_scale * (mol > 1 ? (double)mol : 1.0);
        const double lo = gaps[2 * k] < gaps[2 * k + 1] ? gaps[2 * k] : gaps[2 * k + 1];
        if (lo <= margin) {
            flags[k] = 1;
            ++nflag;
        }
    }
    free(w);
    free(pr);
    return nflag;
}

// Diploid genotype likelihood over haplotype pairs, on the device.
//
// Two kernels per region batch (regions = rows x haplotypes blocks laid out
// like the likelihood matrix of the forward):
//
// k_molecule   one wave per region, one lane per row: normalized rows with the
//              mapping-quality mixture and the certainty (molecule_rows.py:6-14),
//              the certainty filter and the single-alignment molecule row
//              exp(log(max(1e-300, v)) - max) (whole_read.py:36-46).
// k_mixture    one wave per 64 haplotype pairs of a region, one lane per pair:
//              the allele-balance and artifact mixtures of
//              mixture_general.cpp:9-40 / mixture_two.cpp:9-29 in the triangular
//              pair order of triangular.py:14 (or the square a-major order).
//
// Each mixture total sum_r log(max(1e-300, m_r)) is carried as an FP64
// product with its binary exponent split off after every row (frexp), and
// logged once at the end: log(p) + e ln2 with a two-part ln2.  Every row's
// mixture value m_r is formed with the reference's operation order and no
// contraction, so the per-row inputs are the reference's; the total then
// differs from the reference's sequential sum of logs only by that sum's own
// rounding (one log per row, ~1 ulp each), at a fifth of the arithmetic.
#include "vdhmm_internal.h"

namespace vd {

// artifact.py:12-15 (BALANCE, BALANCE_WEIGHT, CONTAMINATION, CONTAMINATION_WEIGHT)
__constant__ double G_BAL[7] = {.2, .3, .4, .5, .6, .7, .8};
__constant__ double G_BW[7] = {.025, .1, .225, .3, .225, .1, .025};
__constant__ double G_THETA[4] = {.02, .1, .2, .3};
__constant__ double G_TW[4] = {.4, .35, .2, .05};

// ln 2 = LN2_HI + LN2_LO; LN2_HI has 21 trailing zero bits, so e * LN2_HI is
// exact for |e| < 2^21 (Cody and Waite's argument split).
constexpr double LN2_HI = 0x1.62e42fee00000p-1;
constexpr double LN2_LO = 0x1.a39ef35793c76p-33;

__device__ __forceinline__ double log_add(double x, double y) {  // mixture_general.cpp:8
  if (x == y) return x + log(2.);
  return (x > y ? x : y) + log1p(exp(-fabs(x - y)));
}

__device__ __forceinline__ double floor300(double v) { return (1e-300 < v) ? v : 1e-300; }  // std::max(1e-300, v)
__device__ __forceinline__ double np_floor300(double v) { return (v != v) ? v : (v > 1e-300 ? v : 1e-300); }
__device__ __forceinline__ double np_max2(double m, double v) { return (m != m || v != v) ? (m + v) : (v > m ? v : m); }
__device__ __forceinline__ double np_min2(double m, double v) { return (m != m || v != v) ? (m + v) : (v < m ? v : m); }

struct LogProd {
  double p;
  int e;
  __device__ __forceinline__ void init() {
    p = 1.0;
    e = 0;
  }
  __device__ __forceinline__ void mul(double v) {
    const double t = p * v;  // finite and > 0: v >= 1e-300, p in [0.5, 1)
    p = __builtin_amdgcn_frexp_mant(t);
    e += __builtin_amdgcn_frexp_exp(t);
  }
  __device__ __forceinline__ double total() const {
    const double de = (double)e;
    return (de * LN2_HI + log(p)) + de * LN2_LO;
  }
};

// ---------------------------------------------------------------------------
// molecule rows

__global__ __launch_bounds__(256) void k_molecule(GenoDev g, const double* __restrict__ ll,
                                                  const uint8_t* __restrict__ mq,
                                                  const double* __restrict__ mq_err, double min_cert,
                                                  double* __restrict__ values, double* __restrict__ cert,
                                                  double* __restrict__ mol, int8_t* __restrict__ keep,
                                                  unsigned long long* __restrict__ ctr) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k = wave; k < g.nregions; k += nwaves) {
    const int64_t r0 = g.row_off[k], nr = g.row_off[k + 1] - r0;
    const int n = (int)(g.col_off[k + 1] - g.col_off[k]);
    const int64_t base = g.mat_off[k];
    if (r0 < 0 || nr < 0 || r0 + nr > g.nrows) {  // uniform; rows outside the row arrays
      if (lane == 0 && ctr) atomicAdd(&ctr[GCTR_LAYOUT], 1ull);
      continue;
    }
    if (base < 0 || base + nr * (n > 0 ? n : 0) > g.mat_size) {
      for (int64_t r = lane; r < nr; r += 64) keep[r0 + r] = 0;
      if (lane == 0 && ctr) atomicAdd(&ctr[GCTR_LAYOUT], 1ull);
      continue;
    }
    if (n <= 0) {
      for (int64_t r = lane; r < nr; r += 64) {
        if (cert) cert[r0 + r] = 0.0;  // numpy would raise on the empty ptp; no pair consumes the row
        keep[r0 + r] = 0;
      }
      continue;
    }
    int kept = 0;
    for (int64_t r = lane; r < nr; r += 64) {
      const double* L = ll + base + r * n;
      double mx = L[0];
      for (int h = 1; h < n; ++h) mx = np_max2(mx, L[h]);
      double s = 0.0;  // values.mean: sequential sum (numpy sums pairwise; differs in the last bits)
      for (int h = 0; h < n; ++h) s += exp(L[h] - mx);
      const double mean = s / n;
      const double e = mq_err[mq[r0 + r]];
      const double om = 1 - e;
      double lmax = 0, lmin = 0;
      for (int h = 0; h < n; ++h) {
        const double v = om * exp(L[h] - mx) + e * mean;
        if (values) values[base + r * n + h] = v;
        const double lv = log(np_floor300(v));
        lmax = h ? np_max2(lmax, lv) : lv;
        lmin = h ? np_min2(lmin, lv) : lv;
      }
      const double c = lmax - lmin;  // np.ptp(log(max(1e-300, values)))
      if (cert) cert[r0 + r] = c;
      const bool kp = !(c < min_cert);  // whole_read.py:37 (NaN kept)
      keep[r0 + r] = kp ? 1 : 0;
      kept += kp;
      // whole_read.py:45-46, one accepted alignment: logl = 0 + log(max(1e-300, v))
      double* O = mol + base + r * n;
      for (int h = 0; h < n; ++h) {
        const double v = om * exp(L[h] - mx) + e * mean;
        O[h] = exp((0.0 + log(np_floor300(v))) - lmax);
      }
    }
    if (ctr) {
      for (int o = 32; o; o >>= 1) kept += __shfl_xor(kept, o);
      if (lane == 0 && kept) atomicAdd(&ctr[GCTR_KEPT], (unsigned long long)kept);
    }
  }
}

// ---------------------------------------------------------------------------
// kept rows of every region, in row order (one wave per region): the mixture
// loops then visit only the rows the certainty filter kept, with no branch on
// a keep load per row (at 30x most rows are dropped)

__global__ __launch_bounds__(256) void k_keep_compact(GenoDev g, const int8_t* __restrict__ keep,
                                                      int32_t* __restrict__ kidx, int32_t* __restrict__ nkept) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k = wave; k < g.nregions; k += nwaves) {
    const int64_t r0 = g.row_off[k], nr = g.row_off[k + 1] - r0;
    if (r0 < 0 || nr < 0 || r0 + nr > g.nrows) {  // uniform; k_mixture reports the layout
      if (lane == 0) nkept[k] = 0;
      continue;
    }
    int base = 0;
    for (int64_t c = 0; c < nr; c += 64) {
      const int64_t r = c + lane;
      const bool kp = r < nr && keep[r0 + r];
      const unsigned long long m = __builtin_amdgcn_ballot_w64(kp);
      if (kp) kidx[r0 + base + __builtin_popcountll(m & ((1ull << lane) - 1))] = (int32_t)r;
      base += __builtin_popcountll(m);
    }
    if (lane == 0) nkept[k] = base;
  }
}

// ---------------------------------------------------------------------------
// wave units over pairs

__global__ void k_mix_count(GenoDev g, int triangular, int64_t* __restrict__ unit_cnt) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < g.nregions;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = g.col_off[k + 1] - g.col_off[k];
    const int64_t pairs = n <= 0 ? 0 : (triangular ? n * (n + 1) / 2 : n * n);
    unit_cnt[k] = (pairs + 63) / 64;
  }
}

// exclusive scan of a[0..n) in place, total into a[n] and ctr[GCTR_UNITS]
__global__ __launch_bounds__(1024) void k_mix_scan(int64_t* __restrict__ a, int64_t n,
                                                   unsigned long long* __restrict__ ctr) {
  __shared__ int64_t part[1024];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t chunk = (n + nt - 1) / nt;
  const int64_t lo = tid * chunk, hi = lo + chunk < n ? lo + chunk : n;
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += a[i];
  part[tid] = s;
  __syncthreads();
  for (int d = 1; d < nt; d <<= 1) {
    const int64_t v = tid >= d ? part[tid - d] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int64_t run = tid ? part[tid - 1] : 0;
  for (int64_t i = lo; i < hi; ++i) {
    const int64_t v = a[i];
    a[i] = run;
    run += v;
  }
  if (tid == nt - 1) {
    a[n] = part[nt - 1];
    ctr[GCTR_UNITS] = (unsigned long long)part[nt - 1];
  }
}

__global__ void k_mix_map(GenoDev g, const int64_t* __restrict__ unit_off, int32_t* __restrict__ unit_region) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < g.nregions;
       k += (int64_t)gridDim.x * blockDim.x)
    for (int64_t u = unit_off[k]; u < unit_off[k + 1] && u < g.unit_cap; ++u) unit_region[u] = (int32_t)k;
}

// (a, b) of pair index q: triangular.py:14 / mixture_general.cpp:13-14
struct PairAB {
  int a, b;
};

__device__ __forceinline__ PairAB pair_of(int64_t q, int n, int triangular) {
  if (!triangular) return PairAB{(int)(q / n), (int)(q % n)};
  int b = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  b -= (int64_t)b * (b + 1) / 2 > q;  // sqrt rounding: at most one step either way
  b += (int64_t)(b + 1) * (b + 2) / 2 <= q;
  return PairAB{(int)(q - (int64_t)b * (b + 1) / 2), b};
}

constexpr int MIXW = 8;  // artifact totals carried per pass over the rows

__global__ __launch_bounds__(256) void k_mixture(GenoDev g, const double* __restrict__ mat,
                                                 const int8_t* __restrict__ keep, double prior, int triangular,
                                                 const int64_t* __restrict__ unit_off,
                                                 const int32_t* __restrict__ unit_region,
                                                 const int32_t* __restrict__ kidx, const int32_t* __restrict__ nkept,
                                                 double* __restrict__ gl, unsigned long long* __restrict__ ctr) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t nunits = (int64_t)ctr[GCTR_UNITS] < g.unit_cap ? (int64_t)ctr[GCTR_UNITS] : g.unit_cap;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < nunits; u += nwaves) {
    const int k = __builtin_amdgcn_readfirstlane(unit_region[u]);
    const int64_t u0 = unit_off[k];
    const int n = (int)(g.col_off[k + 1] - g.col_off[k]);
    const int64_t r0 = g.row_off[k], nr = g.row_off[k + 1] - r0;
    const int64_t pairs = triangular ? (int64_t)n * (n + 1) / 2 : (int64_t)n * n;
    const int64_t q = (u - u0) * 64 + lane;
    const bool active = q < pairs;
    const PairAB ab = pair_of(active ? q : pairs - 1, n, triangular);  // idle lanes shadow the last pair
    const int a = ab.a, b = ab.b;
    const int64_t mo = g.mat_off[k], go = g.gl_off[k];
    if (r0 < 0 || nr < 0 || r0 + nr > g.nrows || mo < 0 || go < 0 || mo + nr * n > g.mat_size ||
        go + pairs > g.gl_size) {  // uniform
      if (u == u0 && lane == 0) atomicAdd(&ctr[GCTR_LAYOUT], 1ull);
      continue;
    }
    const double* M = mat + mo;
    double* out = gl + go;

    // mixture_cpu.py:14,25: every entry of the matrix finite and non-negative
    bool bad = false;
    for (int64_t i = lane; i < nr * n; i += 64) {
      const int64_t r = i / n;
      if (keep && !keep[r0 + r]) continue;
      const double v = M[i];
      bad |= !(v >= 0 && v <= 1.7976931348623157e308);
    }
    if (__builtin_amdgcn_ballot_w64(bad)) {
      if (active) out[q] = __builtin_nan("");
      if (u == u0 && lane == 0) atomicAdd(&ctr[GCTR_INVALID], 1ull);
      continue;
    }

    // the kept rows in row order: kidx[r0 .. r0 + nk) (all rows without a keep mask)
    const int64_t nk = keep ? nkept[k] : nr;
    auto row_at = [&](int64_t i) { return M + (keep ? (int64_t)kidx[r0 + i] : i) * n; };

    if (prior < 0) {  // the pair's mean, no mixture (mixture_general.cpp:23)
      LogProd t;
      t.init();
      for (int64_t i = 0; i < nk; ++i) {
        const double* row = row_at(i);
        t.mul(floor300((row[a] + row[b]) / 2));
      }
      if (active) out[q] = t.total();
      continue;
    }

    // allele balance (mixture_general.cpp:25, :37)
    LogProd tb[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) tb[j].init();
    for (int64_t i = 0; i < nk; ++i) {
      const double* row = row_at(i);
      const double la = row[a], lb = row[b];
#pragma unroll
      for (int j = 0; j < 7; ++j) tb[j].mul(floor300(G_BAL[j] * la + (1 - G_BAL[j]) * lb));
    }
    double clean = tb[0].total() + log(G_BW[0]);
#pragma unroll
    for (int j = 1; j < 7; ++j) clean = log_add(clean, tb[j].total() + log(G_BW[j]));
    if (prior == 0) {
      if (active) out[q] = clean;
      continue;
    }

    // artifact: index j = t * n + c in order (mixture_general.cpp:26, :38)
    double art = 0;
    bool first = true;
    for (int t = 0; t < 4; ++t) {
      const double th = G_THETA[t], om = 1 - G_THETA[t];
      const double lt = log(G_TW[t] / n);
      for (int c0 = 0; c0 < n; c0 += MIXW) {
        const int nc = n - c0 < MIXW ? n - c0 : MIXW;
        LogProd ta[MIXW];
#pragma unroll
        for (int i = 0; i < MIXW; ++i) ta[i].init();
        for (int64_t i = 0; i < nk; ++i) {
          const double* row = row_at(i);
          const double half = om * (row[a] + row[b]) / 2;
#pragma unroll
          for (int i = 0; i < MIXW; ++i) {
            const int c = c0 + i < n ? c0 + i : n - 1;  // in bounds; slots past n are not folded
            ta[i].mul(floor300(half + th * row[c]));
          }
        }
#pragma unroll
        for (int i = 0; i < MIXW; ++i) {  // static indices keep ta[] in registers
          if (i < nc) {                   // uniform
            const double x = ta[i].total() + lt;
            art = first ? x : log_add(art, x);
            first = false;
          }
        }
      }
    }
    if (active) out[q] = log_add(log1p(-prior) + clean, log(prior) + art);
  }
}

// Near-tie guard (partial_anchor.py:119-133): per region, the gaps between the
// two largest values of gl and of lp = gl + log(prior), prior = 1 (a == b) or 2
// (a != b) over n^2 (whole_read.py:51), against margin_scale * max(1, molecules),
// molecules = rows whose certainty is not below min_cert.  Regions with fewer
// than two pairs are never flagged.
// Haplotype prior of the guard: with allele frequencies af (per haplotype
// column) the one of partial_anchor.py:56-57 -- a_c = max(.001, af) normalised,
// prior(a, b) = a_a a_b (1 if a == b else 2) normalised -- else the uniform
// frequencies of whole_read.py:51 (1/n^2 hom, 2/n^2 het).  Sums run in index
// order.
__device__ __forceinline__ double guard_log_prior(const double* __restrict__ af, int64_t c0, int64_t n, int64_t a,
                                                  int64_t b, double S, double P) {
  if (!af) return log((a == b ? 1.0 : 2.0) / (double)(n * n));
  const double fa = fmax(.001, af[c0 + a]) / S, fb = fmax(.001, af[c0 + b]) / S;
  return log(fa * fb * (a == b ? 1.0 : 2.0) / P);
}

__global__ void k_guard(GenoDev g, const double* __restrict__ gl, const double* __restrict__ cert,
                        const double* __restrict__ af, double min_cert, double margin_scale,
                        int8_t* __restrict__ flag, double* __restrict__ gaps, unsigned long long* __restrict__ ctr) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < g.nregions;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = g.col_off[k + 1] - g.col_off[k];
    const int64_t pairs = n * (n + 1) / 2;
    const int64_t r0 = g.row_off[k], r1 = g.row_off[k + 1];
    const int64_t go = g.gl_off[k];
    flag[k] = 0;
    gaps[2 * k] = gaps[2 * k + 1] = __builtin_inf();
    if (pairs < 2 || go < 0 || go + pairs > g.gl_size || r0 < 0 || r1 > g.nrows) continue;
    int64_t mol = 0;
    for (int64_t r = r0; r < r1; ++r) mol += !(cert[r] < min_cert);
    const int64_t c0 = g.col_off[k];
    double S = 0.0, P = 0.0;
    if (af) {
      for (int64_t c = 0; c < n; ++c) S += fmax(.001, af[c0 + c]);
      for (int64_t bb = 0; bb < n; ++bb)
        for (int64_t aa = 0; aa <= bb; ++aa) {
          const double fa = fmax(.001, af[c0 + aa]) / S, fb = fmax(.001, af[c0 + bb]) / S;
          P += fa * fb * (aa == bb ? 1.0 : 2.0);
        }
    }
    double a1 = -__builtin_inf(), a2 = -__builtin_inf(), b1 = -__builtin_inf(), b2 = -__builtin_inf();
    int64_t q = 0;
    for (int64_t bb = 0; bb < n; ++bb)
      for (int64_t aa = 0; aa <= bb; ++aa, ++q) {
        const double v = gl[go + q];
        const double w = v + guard_log_prior(af, c0, n, aa, bb, S, P);
        if (v > a1) { a2 = a1; a1 = v; } else if (v > a2) { a2 = v; }
        if (w > b1) { b2 = b1; b1 = w; } else if (w > b2) { b2 = w; }
      }
    const double ga = a1 - a2, gb = b1 - b2;
    gaps[2 * k] = ga;
    gaps[2 * k + 1] = gb;
    const double margin = margin_scale * (mol > 1 ? (double)mol : 1.0);
    if ((ga < gb ? ga : gb) <= margin) {
      flag[k] = 1;
      atomicAdd(&ctr[GCTR_GUARD], 1ull);
    }
  }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_guard(const GenoDev& g, const double* gl, const double* cert, const double* af, double min_cert,
                        double margin_scale, int8_t* flag, double* gaps, unsigned long long* ctr, hipStream_t st) {
  if (g.nregions <= 0) return hipSuccess;
  const int blocks = (int)((g.nregions + 255) / 256 < 4096 ? (g.nregions + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_guard, dim3(blocks), dim3(256), 0, st, g, gl, cert, af, min_cert, margin_scale, flag, gaps,
                     ctr);
  return hipGetLastError();
}

hipError_t launch_molecule(const GenoDev& g, const double* ll, const uint8_t* mq, const double* mq_err,
                           double min_cert, double* values, double* cert, double* mol, int8_t* keep,
                           unsigned long long* ctr, int cus, hipStream_t st) {
  if (g.nregions <= 0) return hipSuccess;
  const int64_t want = (g.nregions + 3) / 4;
  const int blocks = (int)(want < (int64_t)cus * 32 ? want : (int64_t)cus * 32);
  hipLaunchKernelGGL(k_molecule, dim3(blocks), dim3(256), 0, st, g, ll, mq, mq_err, min_cert, values, cert, mol,
                     keep, ctr);
  return hipGetLastError();
}

hipError_t launch_mixture(const GenoDev& g, const double* mat, const int8_t* keep, double prior, int triangular,
                          int64_t* unit_off, int32_t* unit_region, int32_t* kidx, int32_t* nkept, double* gl,
                          unsigned long long* ctr, int cus, hipStream_t st) {
  if (g.nregions <= 0) return hipSuccess;
  const int cb = (int)((g.nregions + 255) / 256 < 4096 ? (g.nregions + 255) / 256 : 4096);
  if (keep) {
    const int64_t want = (g.nregions + 3) / 4;  // 4 waves (regions) per block
    const int kb = (int)(want < (int64_t)cus * 32 ? want : (int64_t)cus * 32);
    hipLaunchKernelGGL(k_keep_compact, dim3(kb), dim3(256), 0, st, g, keep, kidx, nkept);
  }
  hipLaunchKernelGGL(k_mix_count, dim3(cb), dim3(256), 0, st, g, triangular, unit_off);
  hipLaunchKernelGGL(k_mix_scan, dim3(1), dim3(1024), 0, st, unit_off, g.nregions, ctr);
  // a unit count beyond the map's capacity (a layout whose blocks overlap) is
  // clipped by the map and the mixture; the host reads GCTR_UNITS to report it
  hipLaunchKernelGGL(k_mix_map, dim3(cb), dim3(256), 0, st, g, (const int64_t*)unit_off, unit_region);
  hipLaunchKernelGGL(k_mixture, dim3(cus * 8), dim3(256), 0, st, g, mat, keep, prior, triangular,
                     (const int64_t*)unit_off, (const int32_t*)unit_region, (const int32_t*)kidx,
                     (const int32_t*)nkept, gl, ctr);
  return hipGetLastError();
}

}  // namespace vd

// Diploid genotype likelihood over haplotype pairs, on the device.
//
// Two kernels per region batch (regions = rows x haplotypes blocks laid out
// like the likelihood matrix of the forward):
//
// k_molecule   one wave per region, one lane per row: normalized rows with the
//              mapping-quality mixture and the certainty (molecule_rows.py:6-14),
//              the certainty filter and the single-alignment molecule row
//              exp(log(max(1e-300, v)) - max) (whole_read.py:36-46).
// k_mixture    one lane per haplotype pair, 64 consecutive pairs per wave
// k_mix_wide   (small regions share a wave; a pair of a region with more than
//              8 haplotypes takes one lane per 8 artifact mixtures instead):
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

// std::max(1e-300, v) = (1e-300 < v) ? v : 1e-300, which is fmax's result for
// every v (NaN included: fmax returns the other operand): one v_max_f64
__device__ __forceinline__ double floor300(double v) { return fmax(v, 1e-300); }
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
// The mixture of mixture_general.cpp:9-40 over wave units.
//
// k_mix_wide  regions with more haplotypes (n > MIX_NARROW_MAX; artifact
//             prior > 0), L lanes per pair: in round r lane l sweeps the kept
//             rows for artifact pass p = rL + l (indices j = 8p .. 8p + 7);
//             after each round the L passes' totals are log-added in index
//             order into every lane of the pair by cross-lane reads.  The
//             artifact total goes to the pair's output slot.  A pair then
//             takes about MIX_WIDE_SWEEPS sweeps of wave time instead of
//             1 + n/2.
// k_mixture   one lane per pair: 64 consecutive pairs of the regions' pair
//             sequence per wave, so a wave holds the pairs of several small
//             regions; a lane sweeps the kept rows once for the 7
//             allele-balance mixtures and, in a narrow region, once per 8 of
//             the 4n artifact mixtures, folding the totals in order; in a wide
//             region it takes the artifact total k_mix_wide left in the
//             output slot.
//
// Either way every total is the same sequence of operations as one lane
// carrying the whole pair, and the folds run in the reference's order.

enum : int8_t { MIX_OK = 0, MIX_INVALID = 1, MIX_LAYOUT = 2 };

constexpr int MIXW = 8;            // artifact totals per sweep of the rows
constexpr int MIX_NARROW_MAX = 8;   // haplotypes of a region whose pairs take one lane each
constexpr int MIX_WIDE_ROWS = 16;   // kept rows below which a pair's sweeps cost less than the
                                    // wide kernel's in-order fold (also one lane per pair)
constexpr int MIX_WIDE_SWEEPS = 8;  // artifact passes per lane of a wide region (more when a pair
                                    // would need more than 64 lanes)

// per-region counts, arrays of nregions + 1 (exclusive offsets after the scan)
enum { MIXC_PAIRS = 0, MIXC_WIDE = 1, MIXC_N = 2 };

// lanes per pair of k_mix_wide for a region of n haplotypes and nk kept rows; 0: k_mixture alone
__device__ __forceinline__ int mix_wide_lanes(int64_t n, int64_t nk, double prior, int wide) {
  if (!wide || !(prior > 0) || n <= MIX_NARROW_MAX || nk < MIX_WIDE_ROWS) return 0;
  const int64_t npass = (4 * n + MIXW - 1) / MIXW;
  int L = 1;
  while (L < 64 && (int64_t)L * MIX_WIDE_SWEEPS < npass) L <<= 1;
  return L;
}

// per region (one wave): the layout check, the validation of mixture_cpu.py:14,25
// (every entry of a kept row finite and non-negative), the kept rows in row
// order (the mixture loops then visit only the rows the certainty filter kept,
// with no branch on a keep load per row; at 30x most rows are dropped), the
// region's pairs for k_mixture or its wave units for k_mix_wide
__global__ __launch_bounds__(256) void k_mix_prep(GenoDev g, const double* __restrict__ mat,
                                                  const int8_t* __restrict__ keep, int triangular, double prior,
                                                  int wide, int32_t* __restrict__ kidx, int32_t* __restrict__ nkept,
                                                  int8_t* __restrict__ status, int64_t* __restrict__ cnt,
                                                  unsigned long long* __restrict__ ctr) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t stride = g.nregions + 1;
  for (int64_t k = wave; k < g.nregions; k += nwaves) {
    const int64_t r0 = g.row_off[k], nr = g.row_off[k + 1] - r0;
    const int64_t n = g.col_off[k + 1] - g.col_off[k];
    const int64_t pairs = n <= 0 ? 0 : (triangular ? n * (n + 1) / 2 : n * n);
    const int64_t mo = g.mat_off[k], go = g.gl_off[k];
    int8_t st = MIX_OK;
    int64_t nk = 0;
    if (pairs > 0) {
      if (r0 < 0 || nr < 0 || r0 + nr > g.nrows || mo < 0 || go < 0 || mo + nr * n > g.mat_size ||
          go + pairs > g.gl_size) {  // uniform
        st = MIX_LAYOUT;
      } else {
        const double* M = mat + mo;
        bool bad = false;
        for (int64_t i = lane; i < nr * n; i += 64) {
          const int64_t r = i / n;
          if (keep && !keep[r0 + r]) continue;
          const double v = M[i];
          bad |= !(v >= 0 && v <= 1.7976931348623157e308);
        }
        if (__builtin_amdgcn_ballot_w64(bad)) st = MIX_INVALID;
        if (keep) {
          for (int64_t c = 0; c < nr; c += 64) {
            const int64_t r = c + lane;
            const bool kp = r < nr && keep[r0 + r];
            const unsigned long long m = __builtin_amdgcn_ballot_w64(kp);
            if (kp) kidx[r0 + nk + __builtin_popcountll(m & ((1ull << lane) - 1))] = (int32_t)r;
            nk += __builtin_popcountll(m);
          }
        } else {
          nk = nr;
        }
      }
    }
    if (lane == 0) {
      status[k] = st;
      nkept[k] = (int32_t)nk;
      // an invalid region's pairs are written NaN by k_mixture
      const int L = st == MIX_OK ? mix_wide_lanes(n, nk, prior, wide) : 0;
      cnt[MIXC_PAIRS * stride + k] = st == MIX_LAYOUT ? 0 : pairs;
      cnt[MIXC_WIDE * stride + k] = L ? (pairs * L + 63) / 64 : 0;
      if (st == MIX_LAYOUT) atomicAdd(&ctr[GCTR_LAYOUT], 1ull);
      if (st == MIX_INVALID) atomicAdd(&ctr[GCTR_INVALID], 1ull);
    }
  }
}

// exclusive scans of the MIXC_N count arrays in place (n entries each, total
// into entry n); ctr[GCTR_UNITS] = k_mixture's wave units (64 pairs each),
// ctr[GCTR_UNITS_WIDE] = k_mix_wide's
__global__ __launch_bounds__(1024) void k_mix_scan(int64_t* __restrict__ cnt, int64_t n,
                                                   unsigned long long* __restrict__ ctr) {
  __shared__ int64_t part[1024];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t chunk = (n + nt - 1) / nt;
  const int64_t lo = tid * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (int arr = 0; arr < MIXC_N; ++arr) {
    int64_t* a = cnt + arr * (n + 1);
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
      ctr[arr == MIXC_PAIRS ? GCTR_UNITS : GCTR_UNITS_WIDE] =
          (unsigned long long)(arr == MIXC_PAIRS ? (part[nt - 1] + 63) / 64 : part[nt - 1]);
    }
    __syncthreads();
  }
}

// unit_first[u] = the region holding pair 64u (the first pair of k_mixture's unit u)
__global__ void k_mix_map(GenoDev g, const int64_t* __restrict__ pair_off, int32_t* __restrict__ unit_first) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < g.nregions;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p0 = pair_off[k], p1 = pair_off[k + 1];
    for (int64_t u = (p0 + 63) / 64; u * 64 < p1 && u < g.unit_cap; ++u) unit_first[u] = (int32_t)k;
  }
}

// the last k in [lo, hi] with off[k] <= P (off non-decreasing, off[lo] <= P)
__device__ __forceinline__ int last_at_or_below(const int64_t* __restrict__ off, int lo, int hi, int64_t P) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= P) lo = mid;
    else hi = mid - 1;
  }
  return lo;
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

struct MixView {
  const double* M;      // the region's matrix
  const int32_t* kidx;  // the region's kept rows (null: every row)
  int64_t nk;           // kept rows
  int n;
  __device__ __forceinline__ const double* row(int64_t i) const { return M + (kidx ? (int64_t)kidx[i] : i) * n; }
};

// the pair's mean without a mixture (mixture_general.cpp:23); each row's two
// values are loaded one row ahead, as in the sweeps below
__device__ __forceinline__ double mix_mean(const MixView& v, int a, int b) {
#pragma clang fp contract(off)
  LogProd t;
  t.init();
  double na = 0, nb = 0;
  if (v.nk > 0) {
    const double* row = v.row(0);
    na = row[a];
    nb = row[b];
  }
  for (int64_t i = 0; i < v.nk; ++i) {
    const double la = na, lb = nb;
    if (i + 1 < v.nk) {
      const double* row = v.row(i + 1);
      na = row[a];
      nb = row[b];
    }
    t.mul(floor300((la + lb) / 2));
  }
  return t.total();
}

// allele balance (mixture_general.cpp:25, :37): the 7 totals log-added in order
__device__ __forceinline__ double mix_balance(const MixView& v, int a, int b) {
#pragma clang fp contract(off)
  LogProd tb[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) tb[j].init();
  double na = 0, nb = 0;
  if (v.nk > 0) {
    const double* row = v.row(0);
    na = row[a];
    nb = row[b];
  }
  for (int64_t i = 0; i < v.nk; ++i) {
    const double la = na, lb = nb;
    if (i + 1 < v.nk) {
      const double* row = v.row(i + 1);
      na = row[a];
      nb = row[b];
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) tb[j].mul(floor300(G_BAL[j] * la + (1 - G_BAL[j]) * lb));
  }
  double clean = tb[0].total() + log(G_BW[0]);
#pragma unroll
  for (int j = 1; j < 7; ++j) clean = log_add(clean, tb[j].total() + log(G_BW[j]));
  return clean;
}

// artifact totals j0 .. j0 + MIXW - 1 (those below 4n) of index j = t n + c
// (mixture_general.cpp:26, :38), each with its weight log(w_t / n)
__device__ __forceinline__ void mix_artifact(const MixView& v, int a, int b, int j0, double (&x)[MIXW]) {
#pragma clang fp contract(off)
  const int n = v.n, nj = 4 * n;
  double th[MIXW], om[MIXW];
  int cc[MIXW];
  LogProd ta[MIXW];
#pragma unroll
  for (int s = 0; s < MIXW; ++s) {
    const int j = j0 + s < nj ? j0 + s : nj - 1;  // in bounds; slots past nj are not folded
    const int t = j / n;
    th[s] = G_THETA[t];
    om[s] = 1 - G_THETA[t];
    cc[s] = j - t * n;
    ta[s].init();
  }
  double nsab = 0, nc[MIXW];
#pragma unroll
  for (int s = 0; s < MIXW; ++s) nc[s] = 0;
  if (v.nk > 0) {
    const double* row = v.row(0);
    nsab = row[a] + row[b];
#pragma unroll
    for (int s = 0; s < MIXW; ++s) nc[s] = row[cc[s]];
  }
  for (int64_t i = 0; i < v.nk; ++i) {
    const double sab = nsab;
    double rc[MIXW];
#pragma unroll
    for (int s = 0; s < MIXW; ++s) rc[s] = nc[s];
    if (i + 1 < v.nk) {
      const double* row = v.row(i + 1);
      nsab = row[a] + row[b];
#pragma unroll
      for (int s = 0; s < MIXW; ++s) nc[s] = row[cc[s]];
    }
#pragma unroll
    for (int s = 0; s < MIXW; ++s) {
      const double half = om[s] * sab / 2;  // om * (row[a] + row[b]) / 2
      ta[s].mul(floor300(half + th[s] * rc[s]));
    }
  }
#pragma unroll
  for (int s = 0; s < MIXW; ++s) {
    const int j = j0 + s < nj ? j0 + s : nj - 1;
    x[s] = ta[s].total() + log(G_TW[j / n] / n);
  }
}

// The artifact total of a pair whose wave holds one region only: the region's
// row pointers and the artifact column are then the same in every lane (so
// the column values are scalar loads), and the passes follow the
// contamination levels (8 columns of one level per pass) with each row's
// (1 - theta)(row[a] + row[b]) / 2 formed once per level.
__device__ __forceinline__ double mix_artifact_one_region(const MixView& v, int a, int b) {
#pragma clang fp contract(off)
  const int n = v.n;
  double art = 0;
  for (int t = 0; t < 4; ++t) {
    const double th = G_THETA[t], om = 1 - G_THETA[t];
    const double lt = log(G_TW[t] / n);
    for (int c0 = 0; c0 < n; c0 += MIXW) {
      const int nc = n - c0 < MIXW ? n - c0 : MIXW;
      LogProd ta[MIXW];
#pragma unroll
      for (int s = 0; s < MIXW; ++s) ta[s].init();
      double na = 0, nb = 0;
      if (v.nk > 0) {
        const double* row = v.row(0);
        na = row[a];
        nb = row[b];
      }
      for (int64_t i = 0; i < v.nk; ++i) {
        const double* row = v.row(i);
        const double half = om * (na + nb) / 2;
        if (i + 1 < v.nk) {
          const double* nrow = v.row(i + 1);
          na = nrow[a];
          nb = nrow[b];
        }
#pragma unroll
        for (int s = 0; s < MIXW; ++s) {
          const int c = c0 + s < n ? c0 + s : n - 1;  // in bounds; slots past n are not folded
          ta[s].mul(floor300(half + th * row[c]));
        }
      }
#pragma unroll
      for (int s = 0; s < MIXW; ++s) {  // static indices keep ta[] in registers
        if (s < nc) {
          const double x = ta[s].total() + lt;
          art = (t == 0 && c0 + s == 0) ? x : log_add(art, x);
        }
      }
    }
  }
  return art;
}

__global__ __launch_bounds__(256) void k_mixture(GenoDev g, const double* __restrict__ mat,
                                                 const int8_t* __restrict__ keep, double prior, int triangular,
                                                 int wide,
                                                 const int64_t* __restrict__ pair_off,
                                                 const int32_t* __restrict__ unit_first,
                                                 const int32_t* __restrict__ kidx, const int32_t* __restrict__ nkept,
                                                 const int8_t* __restrict__ status, double* __restrict__ gl,
                                                 unsigned long long* __restrict__ ctr) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t nunits = (int64_t)ctr[GCTR_UNITS] < g.unit_cap ? (int64_t)ctr[GCTR_UNITS] : g.unit_cap;
  const int64_t total = pair_off[g.nregions];
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < nunits; u += nwaves) {
    const int64_t P = u * 64 + lane;
    const bool active = P < total;
    const int64_t Pc = active ? P : total - 1;  // idle lanes (last unit only) shadow the last pair
    // the region of pair Pc lies in [unit_first[u], unit_first[u + 1]]
    const int k = last_at_or_below(pair_off, unit_first[u], u + 1 < nunits ? unit_first[u + 1] : (int)g.nregions - 1,
                                   Pc);
    const int n = (int)(g.col_off[k + 1] - g.col_off[k]);
    const int64_t q = Pc - pair_off[k];
    const PairAB ab = pair_of(q, n, triangular);
    const bool invalid = status[k] != MIX_OK;
    const MixView v{mat + g.mat_off[k], keep ? kidx + g.row_off[k] : nullptr, invalid ? 0 : (int64_t)nkept[k], n};
    double* out = gl + g.gl_off[k] + q;
    double res;
    if (prior < 0) {
      res = mix_mean(v, ab.a, ab.b);
    } else {
      const double clean = mix_balance(v, ab.a, ab.b);
      if (prior == 0) {
        res = clean;
      } else if (!invalid && mix_wide_lanes(n, v.nk, prior, wide)) {
        res = log_add(log1p(-prior) + clean, log(prior) + *out);  // the artifact total of k_mix_wide
      } else if (!__builtin_amdgcn_ballot_w64(k != __builtin_amdgcn_readfirstlane(k))) {
        const int ku = __builtin_amdgcn_readfirstlane(k);
        const int nu = (int)(g.col_off[ku + 1] - g.col_off[ku]);
        const MixView vu{mat + g.mat_off[ku], keep ? kidx + g.row_off[ku] : nullptr,
                         status[ku] != MIX_OK ? 0 : (int64_t)nkept[ku], nu};
        res = log_add(log1p(-prior) + clean, log(prior) + mix_artifact_one_region(vu, ab.a, ab.b));
      } else {
        double art = 0;
        for (int j0 = 0; j0 < 4 * n; j0 += MIXW) {
          double x[MIXW];
          mix_artifact(v, ab.a, ab.b, j0, x);
#pragma unroll
          for (int s = 0; s < MIXW; ++s)  // static indices keep x[] in registers
            if (j0 + s < 4 * n) art = j0 + s == 0 ? x[s] : log_add(art, x[s]);
        }
        res = log_add(log1p(-prior) + clean, log(prior) + art);
      }
    }
    if (active) *out = invalid ? __builtin_nan("") : res;
  }
}

constexpr int MIX_STAGE = 2048;  // doubles of kept rows staged in LDS per wave (16 KiB)

// one wave per block: a block's LDS stage is released as soon as its wave ends
__global__ __launch_bounds__(64) void k_mix_wide(GenoDev g, const double* __restrict__ mat,
                                                  const int8_t* __restrict__ keep, double prior, int triangular,
                                                  const int64_t* __restrict__ wide_off,
                                                  const int32_t* __restrict__ kidx, const int32_t* __restrict__ nkept,
                                                  double* __restrict__ gl, unsigned long long* __restrict__ ctr) {
#pragma clang fp contract(off)
  __shared__ double stage[MIX_STAGE];
  const int lane = threadIdx.x & 63;
  const int64_t nunits = (int64_t)ctr[GCTR_UNITS_WIDE];
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < nunits; u += nwaves) {
    const int k = __builtin_amdgcn_readfirstlane(last_at_or_below(wide_off, 0, (int)g.nregions - 1, u));
    const int n = (int)(g.col_off[k + 1] - g.col_off[k]);
    const int64_t nk = nkept[k];  // uniform: one region per wave
    const int L = mix_wide_lanes(n, nk, prior, 1);  // a unit of k_mix_wide exists only when wide
    const int64_t pairs = triangular ? (int64_t)n * (n + 1) / 2 : (int64_t)n * n;
    const int slot = lane / L, l = lane - slot * L;
    const int64_t q = (u - wide_off[k]) * (64 / L) + slot;
    const bool active = q < pairs;
    const int npass = (4 * n + MIXW - 1) / MIXW, nj = 4 * n;
    const PairAB ab = pair_of(active ? q : pairs - 1, n, triangular);
    const double* M = mat + g.mat_off[k];
    const int32_t* rows = keep ? kidx + g.row_off[k] : nullptr;
    const int rpc = MIX_STAGE / n;  // kept rows per staged chunk
    double art = 0;
    for (int r = 0; r * L < npass; ++r) {
      const int p = r * L + l;  // this lane's pass of the round
      const bool work = active && p < npass;
      const int j0 = p < npass ? p * MIXW : 0;
      double th[MIXW], om[MIXW];
      int cc[MIXW];
      LogProd ta[MIXW];
#pragma unroll
      for (int s = 0; s < MIXW; ++s) {
        const int j = j0 + s < nj ? j0 + s : nj - 1;  // in bounds; slots past nj are not folded
        const int t = j / n;
        th[s] = G_THETA[t];
        om[s] = 1 - G_THETA[t];
        cc[s] = j - t * n;
        ta[s].init();
      }
      if (rpc > 0) {
        // the kept rows pass through LDS in chunks: one coalesced load of a
        // chunk per wave, then every lane reads its pair's entries from LDS
        for (int64_t c0 = 0; c0 < nk; c0 += rpc) {
          const int nr = (int)(nk - c0 < rpc ? nk - c0 : rpc);
          for (int rr = 0; rr < nr; ++rr) {
            const double* src = M + (rows ? (int64_t)rows[c0 + rr] : c0 + rr) * n;
            for (int col = lane; col < n; col += 64) stage[rr * n + col] = src[col];
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (work) {
            for (int rr = 0; rr < nr; ++rr) {
              const double* row = stage + rr * n;
              const double sab = row[ab.a] + row[ab.b];
#pragma unroll
              for (int s = 0; s < MIXW; ++s) {
                const double half = om[s] * sab / 2;  // om * (row[a] + row[b]) / 2
                ta[s].mul(floor300(half + th[s] * row[cc[s]]));
              }
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();  // the chunk is read before the next overwrites it
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
      } else if (work) {  // rows wider than the stage: straight from memory
        for (int64_t i = 0; i < nk; ++i) {
          const double* row = M + (rows ? (int64_t)rows[i] : i) * n;
          const double sab = row[ab.a] + row[ab.b];
#pragma unroll
          for (int s = 0; s < MIXW; ++s) {
            const double half = om[s] * sab / 2;
            ta[s].mul(floor300(half + th[s] * row[cc[s]]));
          }
        }
      }
      double x[MIXW];
#pragma unroll
      for (int s = 0; s < MIXW; ++s) {
        const int j = j0 + s < nj ? j0 + s : nj - 1;
        x[s] = ta[s].total() + log(G_TW[j / n] / n);
      }
      // the round's totals in index order: pass r L + l2 from lane slot L + l2
      for (int l2 = 0; l2 < L && r * L + l2 < npass; ++l2) {
        const int src = slot * L + l2;
#pragma unroll
        for (int s = 0; s < MIXW; ++s) {
          const double y = __shfl(x[s], src);
          const int j = (r * L + l2) * MIXW + s;
          if (j < nj) art = j == 0 ? y : log_add(art, y);
        }
      }
    }
    if (active && l == 0) gl[g.gl_off[k] + q] = art;
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
                          const MixScratch& w, double* gl, unsigned long long* ctr, int cus, hipStream_t st) {
  if (g.nregions <= 0) return hipSuccess;
  const int cb = (int)((g.nregions + 255) / 256 < 4096 ? (g.nregions + 255) / 256 : 4096);
  const int64_t want = (g.nregions + 3) / 4;  // 4 waves (regions) per block
  const int kb = (int)(want < (int64_t)cus * 32 ? want : (int64_t)cus * 32);
  const int64_t stride = g.nregions + 1;
  // With fewer pair units than about four per wave slot of the GPU, the
  // units of the regions with many haplotypes (1 + n/2 sweeps of their kept
  // rows per lane) outlast the rest and set the kernel's time: those regions
  // then go to k_mix_wide.  Above that the in-lane form, whose fold runs once
  // per pair, does less work in all.
  const int wide = prior > 0 && g.gl_size < (int64_t)64 * 4 * 8 * cus;
  hipLaunchKernelGGL(k_mix_prep, dim3(kb), dim3(256), 0, st, g, mat, keep, triangular, prior, wide, w.kidx,
                     w.nkept, w.status, w.cnt, ctr);
  hipLaunchKernelGGL(k_mix_scan, dim3(1), dim3(1024), 0, st, w.cnt, g.nregions, ctr);
  // a unit count beyond the map's capacity (a layout whose blocks overlap) is
  // clipped by the map and the mixture; the host reads GCTR_UNITS to report it
  hipLaunchKernelGGL(k_mix_map, dim3(cb), dim3(256), 0, st, g, (const int64_t*)(w.cnt + MIXC_PAIRS * stride),
                     w.unit_first);
  if (wide)  // the wide regions' units first: they are the longest
    hipLaunchKernelGGL(k_mix_wide, dim3(cus * 32), dim3(64), 0, st, g, mat, keep, prior, triangular,
                       (const int64_t*)(w.cnt + MIXC_WIDE * stride), (const int32_t*)w.kidx,
                       (const int32_t*)w.nkept, gl, ctr);
  hipLaunchKernelGGL(k_mixture, dim3(cus * 8), dim3(256), 0, st, g, mat, keep, prior, triangular, wide,
                     (const int64_t*)(w.cnt + MIXC_PAIRS * stride), (const int32_t*)w.unit_first,
                     (const int32_t*)w.kidx, (const int32_t*)w.nkept, (const int8_t*)w.status, gl, ctr);
  return hipGetLastError();
}

}  // namespace vd

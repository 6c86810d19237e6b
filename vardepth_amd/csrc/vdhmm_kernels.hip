// MI355X (gfx950) PairHMM kernels.
//
// The forward recurrence of the reference (batch_hmm.cpp:31-70 dense form,
// trie_vector.cpp:259-277 column form) is evaluated along anti-diagonals:
//
// * A 16-lane group (one DPP row) owns one read x haplotype pair; lane l holds
//   R consecutive read rows, all at haplotype column t-l-1 at step t.  Rows
//   inside a lane exchange values through registers; the row above a lane's
//   first row lives in lane l-1 and arrives through a DPP row_shr:1 move whose
//   `old` operand injects the boundary row into lane 0.  A wave runs 4 pairs.
// * The recurrence is rescaled so that every cell costs 5 FMA/MUL plus the
//   emission select: with per-row power-of-two normalisers rho_i,
//     D^_i(j) = M^_i(j-1) + x D^_i(j-1)
//     P_i(j)  = M^_i(j) + mu_{i+1} I^_i(j)        (= I^_{i+1}(j))
//     Q_i(j)  = M^_i(j) + beta_{i+1} I^_i(j) + gamma D^_i(j)
//     M^_i(j) = e'_ij Q_{i-1}(j-1)
//   (M = kappa_i M^, I = c kappa_{i-1} I^, D = g kappa_i D^, kappa_i =
//   kappa_{i-1} rho_i), which is the reference model exactly; only rounding
//   differs.
// * Reads are bottom-aligned in the group's 16R rows so that the last read
//   row is always lane 15 / register R-1.  Rows above the read are
//   pass-through rows (e'=0, mu=beta=1) that carry the injected boundary row
//   down unchanged.  The semiglobal terminal sum_j (M+I)(n,j) is the last
//   row's P stream (mu_last = c/rho_n) summed over the steps; columns past the
//   haplotype end get zero emission so that the sum needs no masking.
//   Anchored mode prepends a virtual row 0 (e'=1, mu=beta=0) that builds the
//   reference's init row M(0,0)=1, D(0,j)=g x^(j-1).
//
// FP32 runs the bulk with a start scale of 2^108; a pair whose scaled
// terminal falls below 2^-100 is appended to the FP64 work list.  The FP64
// kernel uses the same recurrence with 64-row tiles, renormalising the
// boundary row between tiles, so it covers any read length.
#include <type_traits>

#include "vdhmm_device.h"

namespace vd {

// ---------------------------------------------------------------------------
// planning

// Class and normaliser sums of every read (route64: bit 0 sends every read to
// the FP64 path, bit c the reads of class c): one 16-lane group per read, the
// per-bin tables in LDS, lane partial sums in a fixed order, then a fixed
// 16-lane butterfly, so the FP64 sum of a read does not depend on the batch.
template <int XM>
__device__ __forceinline__ int swz_xor_i(int v) {
  return __builtin_amdgcn_ds_swizzle(v, (XM << 10) | 0x1f);  // bitmask mode: xor XM within 32 lanes
}

__global__ __launch_bounds__(256) void k_prep_reads(BatchDev b, const int32_t* __restrict__ qE,
                                                    const QTab<float>* __restrict__ tab32, int anchored, int route64,
                                                    int8_t* __restrict__ cls, int32_t* __restrict__ exps,
                                                    double* __restrict__ ln32) {
  __shared__ int sE[64];
  __shared__ double sL[64];
  for (int i = threadIdx.x; i < 64; i += blockDim.x) {
    sE[i] = qE[i];
    sL[i] = tab32->lnrho[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
  const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t rb = wave0 * 4; rb < b.nreads; rb += nwaves * 4) {  // uniform over the wave
    const int64_t r = rb + g;
    const bool ok = r < b.nreads;
    const int64_t o0 = ok ? b.read_off[r] : 0, o1 = ok ? b.read_off[r + 1] : 0;
    int e = 0;
    double lsum = 0.0;
    for (int64_t k = o0 + l16; k < o1; k += 16) {
      int q = b.qual_bytes[k];
      q = q < 63 ? q : 63;
      e += sE[q];
      lsum += sL[q];
    }
    e += swz_xor_i<1>(e);
    e += swz_xor_i<2>(e);
    e += swz_xor_i<4>(e);
    e += swz_xor_i<8>(e);
    lsum = group_sum(lsum);
    if (ok && l16 == 0) {
      const int n = (int)(o1 - o0);
      const int np = n + anchored;
      const int c32 = (np + GROUP - 1) / GROUP;
      const bool to64 = (route64 & 1) || c32 > MAX_R32 || ((route64 >> c32) & 1);
      const int c = n == 0 ? CLS_EMPTY : (to64 ? CLS_F64 : c32);
      cls[r] = (int8_t)c;
      exps[r] = e;
      ln32[r] = lsum;
    }
  }
}

// One block per region: class histogram, per-region class offsets, the read
// permutation grouped by class, FP32 unit counts, -inf for empty reads and the
// FP64 list entries of long reads.
__global__ void k_bucket(BatchDev b, PlanDev p, double* __restrict__ out) {
  __shared__ int cnt[NCLS], off[NCLS + 1], fill[NCLS];
  __shared__ unsigned long long base64;
  for (int64_t k = blockIdx.x; k < b.nregions; k += gridDim.x) {
    if (threadIdx.x < NCLS) {
      cnt[threadIdx.x] = 0;
      fill[threadIdx.x] = 0;
    }
    __syncthreads();
    const int64_t r0 = b.region_reads[k], r1 = b.region_reads[k + 1];
    const int64_t h0 = b.region_haps[k], H = b.region_haps[k + 1] - h0;
    const int64_t o0 = b.out_off[k];
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      int c = p.read_cls[r];
      atomicAdd(&cnt[c], 1);
      if (c == CLS_EMPTY)
        for (int64_t h = 0; h < H; ++h) out[o0 + (r - r0) * H + h] = -__builtin_inf();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      off[0] = 0;
      for (int c = 0; c < NCLS; ++c) off[c + 1] = off[c] + cnt[c];
      for (int c = 0; c <= NCLS; ++c) p.class_off[k * (NCLS + 1) + c] = off[c];
      unsigned long long n64 = (unsigned long long)cnt[CLS_F64] * (unsigned long long)H;
      base64 = n64 ? atomicAdd(&p.ctr[CTR_F64], n64) : 0ull;
    }
    __syncthreads();
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      int c = p.read_cls[r];
      int s = atomicAdd(&fill[c], 1);
      p.read_perm[r0 + off[c] + s] = r;
      if (c == CLS_F64) {
        for (int64_t h = 0; h < H; ++h) {
          unsigned long long idx = base64 + (unsigned long long)s * H + h;
          if ((int64_t)idx < p.f64_cap) {
            p.f64_list[idx] = PairRef{r, h0 + h, o0 + (r - r0) * H + h};
          } else {
            atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
          }
        }
      }
    }
    if (threadIdx.x < NCLS) {
      int c = threadIdx.x;
      // pair units (4 read x haplotype pairs), or with prefix sharing read
      // units (4 reads, each against every haplotype of the region)
      const bool tc = (p.trie >> c) & 1;
      int64_t units = 0;
      if (c >= 1 && c <= MAX_R32 && H > 0)
        units = p.pack_pairs ? (int64_t)cnt[c] * H
                : tc ? ((int64_t)cnt[c] + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT * ((H + p.trie_chunk - 1) / p.trie_chunk)
                     : ((int64_t)cnt[c] * H + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT;
      p.unit_cnt[bucket_of(c, k, b.nregions, p.class_major)] = units;
    }
    __syncthreads();
  }
}

// Exclusive scan of unit_cnt[0..n) in place, in three passes: the sum of
// every SCAN_TILE-entry tile, a scan of those sums (one block), each tile
// scanned from its offset.  unit_cnt[n] and ctr[CTR_NUNITS] receive the
// total; the tile sums sit after unit_cnt[n].
constexpr int SCAN_THREADS = 256, SCAN_PER_THREAD = SCAN_TILE / SCAN_THREADS;

// Exclusive scan of one value per thread over a block of NT threads (sh: NT
// entries of LDS); `total` gets the block's sum.  Every thread must call it.
template <int NT>
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* sh, int64_t& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < NT; d <<= 1) {
    const int64_t x = t >= d ? sh[t - d] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const int64_t incl = sh[t];
  total = sh[NT - 1];
  __syncthreads();  // sh free for the caller's next scan
  return incl - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums(const int64_t* __restrict__ a, int64_t n,
                                                                 int64_t* __restrict__ tsum) {
  __shared__ int64_t sh[SCAN_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
  for (int j = threadIdx.x; j < SCAN_TILE; j += SCAN_THREADS) {
    const int64_t i = base + j;
    if (i < n) s += a[i];
  }
  int64_t tot;
  block_excl_scan<SCAN_THREADS>(s, sh, tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(int64_t* __restrict__ tsum, int64_t nt, int64_t* __restrict__ a,
                                                     int64_t n, unsigned long long* ctr) {
  __shared__ int64_t sh[1024];
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < nt; c0 += 1024) {
    const int64_t i = c0 + threadIdx.x;
    const int64_t v = i < nt ? tsum[i] : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan<1024>(v, sh, tot);
    if (i < nt) tsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    a[n] = carry;
    ctr[CTR_NUNITS] = (unsigned long long)carry;
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(int64_t* __restrict__ a, int64_t n,
                                                             const int64_t* __restrict__ tsum) {
  __shared__ int64_t sh[SCAN_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER_THREAD;
  int64_t v[SCAN_PER_THREAD];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) {
    const int64_t i = base + k;
    v[k] = i < n ? a[i] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t run = tsum[blockIdx.x] + block_excl_scan<SCAN_THREADS>(s, sh, tot);
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) {
    const int64_t i = base + k;
    if (i < n) a[i] = run;
    run += v[k];
  }
}

// ---- haplotype prefix sharing ---------------------------------------------

constexpr int TRIE_MAXH = 256;  // regions with more haplotypes are swept without sharing

// Common prefix length of haplotypes i and j (global indices) and whether i
// sorts before j (byte order, a proper prefix first, ties by index), compared
// 64 bytes per step by the whole wave from byte `from` on (a prefix both are
// known to share).  All lanes must call it with the same arguments.
__device__ __forceinline__ int hap_cmp_wave(const BatchDev& b, int64_t i, int64_t j, int64_t from, bool& before) {
  const int lane = threadIdx.x & 63;
  const int64_t oi = b.hap_off[i], oj = b.hap_off[j];
  const int64_t li = b.hap_off[i + 1] - oi, lj = b.hap_off[j + 1] - oj;
  const int64_t n = li < lj ? li : lj;
  for (int64_t t0 = from; t0 < n; t0 += 64) {
    const int64_t t = t0 + lane;
    const bool diff = t < n && b.hap_bytes[oi + t] != b.hap_bytes[oj + t];
    const unsigned long long mask = __builtin_amdgcn_ballot_w64(diff);
    if (mask) {
      const int64_t at = t0 + __builtin_ctzll(mask);
      before = b.hap_bytes[oi + at] < b.hap_bytes[oj + at];
      return (int)at;
    }
  }
  before = li != lj ? li < lj : i < j;
  return (int)n;
}

// Common suffix length of haplotypes i and j, compared from the ends 64 bytes
// per step by the whole wave.  All lanes must call it with the same arguments.
__device__ __forceinline__ int hap_sfx_wave(const BatchDev& b, int64_t i, int64_t j) {
  const int lane = threadIdx.x & 63;
  const int64_t ei = b.hap_off[i + 1], ej = b.hap_off[j + 1];
  const int64_t li = ei - b.hap_off[i], lj = ej - b.hap_off[j];
  const int64_t n = li < lj ? li : lj;
  for (int64_t t0 = 0; t0 < n; t0 += 64) {
    const int64_t t = t0 + lane;
    const bool diff = t < n && b.hap_bytes[ei - 1 - t] != b.hap_bytes[ej - 1 - t];
    const unsigned long long mask = __builtin_amdgcn_ballot_w64(diff);
    if (mask) return (int)(t0 + __builtin_ctzll(mask));
  }
  return (int)n;
}

// One wave (block) per region: haplotypes in byte order (trie_order, by binary
// insertion: O(H log H) wave-wide compares, each starting past the prefix all
// of the region's haplotypes share), the column each resumes from
// (trie_start: the common prefix with its predecessor, at most its length - 1,
// so that at least one column is swept), and the region's distinct fork
// columns in increasing order (trie_fork), one state slot each; a haplotype
// whose fork would need a slot beyond TRIE_SLOTS is swept in full.  Any order
// is correct for the resume (a column j is read from the haplotype that last
// computed it, and every haplotype in between resumed past j, so by the chain
// of common prefixes it shares the reader's first j bases); byte order makes
// neighbours share the most.
//
// With the suffix continuation on, the suffix S every haplotype of the region
// ends in (trie_sfx; 0 below SFX_MIN or with one haplotype) is not swept per
// haplotype: each haplotype h runs to column J_h = m_h - S.  Its start is then
// also at most J_h - 1 and at most J_{h-1} - 1 of its predecessor, so that
// along any chain of resumes the haplotype that last computed a fork column f
// computes it before its own column J (f + 1 <= J).
__global__ __launch_bounds__(64) void k_hap_trie(BatchDev b, PlanDev p) {
  constexpr int PER = (TRIE_MAXH + 63) / 64;  // ranks per lane
  __shared__ int sorted_sh[TRIE_MAXH];
  const int lane = threadIdx.x;
  for (int64_t k = blockIdx.x; k < b.nregions; k += gridDim.x) {
    const int64_t h0 = b.region_haps[k], H = b.region_haps[k + 1] - h0;
    const bool share = H <= TRIE_MAXH;
    if (share) {
      // prefix shared by every haplotype of the region
      int64_t common = 0x7fffffffffffffffll;
      for (int64_t i = 1; i < H; ++i) {
        bool ib;
        const int64_t c = hap_cmp_wave(b, h0, h0 + i, 0, ib);
        common = c < common ? c : common;
      }
      if (H <= 1) common = 0;
      int sfx = 0;
      if (p.sfx && H > 1) {
        sfx = 0x7fffffff;
        for (int64_t i = 1; i < H; ++i) {
          const int c = hap_sfx_wave(b, h0, h0 + i);
          sfx = c < sfx ? c : sfx;
        }
        if (sfx < SFX_MIN) sfx = 0;
      }
      if (lane == 0) p.trie_sfx[k] = sfx;
      // binary insertion into sorted_sh[0..i)
      for (int i = 0; i < (int)H; ++i) {
        int lo = 0, hi = i;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          bool before;
          hap_cmp_wave(b, h0 + i, h0 + sorted_sh[mid], common, before);
          if (before) hi = mid; else lo = mid + 1;
        }
        // shift sorted_sh[lo..i) up by one, then insert
        int v[PER];
#pragma unroll
        for (int w = 0; w < PER; ++w) {
          const int q = lane + 64 * w;
          v[w] = (q > lo && q <= i) ? sorted_sh[q - 1] : 0;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < PER; ++w) {
          const int q = lane + 64 * w;
          if (q > lo && q <= i) sorted_sh[q] = v[w];
        }
        if (lane == 0) sorted_sh[lo] = i;
        __syncthreads();
      }
#pragma unroll
      for (int w = 0; w < PER; ++w) {
        const int q = lane + 64 * w;
        if (q < H) p.trie_order[h0 + q] = (int32_t)(h0 + sorted_sh[q]);
      }
    } else {
      for (int64_t i = lane; i < H; i += 64) p.trie_order[h0 + i] = (int32_t)(h0 + i);
      if (lane == 0) p.trie_sfx[k] = 0;
    }
    __syncthreads();
    const int S = p.trie_sfx[k];
    __syncthreads();
    __shared__ int st_sh[TRIE_MAXH];
    __shared__ int first_sh[TRIE_MAXH];
    for (int64_t q = 0; q < H; ++q) {  // start columns, one wave-wide compare each
      int st = 0;
      if (share && q > 0) {
        const int64_t a = p.trie_order[h0 + q - 1], c = p.trie_order[h0 + q];
        const int m = (int)(b.hap_off[c + 1] - b.hap_off[c]) - S;
        const int ma = (int)(b.hap_off[a + 1] - b.hap_off[a]) - S;
        bool ib;
        st = hap_cmp_wave(b, a, c, 0, ib);
        st = st < m - 1 ? st : m - 1;
        if (S) st = st < ma - 1 ? st : ma - 1;
        st = st > 0 ? st : 0;
      }
      if (lane == 0) {
        p.trie_start[h0 + q] = st;
        if (q < TRIE_MAXH) st_sh[q] = st;
      }
    }
    __syncthreads();
    const int hs = share ? (int)H : 0;  // ranks with a start column (all zero without sharing)
    // first occurrence of each distinct fork column, then the slot of each rank:
    // the number of distinct fork columns below its own
#pragma unroll
    for (int w = 0; w < PER; ++w) {
      const int q = lane + 64 * w;
      if (q < hs) {
        const int st = st_sh[q];
        bool first = st > 0;
        for (int q3 = 0; q3 < q && first; ++q3) first = st_sh[q3] != st;
        first_sh[q] = first ? 1 : 0;
      }
    }
    __syncthreads();
    int slot_of[PER];
    int reps = 0;
#pragma unroll
    for (int w = 0; w < PER; ++w) {
      const int q = lane + 64 * w;
      int slot = -1;
      if (q < hs && st_sh[q] > 0) {
        const int st = st_sh[q];
        slot = 0;
        for (int q2 = 0; q2 < hs; ++q2) slot += (first_sh[q2] && st_sh[q2] < st) ? 1 : 0;
        if (first_sh[q]) {
          ++reps;
          if (slot < TRIE_SLOTS) p.trie_fork[h0 + slot] = st;
        }
      }
      slot_of[w] = slot;
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) reps += __shfl_xor(reps, o);
    __syncthreads();
#pragma unroll
    for (int w = 0; w < PER; ++w) {
      const int64_t q = lane + 64 * w;
      if (q < hs && slot_of[w] >= TRIE_SLOTS) p.trie_start[h0 + q] = 0;
    }
    if (lane == 0) p.trie_nfork[k] = reps < TRIE_SLOTS ? reps : TRIE_SLOTS;
    __syncthreads();
  }
}

// unit -> bucket map by binary search over the scanned counts.
__global__ void k_unit_map(const int64_t* __restrict__ scanned, int64_t nb, const unsigned long long* ctr,
                           int32_t* __restrict__ unit_bucket) {
  const int64_t total = (int64_t)ctr[CTR_NUNITS];
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < total;
       u += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = nb;  // find last b with scanned[b] <= u
    while (hi - lo > 1) {
      int64_t mid = (lo + hi) >> 1;
      if (scanned[mid] <= u) lo = mid; else hi = mid;
    }
    unit_bucket[u] = (int32_t)lo;
  }
}

// Row constants of lane l's R rows for a read of n bases (bottom-aligned in
// the group's 16R rows; rows above are pass-through rows).
template <int R, bool ANCH>
__device__ __forceinline__ void row_constants(const BatchDev& b, const QTab<float>* __restrict__ tab, int64_t read,
                                              int n, int valid, int l, float (&gd)[R], float (&bd)[R],
                                              float (&mu)[R], float (&be)[R], uint32_t (&rc)[R]) {
  const int np = n + (ANCH ? 1 : 0);
  const int pad = GROUP * R - np;
  const int64_t rb = b.read_off[read];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rel = l * R + r - pad;
    float vg = 0.f, vb = 0.f, vm = 0.f, ve = 0.f;
    uint32_t vc = NOMATCH;
    if (valid) {
      if (rel < 0) {
        vm = ve = 1.f;  // pass-through row
      } else if (ANCH && rel == 0) {
        vg = vb = 1.f;  // virtual init row
      } else {
        const int kk = rel - (ANCH ? 1 : 0);
        int q = b.qual_bytes[rb + kk];
        q = q < 63 ? q : 63;
        vc = b.read_bytes[rb + kk];
        vg = tab->good[q];
        vb = tab->bad[q];
        vm = tab->mu[q];
        ve = tab->beta[q];
        if (kk == n - 1) vm = ve = tab->cfin[q];
      }
    }
    gd[r] = vg; bd[r] = vb; mu[r] = vm; be[r] = ve; rc[r] = vc;
  }
}

template <int R, bool ANCH, int U>
__device__ __forceinline__ void unit32(const BatchDev& b, const PlanDev& p, const QTab<float>* __restrict__ tab,
                                    uint16_t* codes, int smax, const GroupPair gp, double* __restrict__ out) {
  const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
  const int tend = gp.m_max + 16;
  const int tmain = gp.m_min + 2;
  fill_codes(b, gp, codes, smax, tend + 18);
  float gd[R], bd[R], mu[R], be[R];
  uint32_t rc[R];
  row_constants<R, ANCH>(b, tab, gp.read, gp.n, gp.valid, l, gd, bd, mu, be, rc);
  const float gam = tab->gamma, xx = tab->ext;
  const float gam_last = l == GROUP - 1 ? (ANCH ? tab->gap : 0.f) : gam;
  __syncthreads();  // codes visible

  float M[R], D[R], Q[R];
#pragma unroll
  for (int r = 0; r < R; ++r) M[r] = D[r] = Q[r] = 0.f;
  const float C = 0x1p108f;
  float Plast = 0.f, Xprev = 0.f, acc = 0.f, rmax = 0.f, latch = 0.f;
  const uint16_t* cl = codes + g * smax + 16 - l;
  // step 0 (column -1 for lane 0): injection of the boundary row's first values
  {
    const float inj = ANCH ? C : 0.f;
    float Pin = dpp_shr1(inj, Plast);
    float Xn = dpp_shr1(inj, Q[R - 1]);
    Plast = adiag_step<float, R, false>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, cl[0]);
    Xprev = Xn;
    if (ANCH) rmax = fmaxf(rmax, Q[R - 1]); else acc += Plast;
  }
  // the boundary row's injection, in lane 0 of each row only
  const float inj = (ANCH || l != 0) ? 0.f : C;
  // One step at a time; the loops are unrolled by an even count so that the
  // loop-carried M/D/Q registers return to the same physical registers
  // without copies.
  auto main_step = [&](int tt) {
    // four steps' codes are one wide LDS read: keep each in a full register
    const uint32_t code = U == 4 ? code_reg(cl[tt]) : (uint32_t)cl[tt];
    float Pin = dpp_shr1_add(Plast, inj);
    float Xn = dpp_shr1_add(Q[R - 1], inj);
    Plast = adiag_step<float, R, false>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code);
    Xprev = Xn;
    if (ANCH) rmax = fmaxf(rmax, Q[R - 1]); else acc += Plast;
  };
  auto tail_step = [&](int tt) {
    const uint32_t code = cl[tt];
    // past the haplotype end the boundary row injects nothing
    const float pinj = (!ANCH && tt - 1 > gp.m) ? 0.f : inj;
    float Pin = dpp_shr1_add(Plast, pinj);
    float Xn = dpp_shr1_add(Q[R - 1], inj);
    Plast = adiag_step<float, R, true>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code);
    Xprev = Xn;
    if (ANCH) {
      rmax = fmaxf(rmax, Q[R - 1]);
      if (tt - l - 1 == gp.m) latch = Q[R - 1];
    } else {
      acc += Plast;
    }
  };
  // U anti-diagonals per iteration of the main loop (U even keeps the
  // loop-carried registers in place); the autotuner picks U per class
  int t = 1;
  for (; t + U - 1 < tmain; t += U) {
#pragma unroll
    for (int k = 0; k < U; ++k) main_step(t + k);
  }
  for (; t < tmain; ++t) main_step(t);
  for (; t + 1 <= tend; t += 2) {
    tail_step(t);
    tail_step(t + 1);
  }
  if (t <= tend) tail_step(t++);
  count_groups(&p.ctr[CTR_SWEPT], gp.valid ? (unsigned long long)gp.n * (unsigned long long)gp.m : 0ull);
  if (l == GROUP - 1 && gp.valid) {
    const double lnk = p.read_ln32[gp.read] - CEXP32 * LN2;
    double v;
    bool flag;
    if (ANCH) {
      // reference clamp ln(max(1e-300 * rowmax, L)), evaluated in log space
      flag = !(latch >= FLAG32_MIN) || !(rmax < __builtin_inff());
      const double lc = log((double)rmax) + LN_1EM300;
      const double ll = log((double)latch);
      v = (ll > lc ? ll : lc) + lnk;
    } else {
      flag = !(acc >= FLAG32_MIN) || !(acc < __builtin_inff());
      v = log((double)acc) + lnk - log((double)(gp.m + 1));
    }
    if (flag) {
      unsigned long long idx = atomicAdd(&p.ctr[CTR_F64], 1ull);
      atomicAdd(&p.ctr[CTR_FLAGGED], 1ull);
      if ((int64_t)idx < p.f64_cap) p.f64_list[idx] = PairRef{gp.read, gp.hap, gp.out};
      else atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
      out[gp.out] = __builtin_nan("");
    } else {
      out[gp.out] = v;
    }
  }
  __syncthreads();  // codes reused by the next unit
}

// FP32 forward of the units of row class R (one kernel per class, so that each
// runs with its own register footprint): the class's contiguous unit range in
// the class-major plan, taken from a per-class counter.
// ---------------------------------------------------------------------------
// Prefix sharing (PlanDev::trie): a read unit gives group g read j*4+g of a
// (region, class) bucket; the four groups sweep the region's haplotypes in
// byte order in lock step (same haplotype, same loop bounds).  After step T of
// a sweep, lane l holds column T-l-1 of its R rows, so the whole wavefront
// after step f+1 (lane 0 at column f, every lane at a column <= f) depends on
// the haplotype's first f bases only.  A sweep stores that wavefront -- each
// lane's (M, D, Q) rows, its DPP pipeline values (Plast, Xprev) and its
// terminal accumulator -- at every fork column f it computes; a haplotype
// whose trie_start is s restores the wavefront of s and continues at step
// s+2.  It then runs exactly the steps of a full sweep from there on, so the
// results equal full sweeps bit for bit.

struct ReadTask {
  int64_t read, region;
  int n, valid;
  int q0, q1;  // ranks of the unit's haplotype chunk
};

__device__ __forceinline__ int decode_read_unit(const BatchDev& b, const PlanDev& p, unsigned long long u,
                                                ReadTask& rt) {
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int bkt = p.unit_bucket[u];
  const int64_t j = (int64_t)u - p.unit_cnt[bkt];
  const int64_t k = p.class_major ? bkt % b.nregions : bkt / NCLS;
  const int c = p.class_major ? NCLS - 1 - (int)(bkt / b.nregions) : bkt % NCLS;
  const int64_t r0 = b.region_reads[k];
  const int32_t* co = p.class_off + k * (NCLS + 1);
  const int64_t cnt = co[c + 1] - co[c];
  const int64_t H = b.region_haps[k + 1] - b.region_haps[k];
  const int64_t nch = (H + p.trie_chunk - 1) / p.trie_chunk;
  const int64_t jr = j / nch, jc = j % nch;  // read group, haplotype chunk
  const int64_t s = jr * PAIRS_PER_UNIT + g;
  rt.valid = s < cnt;
  rt.read = p.read_perm[r0 + co[c] + (rt.valid ? s : 0)];
  rt.region = k;
  rt.n = (int)(b.read_off[rt.read + 1] - b.read_off[rt.read]);
  rt.q0 = (int)(jc * p.trie_chunk);
  rt.q1 = (int)((jc + 1) * p.trie_chunk < H ? (jc + 1) * p.trie_chunk : H);
  return c;
}

// ---------------------------------------------------------------------------
// Suffix continuation (PlanDev::sfx).  When every haplotype of a region ends
// in the same S bases, the terminal of a pair is linear in the sweep's state
// at column J = m - S: the columns after J see that state only through each
// row's Q_i(J) (the diagonal source of M_{i+1}(J+1)) and D^_i(J+1) =
// M^_i(J) + x D^_i(J), plus the boundary injections at columns > J.  The
// weights are the adjoints of the recurrence over the suffix columns,
// which depend on the read and the suffix only:
//   Qb_i(j) = e'_{i+1}(j+1) Mb_{i+1}(j+1)
//   Db_i(j) = gamma_i Qb_i(j) + x Db_i(j+1)
//   Mb_i(j) = Pb_i(j) + Qb_i(j) + Db_i(j+1)
//   Pb_{i-1}(j) = mu_i Pb_i(j) + beta_i Qb_i(j)
// seeded with Pb_n(j) = 2^SFX_SEED_EXP at every suffix column (semiglobal
// terminal sum) or Qb_n(m) = 2^SFX_SEED_EXP (anchored terminal).  The sweep
// runs anti-diagonals backwards: lane l is at suffix column
// k = S + 15 - l - t after step t (k = 1 is column J + 1) and takes the row
// below it from lane l + 1 through DPP row_shl:1.  Per row it leaves
// Qb_i(J) and Db_i(J+1) in the unit's workspace; lane 0 adds up the adjoints
// of the injected boundary values (dT/dP_in at columns J+1..m, dT/dQ_in at
// columns J..m-1), which the haplotype sweeps add at scale 2^CEXP32.
// The reference's CPU trie does the same reduction over a shared suffix
// (trie_vector.cpp:230-257, 274-275) in FP64 with its own backward arrays.
template <int R, bool ANCH, int U>
__device__ __forceinline__ float sfx_adjoint(const float (&gd)[R], const float (&bd)[R], const float (&mu)[R],
                                             const float (&be)[R], const uint32_t (&rc)[R], float gam,
                                             float gam_last, float xx, const uint16_t* codes, int S,
                                             float2* __restrict__ coef) {
  const int lane = threadIdx.x & 63, l = lane & 15;
  float Db[R], Qc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) Db[r] = Qc[r] = 0.f;
  const float seed = 0x1p100f;
  static_assert(SFX_SEED_EXP == 100, "seed constant");
  if (ANCH && l == GROUP - 1) Qc[R - 1] = seed;
  const float injP = ANCH ? 0.f : seed;
  float PaOut = 0.f, YOut = 0.f, F = 0.f;
  const uint16_t* cr = codes + (S + 32 - l);  // cr[-t]: code of this lane's column at step t
  auto astep = [&](int tt, bool acc) {
    const uint32_t code = cr[-tt];
    float Pa = dpp_shl1(injP, PaOut);
    float pend = dpp_shl1(0.f, YOut);
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
      const float Qb = Qc[r];
      Qc[r] = pend;
      const float Mb = Pa + Qb + Db[r];
      Db[r] = fmaT(r == R - 1 ? gam_last : gam, Qb, xx * Db[r]);
      Pa = fmaT(mu[r], Pa, be[r] * Qb);
      pend = (rc[r] == code ? gd[r] : bd[r]) * Mb;
    }
    if (!ANCH && acc) F += Pa + pend;
    PaOut = Pa;
    YOut = pend;
  };
  // steps 0..14: lane 0 above the suffix; 15..S-2: every lane inside; the last
  // 16 steps: lane l reaches column J + 1 at step S + 14 - l and stores
  int t = 0;
#pragma unroll 1
  for (; t < 15; ++t) astep(t, false);
  for (; t + U - 1 < S - 1; t += U) {
#pragma unroll
    for (int kk = 0; kk < U; ++kk) astep(t + kk, true);
  }
#pragma unroll 1
  for (; t < S - 1; ++t) astep(t, true);
#pragma unroll 1
  for (; t < S + 15; ++t) {
    astep(t, true);
    if (t == S + 14 - l) {
#pragma unroll
      for (int r = 0; r < R; ++r) coef[r] = make_float2(Qc[r], Db[r]);
    }
  }
  return F;
}

template <int R, bool ANCH, int U, bool SFX>
__device__ __forceinline__ void unit32_trie(const BatchDev& b, const PlanDev& p, const QTab<float>* __restrict__ tab,
                                            uint16_t* codes, int smax, const ReadTask& rt,
                                            double* __restrict__ out) {
  constexpr int W = (3 * R + 3 + 3) & ~3;  // floats of one lane's wavefront state, padded to 16 bytes
  const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
  float gd[R], bd[R], mu[R], be[R];
  uint32_t rc[R];
  row_constants<R, ANCH>(b, tab, rt.read, rt.n, rt.valid, l, gd, bd, mu, be, rc);
  const float gam = tab->gamma, xx = tab->ext;
  const float gam_last = l == GROUP - 1 ? (ANCH ? tab->gap : 0.f) : gam;
  const int64_t k = __builtin_amdgcn_readfirstlane((int)rt.region);  // one region per unit
  const int64_t h0 = b.region_haps[k], H = b.region_haps[k + 1] - h0;
  const int64_t r0 = b.region_reads[k];
  const int nfork = p.trie_nfork[k];
  // the region's fork columns, one per lane (sorted; lanes past nfork hold INT_MAX),
  // so that per-haplotype lookups are ballots and lane reads, not memory chains
  const int fork_v = lane < nfork ? p.trie_fork[h0 + lane] : 0x7fffffff;
  auto fork_at = [&](int i) { return __builtin_amdgcn_readlane(fork_v, i); };
  auto slot = [&](int f) {  // this lane's state in fork slot f
    return p.trie_ws + ((((size_t)blockIdx.x * PAIRS_PER_UNIT + g) * TRIE_SLOTS + f) * GROUP + l) * W;
  };
  const float C = 0x1p108f;
  const uint16_t* cl = codes + g * smax + 16 - l;
  // this unit's chunk of ranks, one per lane (trie_chunk <= 64); the chunk's
  // first haplotype sweeps in full, so every fork it restores was written
  // inside the chunk
  const int q0 = __builtin_amdgcn_readfirstlane(rt.q0), q1 = __builtin_amdgcn_readfirstlane(rt.q1);
  int order_v = 0, start_v = 0, len_v = 0;
  {
    const int qq = q0 + lane;
    order_v = qq < q1 ? p.trie_order[h0 + qq] : 0;
    start_v = qq < q1 && lane > 0 ? p.trie_start[h0 + qq] : 0;
    len_v = qq < q1 ? (int)(b.hap_off[order_v + 1] - b.hap_off[order_v]) : 0;
  }
  // suffix continuation: one adjoint sweep per read when the chunk holds two
  // or more haplotypes (S >= SFX_MIN guarantees that columns J+1..J+15 exist)
  const int S = (SFX && q1 - q0 >= 2) ? __builtin_amdgcn_readfirstlane(p.trie_sfx[k]) : 0;
  // this lane's workspace (lane-major, so that every access is an immediate
  // offset from one address): per row (Qb_i(J), Db_i(J+1)); the boundary
  // mass; then Q_i(J), M^_i(J) and D^_i(J) of the current haplotype
  float2* const sbase = reinterpret_cast<float2*>(p.sfx_ws + (size_t)blockIdx.x * SFX_WS);  // uniform
  const int slane = lane * (SFX_WS / 128);  // this lane's first float2
  float* const sstate = reinterpret_cast<float*>(sbase + slane + MAX_R32 + 2);
  if (S > 0) {
    {  // suffix codes of the first haplotype: slot k + 17 holds suffix column k
      const int64_t hb = b.hap_off[order_v], hend = b.hap_off[order_v + 1];
      const int64_t hs = __builtin_amdgcn_readfirstlane((int)(hend - hb)) - S;
      const int64_t base = __builtin_amdgcn_readfirstlane(hb) + hs;
      for (int kk = lane; kk < S + 34; kk += 64) {
        const int col = kk - 17;
        codes[kk] = (uint16_t)(col <= 0 ? SENT_PRE : (col > S ? SENT_POST : (uint32_t)b.hap_bytes[base + col - 1]));
      }
    }
    __syncthreads();
    const float F = sfx_adjoint<R, ANCH, U>(gd, bd, mu, be, rc, gam, gam_last, xx, codes, S, sbase + slane);
    // the group's lane 0 holds the boundary mass: every lane keeps its group's
    // value in its own slot (lane 15 forms the terminal)
    const int fb = __builtin_bit_cast(int, F);
    const float f0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(fb, 0));
    const float f1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(fb, 16));
    const float f2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(fb, 32));
    const float f3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(fb, 48));
    sbase[slane + MAX_R32] = make_float2(g == 0 ? f0 : g == 1 ? f1 : g == 2 ? f2 : f3, 0.f);
    count_groups(&p.ctr[CTR_SWEPT], rt.valid ? (unsigned long long)rt.n * (unsigned long long)S : 0ull);
    __syncthreads();  // codes reused by the haplotypes
  }
  for (int qi = 0; qi < q1 - q0; ++qi) {
    const int64_t hap = __builtin_amdgcn_readlane(order_v, qi);
    const int m = __builtin_amdgcn_readlane(len_v, qi);
    if (m == 0) {  // same haplotype for every group: uniform
      if (rt.valid && l == 0) out[b.out_off[k] + (rt.read - r0) * H + (hap - h0)] = -__builtin_inf();
      continue;
    }
    int st = __builtin_amdgcn_readlane(start_v, qi);
    // the sweep's last column: J = m - S with the suffix continuation, else m
    const int J = m - S;
    // fork slot of column st (a start without a slot sweeps in full), and the
    // first fork this sweep computes: forks in (st, flast] are saved after step
    // f + 1, inside the main loop
    const int flast = S ? J - 1 : m;
    const unsigned long long hit = __builtin_amdgcn_ballot_w64(fork_v == st);
    int rs = hit ? __builtin_ctzll(hit) : 0;
    if (st > 0 && (!hit || st > flast || st >= m)) st = 0;
    int fi = __builtin_popcountll(__builtin_amdgcn_ballot_w64(fork_v <= st));
    // a fork is saved only when a later haplotype of the chunk resumes from it
    auto next_fork = [&]() {
      while (fi < nfork && fork_at(fi) <= flast &&
             !__builtin_amdgcn_ballot_w64(start_v == fork_at(fi) && lane > qi && q0 + lane < q1))
        ++fi;
      return (fi < nfork && fork_at(fi) <= flast) ? fork_at(fi) + 1 : 0x7fffffff;
    };
    int fstep = next_fork();
    const int tend = J + 16, tmain = S ? J + 1 : m + 2;
    {
      GroupPair hp;
      hp.hap = hap;
      hp.m = m;
      hp.valid = 1;
      fill_codes(b, hp, codes, smax, tend + 18);
    }
    __syncthreads();  // codes visible

    float M[R], D[R], Q[R];
    float Plast = 0.f, Xprev = 0.f, acc = 0.f, rmax = 0.f, latch = 0.f;
    // the injection as the DPP moves' old value: a lane-0 register for it
    // (the pair-unit sweep's form) would push these instances past their
    // register budgets
    const float inj = ANCH ? 0.f : C;
    // Without the suffix continuation the P hand-off writes into a register
    // carried across steps: lane 0 of each row is never written by the
    // shift, so it keeps inj and the per-step move of inj is gone
    float Pcar = inj;
    auto main_step = [&](int tt) {
      const uint32_t code = cl[tt];
      float Pin;
      if constexpr (SFX) {
        Pin = dpp_shr1(inj, Plast);
      } else {
        Pcar = dpp_shr1(Pcar, Plast);
        Pin = Pcar;
      }
      float Xn = dpp_shr1(inj, Q[R - 1]);
      Plast = adiag_step<float, R, false>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code);
      Xprev = Xn;
      if (ANCH) rmax = fmaxf(rmax, Q[R - 1]); else acc += Plast;
    };
    auto tail_step = [&](int tt) {
      const uint32_t code = cl[tt];
      const float pinj = (!ANCH && tt - 1 > m) ? 0.f : inj;
      float Pin = dpp_shr1(pinj, Plast);
      float Xn = dpp_shr1(inj, Q[R - 1]);
      Plast = adiag_step<float, R, true>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code);
      Xprev = Xn;
      if (ANCH) {
        rmax = fmaxf(rmax, Q[R - 1]);
        if (tt - l - 1 == m) latch = Q[R - 1];
      } else {
        acc += Plast;
      }
    };
    int t;
    if (st == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) M[r] = D[r] = Q[r] = 0.f;
      {  // step 0 (column -1 for lane 0): the boundary row's first values
        const float inj0 = ANCH ? C : 0.f;
        float Pin = dpp_shr1(inj0, Plast);
        float Xn = dpp_shr1(inj0, Q[R - 1]);
        Plast = adiag_step<float, R, false>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, cl[0]);
        Xprev = Xn;
        if (ANCH) rmax = fmaxf(rmax, Q[R - 1]); else acc += Plast;
      }
      t = 1;
    } else {
      float v[W];
      const float4* src = reinterpret_cast<const float4*>(slot(rs));
#pragma unroll
      for (int i = 0; i < W / 4; ++i) {
        const float4 x = src[i];
        v[4 * i] = x.x;
        v[4 * i + 1] = x.y;
        v[4 * i + 2] = x.z;
        v[4 * i + 3] = x.w;
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        M[r] = v[3 * r];
        D[r] = v[3 * r + 1];
        Q[r] = v[3 * r + 2];
      }
      Plast = v[3 * R];
      Xprev = v[3 * R + 1];
      if (ANCH) rmax = v[3 * R + 2]; else acc = v[3 * R + 2];
      t = st + 2;
    }
    // main steps in segments that end at the fork steps (fstep <= m + 1 < tmain),
    // so that the unrolled loop itself carries no branch
    for (;;) {
      const int seg_end = fstep < tmain ? fstep : tmain - 1;  // last step of the segment
      for (; t + U - 1 <= seg_end; t += U) {
#pragma unroll
        for (int kk = 0; kk < U; ++kk) main_step(t + kk);
      }
      for (; t <= seg_end; ++t) main_step(t);
      if (seg_end != fstep) break;
      if (rt.valid) {  // the wavefront after step f + 1, in 16-byte stores
        float v[W];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          v[3 * r] = M[r];
          v[3 * r + 1] = D[r];
          v[3 * r + 2] = Q[r];
        }
        v[3 * R] = Plast;
        v[3 * R + 1] = Xprev;
        v[3 * R + 2] = ANCH ? rmax : acc;
#pragma unroll
        for (int i = 3 * R + 3; i < W; ++i) v[i] = 0.f;
        float4* dst = reinterpret_cast<float4*>(slot(fi));
#pragma unroll
        for (int i = 0; i < W / 4; ++i) dst[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
      }
      ++fi;
      fstep = next_fork();
    }
    // the tail, steps up to tend = J + 16 (lane 15 at column J last).  With the
    // suffix continuation the tail is steps J+1..J+16: columns <= J+15 < m, so
    // no sentinel codes and no latch; lane l is at column J after step
    // J + l + 1 and leaves Q_i(J), M^_i(J) and D^_i(J) in the workspace
    auto tail_snap = [&](int tt) {
      tail_step(tt);
      if (SFX && S && tt - l - 1 == J) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          sstate[r] = Q[r];
          sstate[MAX_R32 + r] = M[r];
          sstate[2 * MAX_R32 + r] = D[r];
        }
      }
    };
    for (; t + 1 <= tend; t += 2) {
      tail_snap(t);
      tail_snap(t + 1);
    }
    if (t <= tend) tail_snap(t++);
    double dot = 0.0;
    if (SFX && S) {
      // terminal of the suffix columns: the adjoint weights times that state.
      // The weights do not change between haplotypes; the empty asm keeps the
      // compiler from holding them in registers across the sweeps
      int so = slane;
      asm volatile("" : "+v"(so));
      // two rows at a time: the row constants of the next sweep stay live
      // here, so the loads are kept from all being issued at once
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float2 w = sbase[so + r];
        const float* x = reinterpret_cast<const float*>(sbase + so + MAX_R32 + 2);
        const float dn = fmaT(xx, x[2 * MAX_R32 + r], x[MAX_R32 + r]);  // D^_i(J+1), as the sweep forms it
        dot += (double)w.x * (double)x[r] + (double)w.y * (double)dn;
        if (r & 1) asm volatile("" ::: "memory");
      }
      dot = group_sum(dot) * 0x1p-100;
    }
    count_groups(&p.ctr[CTR_SWEPT], rt.valid ? (unsigned long long)rt.n * (unsigned long long)(J - st) : 0ull);
    if (l == GROUP - 1 && rt.valid) {
      const double lnk = p.read_ln32[rt.read] - CEXP32 * LN2;
      const int64_t oidx = b.out_off[k] + (rt.read - r0) * H + (hap - h0);
      double v;
      bool flag;
      if (ANCH && SFX) {
        // with the suffix continuation rmax covers columns <= J only; it bounds
        // nothing there: an unflagged latch >= 2^-100 is within 2^-228 of any
        // finite FP32 row value, far above the reference's 1e-300 clamp
        const double lt = S ? dot : (double)latch;
        flag = !(lt >= FLAG32_MIN) || !(rmax < __builtin_inff()) || !(lt < 0x1p128);
        const double lc = log((double)rmax) + LN_1EM300;
        const double ll = log(lt);
        v = (ll > lc ? ll : lc) + lnk;
      } else if (ANCH) {
        flag = !(latch >= FLAG32_MIN) || !(rmax < __builtin_inff());
        const double lc = log((double)rmax) + LN_1EM300;
        const double ll = log((double)latch);
        v = (ll > lc ? ll : lc) + lnk;
      } else if (SFX) {
        // the boundary injections of the suffix columns carry 2^CEXP32
        int so = slane;
        asm volatile("" : "+v"(so));
        const double tot = S ? (double)acc + dot + (double)sbase[so + MAX_R32].x * 0x1p8 : (double)acc;
        static_assert(CEXP32 - SFX_SEED_EXP == 8, "injection scale");
        flag = !(tot >= FLAG32_MIN) || !(tot < 0x1p128);
        v = log(tot) + lnk - log((double)(m + 1));
      } else {
        flag = !(acc >= FLAG32_MIN) || !(acc < __builtin_inff());
        v = log((double)acc) + lnk - log((double)(m + 1));
      }
      if (flag) {
        unsigned long long idx = atomicAdd(&p.ctr[CTR_F64], 1ull);
        atomicAdd(&p.ctr[CTR_FLAGGED], 1ull);
        if ((int64_t)idx < p.f64_cap) p.f64_list[idx] = PairRef{rt.read, hap, oidx};
        else atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
        out[oidx] = __builtin_nan("");
      } else {
        out[oidx] = v;
      }
    }
    __syncthreads();  // codes reused by the next haplotype
  }
}

// Waves per SIMD requested for the class kernels whose natural register
// allocation stops just short of the next occupancy step (e.g. 171 registers
// where 168 give 3 waves): the highest occupancy at which the compiler fits
// the instance without spilling, found by compiling every instance at 2..7
// waves; for the semiglobal prefix-sharing R = 11, 12 and 16 one step more,
// where the few spilled values (4-12) are reloaded only outside the sweep
// loops and the kernels measured 6-16% faster.  1 = no request (the hint
// would not raise the occupancy, and it changes the schedule).  Index: rows
// per lane R (0 unused).
constexpr int WPE_T[2][2][17] = {{{1, 1, 1, 1, 1, 1, 1, 1, 1, 4, 4, 4, 4, 1, 3, 3, 3},
                                  {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1}},
                                 {{1, 1, 6, 1, 5, 1, 1, 1, 1, 1, 3, 1, 1, 1, 1, 1, 1},
                                  {1, 1, 1, 1, 1, 1, 1, 1, 3, 1, 1, 1, 1, 1, 1, 1, 1}}};  // [anchored][suffix][R]
constexpr int WPE_P[2][2][17] = {{{1, 8, 8, 1, 1, 6, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
                                  {1, 8, 8, 8, 7, 6, 5, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1}},
                                 {{1, 8, 8, 1, 7, 6, 5, 1, 1, 4, 1, 1, 1, 3, 1, 1, 1},
                                  {1, 8, 8, 7, 7, 6, 5, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1}}};  // [anchored][unroll 4][R]
constexpr int wpe_t(bool anch, bool sfx, int r) { return WPE_T[anch][sfx][r]; }
constexpr int wpe_p(bool anch, int u, int r) { return WPE_P[anch][u == 4][r]; }

template <bool ANCH, int R, int U>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(wpe_p(ANCH, U, R)))) void k_fwd32(BatchDev b, PlanDev p, const QTab<float>* __restrict__ gtab,
                                              int smax, double* __restrict__ out) {
  const unsigned long long lo = (unsigned long long)p.unit_cnt[bucket_of(R, 0, b.nregions, 1)];
  const unsigned long long hi = (unsigned long long)p.unit_cnt[bucket_of(R - 1, 0, b.nregions, 1)];
  if (lo >= hi) return;  // no read of this class in the batch
  extern __shared__ __align__(16) unsigned char lds[];
  QTab<float>* tab = reinterpret_cast<QTab<float>*>(lds);
  uint16_t* codes = reinterpret_cast<uint16_t*>(lds + ((qtab_lds_bytes<float>() + 15) & ~size_t(15)));
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(gtab);
    uint32_t* dst = reinterpret_cast<uint32_t*>(tab);
    for (int i = threadIdx.x; i < (int)(qtab_lds_bytes<float>() / 4); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // the class's units [lo, hi), or with packed pairs its pairs [lo, hi) in
  // units of 4 that may span regions
  const unsigned long long nunits = p.pack_pairs ? (hi - lo + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT : hi - lo;
  for (;;) {
    unsigned long long u = 0;
    if (lane == 0) u = atomicAdd(&p.ctr[CTR_CLS + R], 1ull);
    u = first_lane(u);
    if (u >= nunits) break;
    GroupPair gp;
    if (p.pack_pairs) {
      const unsigned long long pi = lo + u * PAIRS_PER_UNIT + (lane >> 4);
      if (decode_pair32(b, p, (int64_t)pi, pi < hi, gp, out) < 0) continue;
    } else if (decode_unit32(b, p, lo + u, gp, out) != R) {
      continue;
    }
    unit32<R, ANCH, U>(b, p, tab, codes, smax, gp, out);
  }
}

// The prefix-sharing forward of class R over the read units of the plan (SFX:
// with the suffix continuation, a separate instantiation because its code
// raises the register count of some classes).
//
// part 0 takes every read unit of the class; with a split class, part 1 (the
// SFX instance) takes the units whose region has a shared suffix and whose
// chunk holds two or more haplotypes, part 2 (the prefix-only instance) the
// others, from a counter of its own.
template <bool ANCH, int R, int U, bool SFX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(wpe_t(ANCH, SFX, R)))) void k_fwd32t(
    BatchDev b, PlanDev p, const QTab<float>* __restrict__ gtab, int smax, double* __restrict__ out, int part) {
  const unsigned long long lo = (unsigned long long)p.unit_cnt[bucket_of(R, 0, b.nregions, 1)];
  const unsigned long long hi = (unsigned long long)p.unit_cnt[bucket_of(R - 1, 0, b.nregions, 1)];
  if (lo >= hi) return;
  extern __shared__ __align__(16) unsigned char lds[];
  QTab<float>* tab = reinterpret_cast<QTab<float>*>(lds);
  uint16_t* codes = reinterpret_cast<uint16_t*>(lds + ((qtab_lds_bytes<float>() + 15) & ~size_t(15)));
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(gtab);
    uint32_t* dst = reinterpret_cast<uint32_t*>(tab);
    for (int i = threadIdx.x; i < (int)(qtab_lds_bytes<float>() / 4); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  unsigned long long* next = &p.ctr[(part == 2 ? CTR_CLS2 : CTR_CLS) + R];
  for (;;) {
    unsigned long long u = 0;
    if (lane == 0) u = lo + atomicAdd(next, 1ull);
    u = first_lane(u);
    if (u >= hi) break;
    ReadTask rt;
    if (decode_read_unit(b, p, u, rt) != R) continue;
    if (part) {  // uniform: one region and chunk per unit
      const bool sfx_unit = p.trie_sfx[rt.region] > 0 && rt.q1 - rt.q0 >= 2;
      if (sfx_unit != (part == 1)) continue;
    }
    unit32_trie<R, ANCH, U, SFX>(b, p, tab, codes, smax, rt, out);
  }
}

// ---------------------------------------------------------------------------
// FP64: the FP32 layout (4 pairs per wave, 16-lane groups, DPP row_shr:1) with
// R64 = 4 rows per lane, i.e. 64-row tiles.  A tile's boundary row (the last
// row's P/Q streams) goes through a global workspace and is renormalised to
// 2^CEXP64 before the next tile, so any read length stays in range: a tile
// decays by at most 64 ln(4/x) nats along the insertion path.

// One tile's sweep of the FP64 forward: tile64 in vdhmm_device.h (shared
// with the tiled FP64 adjoint).

template <bool ANCH>
__device__ __forceinline__ void unit64(const BatchDev& b, const PlanDev& p, const QTab<double>* __restrict__ tab,
                                       uint16_t* codes, int smax, const GroupPair gp, double* __restrict__ ws,
                                       double* __restrict__ out) {
  constexpr int R = R64;
  const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
  const int tend = gp.m_max + 16;
  const int tmain = gp.m_min + 2;
  fill_codes(b, gp, codes, smax, tend + 18);
  const int np = gp.n + (ANCH ? 1 : 0);
  const int ntile_g = gp.valid ? (np + GROUP * R - 1) / (GROUP * R) : 0;
  const int ntiles = max(max(lane16(ntile_g, 0), lane16(ntile_g, 1)), max(lane16(ntile_g, 2), lane16(ntile_g, 3)));
  const int pad = ntiles * GROUP * R - np;  // reads are bottom-aligned: pass-through rows on top
  const int64_t rb = b.read_off[gp.read];
  const double gam = tab->gamma, xx = tab->ext;
  // boundary buffers: [buffer][stream][group][slot], slot = column + 17
  double* bufP[2];
  double* bufQ[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    bufP[k] = ws + ((size_t)(2 * k + 0) * PAIRS_PER_UNIT + g) * smax;
    bufQ[k] = ws + ((size_t)(2 * k + 1) * PAIRS_PER_UNIT + g) * smax;
  }
  int shift_total = 0;   // sum of tile renormalisation exponents (log2)
  double sc1 = 1.0, sc2 = 1.0;  // 2^s of this tile's injected boundary, as two factors
  double acc = 0.0, rmax = 0.0, latch = 0.0;
  const uint16_t* cl = codes + g * smax + 16 - l;
  __syncthreads();
  for (int tile = 0; tile < ntiles; ++tile) {
    const bool first = tile == 0, last = tile == ntiles - 1;
    // the read's last row (lane 15, register R-1 of the last tile) carries the
    // terminal weights; in earlier tiles that register is an interior row
    const double gam_last = (l == GROUP - 1 && last) ? (ANCH ? tab->gap : 0.0) : gam;
    double gd[R], bd[R], mu[R], be[R];
    uint32_t rc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int rel = tile * GROUP * R + l * R + r - pad;
      double vg = 0.0, vb = 0.0, vm = 0.0, ve = 0.0;
      uint32_t vc = NOMATCH;
      if (gp.valid) {
        if (rel < 0) {
          vm = ve = 1.0;
        } else if (ANCH && rel == 0) {
          vg = vb = 1.0;
        } else {
          const int kk = rel - (ANCH ? 1 : 0);
          int q = b.qual_bytes[rb + kk];
          q = q < 63 ? q : 63;
          vc = b.read_bytes[rb + kk];
          vg = tab->good[q];
          vb = tab->bad[q];
          vm = tab->mu[q];
          ve = tab->beta[q];
          if (kk == gp.n - 1) vm = ve = tab->cfin[q];
        }
      }
      gd[r] = vg; bd[r] = vb; mu[r] = vm; be[r] = ve; rc[r] = vc;
    }
    const double* inP = bufP[(tile + 1) & 1];
    const double* inQ = bufQ[(tile + 1) & 1];
    double* oP = bufP[tile & 1];
    double* oQ = bufQ[tile & 1];
    // 2^s is applied as two exact factors (s can exceed one factor's exponent
    // range), once per loaded boundary value
    double bmax = 0.0;
#define VD_TILE(F, L)                                                                                               \
  tile64<ANCH, F, L>(gp, cl, tmain, tend, gd, bd, mu, be, rc, gam, gam_last, xx, sc1, sc2, inP, inQ, oP, oQ, acc,  \
                     rmax, latch, bmax)
    if (first && last) VD_TILE(true, true);
    else if (first) VD_TILE(true, false);
    else if (last) VD_TILE(false, true);
    else VD_TILE(false, false);
#undef VD_TILE
    if (!last) {
      // renormalise the boundary row to 2^CEXP64.  A tile above the group's
      // first read row (a unit mate has more tiles) holds pass-through rows
      // only and hands its injection on exactly, unscaled, so the pair's
      // real tiles see the values and renormalisations of a unit of its own:
      // its result depends on the pair alone.
      int e = 0;
      if (bmax > 0.0) frexp(bmax, &e);
      const bool real_rows = tile >= ntiles - ntile_g;
      int s = (bmax > 0.0 && real_rows) ? CEXP64 - e : 0;
      s = s > 2040 ? 2040 : s;
      const int src = g * GROUP + GROUP - 1;
      const int s1 = s / 2, s2 = s - s1;
      sc1 = __shfl(ldexp(1.0, s1), src);
      sc2 = __shfl(ldexp(1.0, s2), src);
      shift_total -= __shfl(s, src);
      __threadfence();
      __syncthreads();
    }
  }
  if (l == GROUP - 1 && gp.valid) {
    const double lnk = (double)(p.read_exp[gp.read] - CEXP64 + shift_total) * LN2;
    double v;
    bool bad;
    if (ANCH) {
      // reference clamp ln(max(1e-300 * rowmax, L)), evaluated in log space
      bad = !(rmax > 0.0) || !(rmax < __builtin_inf());
      const double lc = log(rmax) + LN_1EM300;
      const double ll = log(latch);
      v = (ll > lc ? ll : lc) + lnk;
    } else {
      bad = !(acc > 0.0) || !(acc < __builtin_inf());
      v = log(acc) + lnk - log((double)(gp.m + 1));
    }
    if (bad) atomicAdd(&p.ctr[CTR_F64BAD], 1ull);
    out[gp.out] = v;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// FP64 work list binned by shape.  A unit of the FP64 kernel runs as many
// 64-row tiles as its tallest pair and as many steps as its longest
// haplotype, so the list is counting-sorted into bins of (tiles, haplotype
// length / 128), largest first (longest work first for the persistent grid),
// before the FP64 kernel takes its units in bin order.  A pair's FP64 result
// does not depend on its unit mates, so the order changes no value.

constexpr int F64_NT = 16;                 // tile-count classes 1 .. 16 (taller reads share the last)
constexpr int F64_NM = 32;                 // haplotype-length classes of 128 columns
constexpr int F64_NBIN = F64_NT * F64_NM;  // bins[0 .. NBIN): counts; bins[NBIN .. 2 NBIN): cursors
static_assert(F64_NBIN == F64_BINS_TOTAL && F64_NBIN <= 1024, "bin layout");

__device__ __forceinline__ int f64_bin(const BatchDev& b, const PairRef& pr, int anchored) {
  const int np = (int)(b.read_off[pr.read + 1] - b.read_off[pr.read]) + anchored;
  const int m = (int)(b.hap_off[pr.hap + 1] - b.hap_off[pr.hap]);
  int nt = (np + GROUP * R64 - 1) / (GROUP * R64);
  nt = nt < 1 ? 1 : (nt > F64_NT ? F64_NT : nt);
  int mb = m / 128;
  mb = mb >= F64_NM ? F64_NM - 1 : mb;
  return (F64_NT - nt) * F64_NM + (F64_NM - 1 - mb);
}

__global__ __launch_bounds__(256) void k_f64_hist(BatchDev b, PlanDev p, int anchored,
                                                  unsigned long long* __restrict__ bins) {
  __shared__ unsigned int loc[F64_NBIN];
  for (int i = threadIdx.x; i < F64_NBIN; i += blockDim.x) loc[i] = 0u;
  __syncthreads();
  unsigned long long count = p.ctr[CTR_F64];
  if ((int64_t)count > p.f64_cap) count = (unsigned long long)p.f64_cap;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (unsigned long long)gridDim.x * blockDim.x)
    atomicAdd(&loc[f64_bin(b, p.f64_list[i], anchored)], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < F64_NBIN; i += blockDim.x)
    if (loc[i]) atomicAdd(&bins[i], (unsigned long long)loc[i]);
}

// one block of F64_NBIN threads: the exclusive scan of the counts into the cursors
__global__ __launch_bounds__(F64_NBIN) void k_f64_scan(unsigned long long* __restrict__ bins) {
  __shared__ unsigned long long s[F64_NBIN];
  const int t = threadIdx.x;
  const unsigned long long own = bins[t];
  s[t] = own;
  __syncthreads();
  for (int d = 1; d < F64_NBIN; d <<= 1) {
    const unsigned long long v = t >= d ? s[t - d] : 0ull;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  bins[F64_NBIN + t] = s[t] - own;
}

__global__ __launch_bounds__(256) void k_f64_scatter(BatchDev b, PlanDev p, int anchored,
                                                     unsigned long long* __restrict__ bins,
                                                     PairRef* __restrict__ sorted) {
  unsigned long long count = p.ctr[CTR_F64];
  if ((int64_t)count > p.f64_cap) count = (unsigned long long)p.f64_cap;
  const int lane = threadIdx.x & 63;
  // one cursor atomic per distinct bin of a wave: the list is dominated by a
  // few bins, where one atomic per entry serialises on the same addresses
  for (unsigned long long w = (unsigned long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w < count;
       w += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long i = w + lane;
    const bool act = i < count;
    PairRef pr{};
    int bin = -1;
    if (act) {
      pr = p.f64_list[i];
      bin = f64_bin(b, pr, anchored);
    }
    unsigned long long pending = __builtin_amdgcn_ballot_w64(act), pos = 0;
    while (pending) {
      const int leader = __builtin_ctzll(pending);
      const int lb = __shfl(bin, leader);
      const unsigned long long same = __builtin_amdgcn_ballot_w64(bin == lb);
      unsigned long long start = 0;
      if (lane == leader) start = atomicAdd(&bins[F64_NBIN + lb], (unsigned long long)__builtin_popcountll(same));
      start = __shfl(start, leader);
      if (bin == lb) pos = start + (unsigned long long)__builtin_popcountll(same & ((1ull << lane) - 1ull));
      pending &= ~same;
    }
    if (act) sorted[pos] = pr;
  }
}

// 3 waves per SIMD, no spills (158-162 VGPRs with the boundary-row conveyor).
// At 4 waves the conveyor's registers spill inside the sweep and the kernel
// measured slower: 57.1 vs 55.6 ms on the 4096-region precision-stress batch
// (152,114 rescued pairs).
template <bool ANCH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_fwd64(BatchDev b, PlanDev p, const QTab<double>* __restrict__ gtab,
                                              int smax, double* __restrict__ wsall, double* __restrict__ out) {
  if (p.ctr[CTR_F64] == 0) return;  // nothing rescued in this pass
  extern __shared__ __align__(16) unsigned char lds[];
  QTab<double>* tab = reinterpret_cast<QTab<double>*>(lds);
  uint16_t* codes = reinterpret_cast<uint16_t*>(lds + ((qtab_lds_bytes<double>() + 15) & ~size_t(15)));
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(gtab);
    uint32_t* dst = reinterpret_cast<uint32_t*>(tab);
    for (int i = threadIdx.x; i < (int)(qtab_lds_bytes<double>() / 4); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  double* ws = wsall + (size_t)blockIdx.x * 4 * PAIRS_PER_UNIT * smax;
  const int lane = threadIdx.x & 63, g = lane >> 4;
  unsigned long long count = p.ctr[CTR_F64];
  if ((int64_t)count > p.f64_cap) count = (unsigned long long)p.f64_cap;
  const unsigned long long units = (count + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT;
  for (;;) {
    unsigned long long u = 0;
    if (lane == 0) u = atomicAdd(&p.ctr[CTR_UNIT64], 1ull);
    u = first_lane(u);
    if (u >= units) break;
    const unsigned long long idx = u * PAIRS_PER_UNIT + g;
    GroupPair gp;
    gp.valid = idx < count;
    const PairRef pr = p.f64_list[gp.valid ? idx : u * PAIRS_PER_UNIT];
    gp.read = pr.read;
    gp.hap = pr.hap;
    gp.out = pr.out;
    gp.n = (int)(b.read_off[gp.read + 1] - b.read_off[gp.read]);
    gp.m = (int)(b.hap_off[gp.hap + 1] - b.hap_off[gp.hap]);
    if (gp.valid && gp.m == 0) {
      if ((lane & 15) == 0) out[gp.out] = -__builtin_inf();
      gp.valid = 0;
    }
    const int mx0 = gp.valid ? gp.m : 0, mn0 = gp.valid ? gp.m : 0x7fffffff;
    const int mx = max(max(lane16(mx0, 0), lane16(mx0, 1)), max(lane16(mx0, 2), lane16(mx0, 3)));
    const int mn = min(min(lane16(mn0, 0), lane16(mn0, 1)), min(lane16(mn0, 2), lane16(mn0, 3)));
    if (mx == 0) continue;
    gp.m_max = mx;
    gp.m_min = mn;
    if (!gp.valid) gp.m = mx;
    unit64<ANCH>(b, p, tab, codes, smax, gp, ws, out);
  }
}

// One read against `count` haplotypes in FP64, for the scalar drop-ins: block
// u runs pairs 4u .. 4u+3 as one unit of the FP64 kernel, with no plan and no
// work list.  A pair's FP64 result depends on the pair alone, so these values
// equal the batched FP64 path's bit for bit.  p supplies read_exp and the
// counters only.
template <bool ANCH>
__global__ __launch_bounds__(64) void k_scalar64(BatchDev b, PlanDev p, const PairRef* __restrict__ pairs, int count,
                                                 const QTab<double>* __restrict__ gtab, int smax,
                                                 double* __restrict__ wsall, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char lds[];
  QTab<double>* tab = reinterpret_cast<QTab<double>*>(lds);
  uint16_t* codes = reinterpret_cast<uint16_t*>(lds + ((qtab_lds_bytes<double>() + 15) & ~size_t(15)));
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(gtab);
    uint32_t* dst = reinterpret_cast<uint32_t*>(tab);
    for (int i = threadIdx.x; i < (int)(qtab_lds_bytes<double>() / 4); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  double* ws = wsall + (size_t)blockIdx.x * 4 * PAIRS_PER_UNIT * smax;
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int idx = blockIdx.x * PAIRS_PER_UNIT + g;
  GroupPair gp;
  gp.valid = idx < count;
  const PairRef pr = pairs[gp.valid ? idx : blockIdx.x * PAIRS_PER_UNIT];
  gp.read = pr.read;
  gp.hap = pr.hap;
  gp.out = pr.out;
  gp.n = (int)(b.read_off[gp.read + 1] - b.read_off[gp.read]);
  gp.m = (int)(b.hap_off[gp.hap + 1] - b.hap_off[gp.hap]);
  if (gp.valid && gp.m == 0) {
    if ((lane & 15) == 0) out[gp.out] = -__builtin_inf();
    gp.valid = 0;
  }
  const int mx0 = gp.valid ? gp.m : 0, mn0 = gp.valid ? gp.m : 0x7fffffff;
  const int mx = max(max(lane16(mx0, 0), lane16(mx0, 1)), max(lane16(mx0, 2), lane16(mx0, 3)));
  const int mn = min(min(lane16(mn0, 0), lane16(mn0, 1)), min(lane16(mn0, 2), lane16(mn0, 3)));
  if (mx == 0) return;
  gp.m_max = mx;
  gp.m_min = mn;
  if (!gp.valid) gp.m = mx;
  unit64<ANCH>(b, p, tab, codes, smax, gp, ws, out);
}

hipError_t launch_scalar64(const BatchDev& b, const PlanDev& p, const PairRef* pairs, int count,
                           const QTab<double>* tab64, int anchored, int smax, double* ws, double* out,
                           hipStream_t st) {
  const dim3 grid((count + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT);
  if (anchored)
    hipLaunchKernelGGL(k_scalar64<true>, grid, dim3(64), lds_fwd64(smax), st, b, p, pairs, count, tab64, smax, ws,
                       out);
  else
    hipLaunchKernelGGL(k_scalar64<false>, grid, dim3(64), lds_fwd64(smax), st, b, p, pairs, count, tab64, smax, ws,
                       out);
  return hipGetLastError();
}

template __global__ void k_fwd64<false>(BatchDev, PlanDev, const QTab<double>*, int, double*, double*);
template __global__ void k_fwd64<true>(BatchDev, PlanDev, const QTab<double>*, int, double*, double*);



}  // namespace vd

// ---------------------------------------------------------------------------
// host-side launchers (called by vdhmm_engine.cpp)

namespace vd {

__global__ void k_zero_ctr(unsigned long long* ctr, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) ctr[i] = 0ull;
}

size_t lds_fwd32(int smax) {
  return ((qtab_lds_bytes<float>() + 15) & ~size_t(15)) + (size_t)PAIRS_PER_UNIT * smax * sizeof(uint16_t);
}
size_t lds_fwd64(int smax) {
  return ((qtab_lds_bytes<double>() + 15) & ~size_t(15)) + (size_t)PAIRS_PER_UNIT * smax * sizeof(uint16_t);
}

template <bool ANCH, int U>
static hipError_t occ_fwd32(int r, int smax, int* n) {
  switch (r) {
#define VD_OCC(RR) \
  case RR: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_fwd32<ANCH, RR, U>, 64, lds_fwd32(smax));
    VD_OCC(1) VD_OCC(2) VD_OCC(3) VD_OCC(4) VD_OCC(5) VD_OCC(6) VD_OCC(7) VD_OCC(8)
    VD_OCC(9) VD_OCC(10) VD_OCC(11) VD_OCC(12) VD_OCC(13) VD_OCC(14) VD_OCC(15) VD_OCC(16)
#undef VD_OCC
    default: *n = 1; return hipSuccess;
  }
}

// blocks per CU of the FP32 kernel of class r (1..16) with main-loop unroll u,
// or of the FP64 kernel (r = 0)
hipError_t occupancy(int anchored, int r, int u, int smax, int* blocks_per_cu) {
  if (r == 0)
    return anchored ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_fwd64<true>, 64, lds_fwd64(smax))
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_fwd64<false>, 64, lds_fwd64(smax));
  if (u == 4)
    return anchored ? occ_fwd32<true, 4>(r, smax, blocks_per_cu) : occ_fwd32<false, 4>(r, smax, blocks_per_cu);
  return anchored ? occ_fwd32<true, 2>(r, smax, blocks_per_cu) : occ_fwd32<false, 2>(r, smax, blocks_per_cu);
}

template <bool ANCH, int U>
static void launch_fwd32_r(int r, int blocks, size_t lds, hipStream_t st, const BatchDev& b, const PlanDev& p,
                           const QTab<float>* tab32, int smax, double* out) {
  switch (r) {
#define VD_L(RR)                                                                                         \
  case RR:                                                                                               \
    hipLaunchKernelGGL((k_fwd32<ANCH, RR, U>), dim3(blocks), dim3(64), lds, st, b, p, tab32, smax, out); \
    break;
    VD_L(1) VD_L(2) VD_L(3) VD_L(4) VD_L(5) VD_L(6) VD_L(7) VD_L(8)
    VD_L(9) VD_L(10) VD_L(11) VD_L(12) VD_L(13) VD_L(14) VD_L(15) VD_L(16)
#undef VD_L
    default: break;
  }
}

template <bool ANCH, bool SFX>
static hipError_t occ_fwd32t(int r, int smax, int* n) {
  switch (r) {
#define VD_OCC(RR) \
  case RR: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_fwd32t<ANCH, RR, 2, SFX>, 64, lds_fwd32(smax));
    VD_OCC(1) VD_OCC(2) VD_OCC(3) VD_OCC(4) VD_OCC(5) VD_OCC(6) VD_OCC(7) VD_OCC(8)
    VD_OCC(9) VD_OCC(10) VD_OCC(11) VD_OCC(12) VD_OCC(13) VD_OCC(14) VD_OCC(15) VD_OCC(16)
#undef VD_OCC
    default: *n = 1; return hipSuccess;
  }
}

hipError_t occupancy_trie(int anchored, int r, int sfx, int smax, int* blocks_per_cu) {
  if (sfx) return anchored ? occ_fwd32t<true, true>(r, smax, blocks_per_cu) : occ_fwd32t<false, true>(r, smax, blocks_per_cu);
  return anchored ? occ_fwd32t<true, false>(r, smax, blocks_per_cu) : occ_fwd32t<false, false>(r, smax, blocks_per_cu);
}

template <bool ANCH, bool SFX>
static void launch_fwd32t_r(int r, int blocks, size_t lds, hipStream_t st, const BatchDev& b, const PlanDev& p,
                            const QTab<float>* tab32, int smax, double* out, int part) {
  switch (r) {
#define VD_L(RR)                                                                                                  \
  case RR:                                                                                                        \
    hipLaunchKernelGGL((k_fwd32t<ANCH, RR, 2, SFX>), dim3(blocks), dim3(64), lds, st, b, p, tab32, smax, out, part); \
    break;
    VD_L(1) VD_L(2) VD_L(3) VD_L(4) VD_L(5) VD_L(6) VD_L(7) VD_L(8)
    VD_L(9) VD_L(10) VD_L(11) VD_L(12) VD_L(13) VD_L(14) VD_L(15) VD_L(16)
#undef VD_L
    default: break;
  }
}

void launch_plan(const BatchDev& b, const PlanDev& p, const int32_t* qE, const QTab<float>* tab32, int anchored,
                 int route64, double* out, hipStream_t st) {
  const int64_t nb = b.nregions * NCLS;
  hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, st, p.ctr, (int)CTR_COUNT);
  if (b.nreads > 0) {
    int64_t g = (b.nreads + 15) / 16;  // 4 waves of 64 per block, a read per 16-lane group
    g = g > 8192 ? 8192 : g;
    hipLaunchKernelGGL(k_prep_reads, dim3((unsigned)g), dim3(256), 0, st, b, qE, tab32, anchored, route64, p.read_cls,
                       p.read_exp, p.read_ln32);
  }
  if (b.nregions > 0) {
    int g = (int)(b.nregions > 65535 ? 65535 : b.nregions);
    hipLaunchKernelGGL(k_bucket, dim3(g), dim3(256), 0, st, b, p, out);
    if (p.trie) hipLaunchKernelGGL(k_hap_trie, dim3(g), dim3(64), 0, st, b, p);
  }
  const int64_t nt = (nb + SCAN_TILE - 1) / SCAN_TILE;
  int64_t* tsum = p.unit_cnt + nb + 1;
  if (nt > 0) hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nt), dim3(SCAN_THREADS), 0, st, p.unit_cnt, nb, tsum);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, tsum, nt, p.unit_cnt, nb, p.ctr);
  if (nt > 0) hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(SCAN_THREADS), 0, st, p.unit_cnt, nb, tsum);
  hipLaunchKernelGGL(k_unit_map, dim3(2048), dim3(256), 0, st, (const int64_t*)p.unit_cnt, nb,
                     (const unsigned long long*)p.ctr, p.unit_bucket);
}

hipError_t launch_pipeline(const BatchDev& b, const PlanDev& p, const int32_t* qE, const QTab<float>* tab32,
                           const QTab<double>* tab64, int anchored, int route64, int smax, const int* blocks32,
                           const int* blocks32s,
                           const int* unroll32, int blocks64, double* ws64, PairRef* f64_sorted,
                           unsigned long long* f64_bins, double* out, hipStream_t st0, hipEvent_t* ev,
                           hipEvent_t* ev_cls, const SideLane* lane2) {
  hipStream_t st = st0;
  if (ev) (void)hipEventRecord(ev[0], st);
  launch_plan(b, p, qE, tab32, anchored, route64, out, st);
  if (ev) (void)hipEventRecord(ev[1], st);
  if (!(route64 & 1)) {
    // per-class timing keeps one stream, so that each class is timed alone
    const bool two = lane2 && !ev_cls;
    const PlanDev& p0 = p;
    PlanDev p1 = p;
    if (two) {
      if (p1.trie_ws) p1.trie_ws += lane2->trie_half;
      if (p1.sfx_ws) p1.sfx_ws += lane2->sfx_half;
      (void)hipEventRecord(lane2->fork, st0);
      (void)hipStreamWaitEvent(lane2->side, lane2->fork, 0);
    }
    for (int r = MAX_R32; r >= 1; --r) {  // largest class first
      if (ev_cls) (void)hipEventRecord(ev_cls[r], st);
      const bool alt = two && (r & 1);
      const PlanDev& p = alt ? p1 : p0;
      st = alt ? lane2->side : st0;
      const bool u4 = unroll32[r] == 4;
      if ((p.trie >> r) & 1) {
        const size_t lds = lds_fwd32(smax);
        const bool sfx = (p.sfx >> r) & 1, split = sfx && ((p.split >> r) & 1);
        if (sfx) {  // all units, or (split) those of regions with a shared suffix
          const int blocks = split ? blocks32s[r] : blocks32[r];
          if (anchored) launch_fwd32t_r<true, true>(r, blocks, lds, st, b, p, tab32, smax, out, split ? 1 : 0);
          else launch_fwd32t_r<false, true>(r, blocks, lds, st, b, p, tab32, smax, out, split ? 1 : 0);
        }
        if (!sfx || split) {  // all units, or (split) the others
          if (anchored) launch_fwd32t_r<true, false>(r, blocks32[r], lds, st, b, p, tab32, smax, out, split ? 2 : 0);
          else launch_fwd32t_r<false, false>(r, blocks32[r], lds, st, b, p, tab32, smax, out, split ? 2 : 0);
        }
      } else if (anchored) {
        if (u4) launch_fwd32_r<true, 4>(r, blocks32[r], lds_fwd32(smax), st, b, p, tab32, smax, out);
        else launch_fwd32_r<true, 2>(r, blocks32[r], lds_fwd32(smax), st, b, p, tab32, smax, out);
      } else {
        if (u4) launch_fwd32_r<false, 4>(r, blocks32[r], lds_fwd32(smax), st, b, p, tab32, smax, out);
        else launch_fwd32_r<false, 2>(r, blocks32[r], lds_fwd32(smax), st, b, p, tab32, smax, out);
      }
    }
    st = st0;
    if (ev_cls) (void)hipEventRecord(ev_cls[0], st);
    if (two) {
      (void)hipEventRecord(lane2->join, lane2->side);
      (void)hipStreamWaitEvent(st0, lane2->join, 0);
    }
  }
  if (ev) (void)hipEventRecord(ev[2], st);
  PlanDev p64 = p;
  if (f64_sorted && f64_bins) {  // bin the FP64 list by shape (fill kernel, then count, scan, scatter)
    hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(256), 0, st, f64_bins, 2 * F64_NBIN);
    hipLaunchKernelGGL(k_f64_hist, dim3(1024), dim3(256), 0, st, b, p, anchored, f64_bins);
    hipLaunchKernelGGL(k_f64_scan, dim3(1), dim3(F64_NBIN), 0, st, f64_bins);
    hipLaunchKernelGGL(k_f64_scatter, dim3(1024), dim3(256), 0, st, b, p, anchored, f64_bins, f64_sorted);
    p64.f64_list = f64_sorted;
  }
  if (anchored)
    hipLaunchKernelGGL(k_fwd64<true>, dim3(blocks64), dim3(64), lds_fwd64(smax), st, b, p64, tab64, smax, ws64, out);
  else
    hipLaunchKernelGGL(k_fwd64<false>, dim3(blocks64), dim3(64), lds_fwd64(smax), st, b, p64, tab64, smax, ws64, out);
  if (ev) (void)hipEventRecord(ev[3], st);
  return hipGetLastError();
}

}  // namespace vd

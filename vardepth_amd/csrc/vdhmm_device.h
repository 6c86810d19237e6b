// Device helpers shared by the forward kernels (vdhmm_kernels.hip) and the
// adjoint kernels (vdhmm_grad.hip).
#pragma once
#include "vdhmm_internal.h"

#include <type_traits>

namespace vd {

#define LN2 0.69314718055994530942
#define LN_1EM300 (-690.77552789821368)  // ln(1e-300): the reference's terminal clamp

__device__ __forceinline__ float dpp_shr1(float old, float src) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, src),
                                         0x111, 0xF, 0xF, false));
}

struct I2 {
  int lo, hi;
};

__device__ __forceinline__ double dpp_shr1(double old, double src) {
  I2 o = __builtin_bit_cast(I2, old), s = __builtin_bit_cast(I2, src);
  I2 r;
  r.lo = __builtin_amdgcn_update_dpp(o.lo, s.lo, 0x111, 0xF, 0xF, false);
  r.hi = __builtin_amdgcn_update_dpp(o.hi, s.hi, 0x111, 0xF, 0xF, false);
  return __builtin_bit_cast(double, r);
}

// row_shl:1, the reverse hand-off (lane l reads lane l+1; `old` enters lane 15)
__device__ __forceinline__ float dpp_shl1(float old, float src) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, src),
                                         0x101, 0xF, 0xF, false));
}

__device__ __forceinline__ double dpp_shl1(double old, double src) {
  I2 o = __builtin_bit_cast(I2, old), s = __builtin_bit_cast(I2, src);
  I2 r;
  r.lo = __builtin_amdgcn_update_dpp(o.lo, s.lo, 0x101, 0xF, 0xF, false);
  r.hi = __builtin_amdgcn_update_dpp(o.hi, s.hi, 0x101, 0xF, 0xF, false);
  return __builtin_bit_cast(double, r);
}

// row_shr:1 for a lane-dependent boundary value: lane 0 of each 16-lane row
// reads zero (bound_ctrl) and every lane adds inj, which is zero except in
// lane 0.  The shift folds into the add (one v_add_f32 with a DPP operand,
// where dpp_shr1 takes a move of the boundary value and a DPP move), and the
// sum is exact: the other lanes add +0 to a value that is never -0.
__device__ __forceinline__ float dpp_shr1_add(float src, float inj) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, src), 0x111, 0xF, 0xF,
                                                               true)) +
         inj;
}

// A haplotype code read from LDS, held in a full register.  Neighbouring
// 16-bit codes are read as one wide LDS load; without the empty asm the
// compiler compares their halves with SDWA word selects, which issue at a
// fraction of the rate of plain compares.  The asm is not volatile, so equal
// codes still merge (the adjoint's recompute and backward share them).
__device__ __forceinline__ uint32_t code_reg(uint32_t c) {
  asm("" : "+v"(c));
  return c;
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------------------
// the rescaled anti-diagonal step for R rows of one lane

template <typename T, int R, bool TAIL>
__device__ __forceinline__ T adiag_step(T (&M)[R], T (&D)[R], T (&Q)[R], const T (&gd)[R], const T (&bd)[R],
                                        const T (&mu)[R], const T (&be)[R], const uint32_t (&rc)[R], T gam,
                                        T gam_last, T xx, T Pin, T Qin, uint32_t code) {
#pragma unroll
  for (int r = 0; r < R; ++r) D[r] = fmaT(xx, D[r], M[r]);
  const bool post = TAIL && code == SENT_POST;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    T qup = r == 0 ? Qin : Q[r - 1];
    T e = rc[r] == code ? gd[r] : bd[r];
    if (TAIL) e = post ? T(0) : e;
    M[r] = e * qup;
  }
  T Ih = Pin;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T P = fmaT(mu[r], Ih, M[r]);
    Q[r] = fmaT(r == R - 1 ? gam_last : gam, D[r], fmaT(be[r], Ih, M[r]));
    Ih = P;
  }
  return Ih;
}

// x summed over the 16 lanes of each group (every lane gets the sum), by
// ds_swizzle xor butterflies: the lane pattern is an immediate, so no lane
// index registers stay live around the call.
template <int XM>
__device__ __forceinline__ double swz_xor(double v) {
  I2 a = __builtin_bit_cast(I2, v);
  constexpr int pat = (XM << 10) | 0x1f;  // bitmask mode: and 0x1f, or 0, xor XM
  a.lo = __builtin_amdgcn_ds_swizzle(a.lo, pat);
  a.hi = __builtin_amdgcn_ds_swizzle(a.hi, pat);
  return __builtin_bit_cast(double, a);
}

__device__ __forceinline__ double group_sum(double v) {
  v += swz_xor<1>(v);
  v += swz_xor<2>(v);
  v += swz_xor<4>(v);
  v += swz_xor<8>(v);
  return v;
}

// The value of lane 16g (g = 0..3), read into scalar registers.  Cross-group
// exchanges use readlane with constant lanes rather than shuffles: shuffle
// lane addresses are loop-invariant VGPRs the compiler keeps live across the
// persistent loops.
__device__ __forceinline__ int lane16(int v, int g) { return __builtin_amdgcn_readlane(v, 16 * g); }
__device__ __forceinline__ int64_t lane16(int64_t v, int g) {
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)v, 16 * g);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), 16 * g);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ unsigned long long first_lane(unsigned long long v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Add v (the value of lane 16g of each group g) to a 64-bit counter: one
// atomic per wave.
__device__ __forceinline__ void count_groups(unsigned long long* ctr, unsigned long long v) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)v;
  const unsigned long long t =
      (unsigned long long)(lane16(w, 0) + lane16(w, 1) + lane16(w, 2) + lane16(w, 3));
  if (lane == 0 && t) atomicAdd(ctr, t);
}

// Per-lane description of the lane's group pair inside a unit.
struct GroupPair {
  int64_t read, hap, out;
  int n, m;        // read / haplotype length (m = unit max for invalid groups)
  int valid;
  int m_min, m_max;  // over the unit's valid groups
};

// Fill the unit's hap-code slots: slot k of group g holds the code of column
// k-17 (SENT_PRE for columns <= 0, SENT_POST past the end).
__device__ __forceinline__ void fill_codes(const BatchDev& b, const GroupPair& gp, uint16_t* codes, int smax,
                                           int nslots) {
  const int lane = threadIdx.x & 63;
  int64_t hb[PAIRS_PER_UNIT];
  int hm[PAIRS_PER_UNIT];
#pragma unroll
  for (int g = 0; g < PAIRS_PER_UNIT; ++g) {
    int64_t hoff = b.hap_off[gp.hap];
    hb[g] = lane16(hoff, g);
    hm[g] = lane16(gp.valid ? gp.m : 0, g);
  }
  for (int kk = lane; kk < nslots; kk += 64) {
    const int col = kk - 17;
#pragma unroll
    for (int g = 0; g < PAIRS_PER_UNIT; ++g) {
      uint32_t c = col <= 0 ? SENT_PRE : (col > hm[g] ? SENT_POST : (uint32_t)b.hap_bytes[hb[g] + col - 1]);
      codes[g * smax + kk] = (uint16_t)c;
    }
  }
}


// Work buckets.  Region-major (the forward): bucket (c, k) = k * NCLS + c, so a
// region's classes are processed together.  Class-major with the largest
// class first (the gradients): bucket (c, k) = (NCLS-1-c) * nregions + k, so
// after the scan the FP32 units of class c are the contiguous range
// [unit_cnt[bucket_of(c, 0)], unit_cnt[bucket_of(c - 1, 0)]) that one
// per-class kernel takes.
__host__ __device__ __forceinline__ int64_t bucket_of(int c, int64_t k, int64_t nregions, int class_major) {
  return class_major ? (int64_t)(NCLS - 1 - c) * nregions + k : k * NCLS + c;
}

// Packed plans (pack_pairs): pair pi of the class-major pair sequence (valid:
// below the class's end) -> the lane's group pair, with the unit's extent
// over its four groups; returns -1 when every group of the unit is empty.
// Pairs with an empty haplotype get -inf here and are marked invalid.
__device__ __forceinline__ int decode_pair32(const BatchDev& b, const PlanDev& p, int64_t pi, bool valid,
                                             GroupPair& gp, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  gp.valid = valid;
  gp.read = 0;
  gp.hap = 0;
  gp.out = 0;
  gp.n = 0;
  gp.m = 0;
  if (valid) {
    const int bkt = p.unit_bucket[pi];
    const int64_t j = pi - p.unit_cnt[bkt];
    const int64_t k = bkt % b.nregions;
    const int c = NCLS - 1 - (int)(bkt / b.nregions);
    const int64_t r0 = b.region_reads[k], h0 = b.region_haps[k];
    const int64_t H = b.region_haps[k + 1] - h0;
    const int32_t* co = p.class_off + k * (NCLS + 1);
    const int64_t s = j / H, h = j % H;
    gp.read = p.read_perm[r0 + co[c] + s];
    gp.hap = h0 + h;
    gp.out = b.out_off[k] + (gp.read - r0) * H + h;
    gp.n = (int)(b.read_off[gp.read + 1] - b.read_off[gp.read]);
    gp.m = (int)(b.hap_off[gp.hap + 1] - b.hap_off[gp.hap]);
    if (gp.m == 0) {
      if ((lane & 15) == 0) out[gp.out] = -__builtin_inf();
      gp.valid = 0;
    }
  }
  const int mx0 = gp.valid ? gp.m : 0, mn0 = gp.valid ? gp.m : 0x7fffffff;
  const int mx = max(max(lane16(mx0, 0), lane16(mx0, 1)), max(lane16(mx0, 2), lane16(mx0, 3)));
  const int mn = min(min(lane16(mn0, 0), lane16(mn0, 1)), min(lane16(mn0, 2), lane16(mn0, 3)));
  if (mx == 0) return -1;
  gp.m_max = mx;
  gp.m_min = mn;
  if (!gp.valid) gp.m = mx;  // an idle group: read 0, no rows, nothing written
  return 0;
}

// FP32 unit u -> the lane's group pair; returns the unit's class (rows per
// lane), or -1 when every group of the unit is empty.  Pairs with an empty
// haplotype get -inf here and are marked invalid.
__device__ __forceinline__ int decode_unit32(const BatchDev& b, const PlanDev& p, unsigned long long u,
                                             GroupPair& gp, double* __restrict__ out) {
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int bkt = p.unit_bucket[u];
  const int64_t j = (int64_t)u - p.unit_cnt[bkt];
  const int64_t k = p.class_major ? bkt % b.nregions : bkt / NCLS;
  const int c = p.class_major ? NCLS - 1 - (int)(bkt / b.nregions) : bkt % NCLS;
  const int64_t r0 = b.region_reads[k], h0 = b.region_haps[k];
  const int64_t H = b.region_haps[k + 1] - h0;
  const int32_t* co = p.class_off + k * (NCLS + 1);
  const int64_t cnt = co[c + 1] - co[c];
  const int64_t pi = j * PAIRS_PER_UNIT + g;
  gp.valid = pi < cnt * H;
  const int64_t s = gp.valid ? pi / H : 0, h = gp.valid ? pi % H : 0;
  gp.read = p.read_perm[r0 + co[c] + s];
  gp.hap = h0 + h;
  gp.out = b.out_off[k] + (gp.read - r0) * H + h;
  gp.n = (int)(b.read_off[gp.read + 1] - b.read_off[gp.read]);
  gp.m = (int)(b.hap_off[gp.hap + 1] - b.hap_off[gp.hap]);
  if (gp.valid && gp.m == 0) {
    if ((lane & 15) == 0) out[gp.out] = -__builtin_inf();
    gp.valid = 0;
  }
  const int mx0 = gp.valid ? gp.m : 0, mn0 = gp.valid ? gp.m : 0x7fffffff;
  const int mx = max(max(lane16(mx0, 0), lane16(mx0, 1)), max(lane16(mx0, 2), lane16(mx0, 3)));
  const int mn = min(min(lane16(mn0, 0), lane16(mn0, 1)), min(lane16(mn0, 2), lane16(mn0, 3)));
  if (mx == 0) return -1;
  gp.m_max = mx;
  gp.m_min = mn;
  if (!gp.valid) gp.m = mx;
  return c;
}

// ---------------------------------------------------------------------------
// FP64 64-row tiles (R64 rows per lane)

// One tile's sweep of the FP64 forward.  FIRST: the boundary row is the
// model's start (free starts, or the anchored origin); otherwise it comes
// from the previous tile's last row, pre-scaled by that tile's
// renormalisation as it is loaded (16 columns per vector load, one chunk
// ahead, handed down the lanes by DPP row_shl:1).  LAST: the tile holds the
// read's last row and accumulates the terminal; otherwise lane 15 stores the
// tile's last row and its maximum.  Steps before the shortest haplotype's end
// run without the end-of-haplotype test; both loops advance two steps per
// iteration so the loop-carried state stays in place.
template <bool ANCH, bool FIRST, bool LAST>
__device__ __forceinline__ void tile64(const GroupPair& gp, const uint16_t* cl, int tmain, int tend,
                                       const double (&gd)[R64], const double (&bd)[R64], const double (&mu)[R64],
                                       const double (&be)[R64], const uint32_t (&rc)[R64], double gam,
                                       double gam_last, double xx, double sc1, double sc2,
                                       const double* __restrict__ inP,
                                       const double* __restrict__ inQ, double* __restrict__ oP,
                                       double* __restrict__ oQ, double& acc, double& rmax, double& latch,
                                       double& bmax) {
  constexpr int R = R64;
  const int l = threadIdx.x & 15;
  double M[R], D[R], Q[R];
#pragma unroll
  for (int r = 0; r < R; ++r) M[r] = D[r] = Q[r] = 0.0;
  double Plast = 0.0, Xprev = 0.0;
  const double C = 0x1p500;
  double bP = 0.0, bQ = 0.0, nP = 0.0, nQ = 0.0;
  if (!FIRST) {
    if (l - 1 <= gp.m_max) {
      bP = inP[l + 16] * sc1 * sc2;
      bQ = inQ[l + 16] * sc1 * sc2;
    }
    if (l + 15 <= gp.m_max) {
      nP = inP[l + 32] * sc1 * sc2;
      nQ = inQ[l + 32] * sc1 * sc2;
    }
  }
  auto step = [&](int t, auto tail) {
    constexpr bool TAIL = decltype(tail)::value;
    const uint32_t code = cl[t];
    double pinj, qinj;
    if (FIRST) {
      if (ANCH) {
        pinj = qinj = t == 0 ? C : 0.0;
      } else {
        pinj = (t == 0 || t - 1 > gp.m) ? 0.0 : C;
        qinj = t == 0 ? 0.0 : C;
      }
    } else {
      if (t > 0 && (t & 15) == 0) {
        bP = nP;
        bQ = nQ;
        nP = nQ = 0.0;
        if (t + l + 15 <= gp.m_max) {
          nP = inP[t + l + 32] * sc1 * sc2;
          nQ = inQ[t + l + 32] * sc1 * sc2;
        }
      }
      pinj = bP;
      qinj = bQ;
      bP = dpp_shl1(0.0, bP);
      bQ = dpp_shl1(0.0, bQ);
    }
    const double Pin = dpp_shr1(pinj, Plast);
    const double Xn = dpp_shr1(qinj, Q[R - 1]);
    Plast = adiag_step<double, R, TAIL>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code);
    Xprev = Xn;
    if (LAST) {
      if (ANCH) {
        rmax = fmax(rmax, Q[R - 1]);
        latch = (t - l - 1 == gp.m) ? Q[R - 1] : latch;
      } else {
        acc += Plast;
      }
    } else {
      if (l == GROUP - 1) {
        oP[t + 1] = Plast;  // column t-16 -> slot t+1
        oQ[t + 1] = Q[R - 1];
      }
      bmax = fmax(bmax, fmax(Plast, Q[R - 1]));
    }
  };
  using head_t = std::integral_constant<bool, false>;
  using tail_t = std::integral_constant<bool, true>;
  int t = 0;
  for (; t + 1 < tmain; t += 2) {
    step(t, head_t{});
    step(t + 1, head_t{});
  }
  if (t < tmain) step(t++, head_t{});
  for (; t + 1 <= tend; t += 2) {
    step(t, tail_t{});
    step(t + 1, tail_t{});
  }
  if (t <= tend) step(t, tail_t{});
}

}  // namespace vd

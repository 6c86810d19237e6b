// Adjoint (reverse-mode) kernels: gradients of sum_p w_p ln L_p with respect
// to the gap open g, the gap continuation x and the 64-bin error table.
//
// There is no counterpart in the reference; the model is the one the forward
// kernels evaluate (batch_hmm.cpp:31-70).  A unit is the forward kernel's
// unit (4 pairs of one region and row class, 16 lanes per pair, R rows per
// lane), processed by one wave in three sweeps:
//
// 1. forward over all anti-diagonal steps, writing the lane state (M^, D^, Q
//    of its R rows and the two DPP carries) to a per-wave checkpoint area
//    every GK steps, and computing each pair's terminal;
// 2. backward over the steps in reverse, block by block: a block of GK steps
//    is recomputed from its checkpoint with the per-step values the adjoint
//    needs (each row's diagonal source Q_{i-1}(j-1), its D^, the lane's
//    injected P) kept in registers, then swept backwards;
// 3. per row, the accumulated adjoint sums are turned into gradient terms
//    through tables of partial derivatives of the scaled row constants.
//
// The adjoint recurrence is the transpose of the forward one.  With
// Qb = dL/dQ, Pb = dL/dP, Mb = dL/dM^, Db = dL/dD^ at cell (i, j):
//   Qb_i(j) = e'_{i+1}(j+1) Mb_{i+1}(j+1)
//   Pb_i(j) = mu_{i+1} Pb_{i+1}(j) + beta_{i+1} Qb_{i+1}(j)
//   Db_i(j) = gamma_i Qb_i(j) + x Db_i(j+1)
//   Mb_i(j) = Pb_i(j) + Qb_i(j) + Db_i(j+1)
// so lane l receives its carries from lane l+1 through DPP row_shl:1, whose
// `old` operand injects the terminal seeds into lane 15.  The seed is
// w / terminal, which makes every sum below a derivative of w ln L:
//   S_good_i = sum_match Mb Q_{i-1}(j-1)   S_bad_i = sum_mismatch (same)
//   S_mu_i   = sum Pb_i I^_i               S_beta_i = sum Qb_i I^_i
//   T_gamma  = sum Qb D^ (gamma rows)       T_x = sum Db_i(j+1) D^_i(j)
#include "vdhmm_device.h"

namespace vd {

constexpr int GK = 4;        // checkpoint interval (anti-diagonal steps)
constexpr int CKW_MAX = 3 * MAX_R32 + 2;

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

template <typename T>
struct GradScale;
template <>
struct GradScale<float> {
  static constexpr int CEXP = CEXP32;
};
template <>
struct GradScale<double> {
  static constexpr int CEXP = CEXPG64;
};

// One forward step that also records each row's diagonal source (qu) and new
// D^ (dd).  Emission is zero past the haplotype end at every step (codes past
// the end only occur in the tail, as in the forward kernels).
template <typename T, int R>
__device__ __forceinline__ T gstep(T (&M)[R], T (&D)[R], T (&Q)[R], const T (&gd)[R], const T (&bd)[R],
                                   const T (&mu)[R], const T (&be)[R], const uint32_t (&rc)[R], T gam, T gam_last,
                                   T xx, T Pin, T Qin, uint32_t code, T (&qu)[R], T (&dd)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    D[r] = fmaT(xx, D[r], M[r]);
    dd[r] = D[r];
  }
  const bool post = code == SENT_POST;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const T qup = r == 0 ? Qin : Q[r - 1];
    T e = rc[r] == code ? gd[r] : bd[r];
    e = post ? T(0) : e;
    qu[r] = qup;
    M[r] = e * qup;
  }
  T Ih = Pin;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const T P = fmaT(mu[r], Ih, M[r]);
    Q[r] = fmaT(r == R - 1 ? gam_last : gam, D[r], fmaT(be[r], Ih, M[r]));
    Ih = P;
  }
  return Ih;
}

// Injected boundary values of the anti-diagonal sweep (lane 0 of each group).
template <typename T, bool ANCH>
__device__ __forceinline__ void injections(int t, int m, T C, T& pinj, T& qinj) {
  if (ANCH) {
    pinj = qinj = t == 0 ? C : T(0);
  } else {
    pinj = (t == 0 || t - 1 > m) ? T(0) : C;
    qinj = t == 0 ? T(0) : C;
  }
}

template <typename T, int R, bool ANCH>
__device__ __forceinline__ void unit_grad(const BatchDev& b, const PlanDev& p, const QTab<T>* __restrict__ tab,
                                          const GradTab* __restrict__ gt, uint16_t* codes, int smax,
                                          const GroupPair gp, T* __restrict__ ck, const double* __restrict__ weights,
                                          double* __restrict__ out, double& dg, double& dx, double* lds_de,
                                          bool rescue) {
  constexpr int CKW = 3 * R + 2;
  const int lane = threadIdx.x & 63, l = lane & 15;
  const int tend = gp.m_max + 16;
  fill_codes(b, gp, codes, smax, tend + 18);
  T gd[R], bd[R], mu[R], be[R];
  uint32_t rc[R];
  const int np = gp.n + (ANCH ? 1 : 0);
  const int pad = GROUP * R - np;
  const int64_t rb = b.read_off[gp.read];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rel = l * R + r - pad;
    T vg = 0, vb = 0, vm = 0, ve = 0;
    uint32_t vc = NOMATCH;
    if (gp.valid) {
      if (rel < 0) {
        vm = ve = 1;
      } else if (ANCH && rel == 0) {
        vg = vb = 1;
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
  const T gam = tab->gamma, xx = tab->ext;
  const T gam_last = l == GROUP - 1 ? (ANCH ? tab->gap : T(0)) : gam;
  const T C = (T)ldexp(1.0, GradScale<T>::CEXP);
  const uint16_t* cl = codes + (lane >> 4) * smax + 16 - l;
  __syncthreads();

  // ---- sweep 1: forward with checkpoints
  T M[R], D[R], Q[R], qu[R], dd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) M[r] = D[r] = Q[r] = 0;
  T Plast = 0, Xprev = 0, acc = 0, rmax = 0, latch = 0;
  for (int t = 0; t <= tend; ++t) {
    if (t % GK == 0) {
      T* c = ck + (size_t)(t / GK) * CKW * 64 + lane;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        c[(size_t)r * 64] = M[r];
        c[(size_t)(R + r) * 64] = D[r];
        c[(size_t)(2 * R + r) * 64] = Q[r];
      }
      c[(size_t)(3 * R) * 64] = Plast;
      c[(size_t)(3 * R + 1) * 64] = Xprev;
    }
    T pinj, qinj;
    injections<T, ANCH>(t, gp.m, C, pinj, qinj);
    const T Pin = dpp_shr1(pinj, Plast);
    const T Xn = dpp_shr1(qinj, Q[R - 1]);
    Plast = gstep<T, R>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, cl[t], qu, dd);
    Xprev = Xn;
    if (ANCH) {
      rmax = fmax(rmax, Q[R - 1]);
      if (t - l - 1 == gp.m) latch = Q[R - 1];
    } else {
      acc += Plast;
    }
  }

  // ---- terminal, flag and seed (lane 15 of each group holds the pair's terminal)
  const T term = ANCH ? latch : acc;
  bool flag = false;
  double seed = 0.0;
  if (l == GROUP - 1 && gp.valid) {
    const double lnk = (double)(p.read_exp[gp.read] - GradScale<T>::CEXP) * LN2;
    double v;
    if (ANCH) {
      const double lc = log((double)rmax) + LN_1EM300;
      const double ll = log((double)latch);
      v = (ll > lc ? ll : lc) + lnk;
    } else {
      v = log((double)acc) + lnk - log((double)(gp.m + 1));
    }
    if (rescue) {
      flag = !((double)term > 0.0) || !((double)term < __builtin_inf());
      if (flag) atomicAdd(&p.ctr[CTR_F64BAD], 1ull);
    } else {
      flag = !(term >= (T)FLAG32_MIN) || !((double)term < __builtin_inf());
      if (flag) {
        unsigned long long idx = atomicAdd(&p.ctr[CTR_F64], 1ull);
        atomicAdd(&p.ctr[CTR_FLAGGED], 1ull);
        if ((int64_t)idx < p.f64_cap) p.f64_list[idx] = PairRef{gp.read, gp.hap, gp.out};
        else atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
      }
    }
    out[gp.out] = flag && !rescue ? __builtin_nan("") : v;
    const double w = weights ? weights[gp.out] : 1.0;
    seed = flag ? 0.0 : w / (double)term;
  }
  const int src = (lane & ~15) + 15;
  seed = __shfl(seed, src);
  flag = __shfl((int)flag, src) != 0;
  const T tseed = (T)seed;

  // ---- sweep 2: backward, block by block
  T EMp[R], Dbr[R], Am[R], Ax[R], Smu[R], Sbe[R];
#pragma unroll
  for (int r = 0; r < R; ++r) EMp[r] = Dbr[r] = Am[r] = Ax[r] = Smu[r] = Sbe[r] = 0;
  T Tg = 0, Tlast = 0, Txd = 0, Ilast = 0;
  // the anchored terminal of a pair whose haplotype is the unit's longest
  // enters before the first backward step
  T Eprev = (ANCH && l == GROUP - 1 && tend - 16 == gp.m) ? tseed : T(0);
  const T Pseed = ANCH ? T(0) : tseed;
  const int nblk = tend / GK + 1;
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int t0 = blk * GK;
    const int cnt = min(GK, tend + 1 - t0);
    {
      const T* c = ck + (size_t)blk * CKW * 64 + lane;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        M[r] = c[(size_t)r * 64];
        D[r] = c[(size_t)(R + r) * 64];
        Q[r] = c[(size_t)(2 * R + r) * 64];
      }
      Plast = c[(size_t)(3 * R) * 64];
      Xprev = c[(size_t)(3 * R + 1) * 64];
    }
    T sQ[GK][R], sD[GK][R], sPin[GK];
#pragma unroll
    for (int s = 0; s < GK; ++s) {
      if (s < cnt) {
        const int t = t0 + s;
        T pinj, qinj;
        injections<T, ANCH>(t, gp.m, C, pinj, qinj);
        const T Pin = dpp_shr1(pinj, Plast);
        const T Xn = dpp_shr1(qinj, Q[R - 1]);
        Plast = gstep<T, R>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, cl[t], sQ[s], sD[s]);
        Xprev = Xn;
        sPin[s] = Pin;
      }
    }
#pragma unroll
    for (int s = GK - 1; s >= 0; --s) {
      if (s < cnt) {
        const int t = t0 + s;
        const uint32_t code = cl[t];
        const bool post = code == SENT_POST;
        const T Pbin = dpp_shl1(Pseed, Ilast);
        const T Qbin = Eprev;
        const T einj = (ANCH && t - 17 == gp.m) ? tseed : T(0);
        const T En = dpp_shl1(einj, EMp[0]);
        T Qb[R], e[R], Iv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) Qb[r] = r == R - 1 ? Qbin : EMp[r + 1];
        T Ih = sPin[s];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          T er = rc[r] == code ? gd[r] : bd[r];
          e[r] = post ? T(0) : er;
          Iv[r] = Ih;
          Ih = fmaT(mu[r], Ih, e[r] * sQ[s][r]);
        }
        T Pb = Pbin;
#pragma unroll
        for (int r = R - 1; r >= 0; --r) {
          const T gr = r == R - 1 ? gam_last : gam;
          const T Dbar = fmaT(xx, Dbr[r], gr * Qb[r]);
          const T Mbar = Pb + Qb[r] + Dbr[r];
          const T mq = Mbar * sQ[s][r];
          const bool mt = rc[r] == code;
          Am[r] += mt ? mq : T(0);
          Ax[r] += (mt || post) ? T(0) : mq;
          Smu[r] = fmaT(Pb, Iv[r], Smu[r]);
          Sbe[r] = fmaT(Qb[r], Iv[r], Sbe[r]);
          const T dq = Qb[r] * sD[s][r];
          if (r == R - 1) Tlast += dq; else Tg += dq;
          Txd = fmaT(Dbr[r], sD[s][r], Txd);
          const T Ibar = fmaT(be[r], Qb[r], mu[r] * Pb);
          Dbr[r] = Dbar;
          EMp[r] = e[r] * Mbar;
          Pb = Ibar;
        }
        Ilast = Pb;
        Eprev = En;
      }
    }
  }

  // ---- sweep 3: gradient terms of the pair's rows
  if (gp.valid && !flag) {
    double lg = 0.0, lx = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int rel = l * R + r - pad;
      const int kk = rel - (ANCH ? 1 : 0);
      if (kk >= 0) {
        int q = b.qual_bytes[rb + kk];
        q = q < 63 ? q : 63;
        const double am = Am[r], ax = Ax[r], smu = Smu[r], sbe = Sbe[r];
        atomicAdd(&lds_de[q], am * gt->dgood_de[q] + ax * gt->dbad_de[q]);
        lg += am * gt->dgood_dg[q] + ax * gt->dbad_dg[q];
        if (kk == gp.n - 1) {
          lg += (smu + sbe) * gt->dcfin_dg[q];
        } else {
          lx += smu * gt->dmu_dx[q] + sbe * gt->dbeta_dx[q];
          lg += sbe * gt->dbeta_dg[q];
        }
      }
    }
    // register R-1 of lane 15 is the read's last row (gamma_last); every other
    // row uses gamma
    const double tg = (double)Tg + (l == GROUP - 1 ? 0.0 : (double)Tlast);
    const double tl = l == GROUP - 1 ? (double)Tlast : 0.0;
    lg += tg * gt->dgam_dg + (ANCH ? tl : 0.0);
    lx += tg * gt->dgam_dx + (double)Txd;
    dg += lg;
    dx += lx;
  }
  __syncthreads();  // codes reused by the next unit
}

__device__ __forceinline__ void grad_flush(double dg, double dx, const double* lds_de, double* __restrict__ gpart) {
  for (int o = 32; o > 0; o >>= 1) {
    dg += __shfl_xor(dg, o);
    dx += __shfl_xor(dx, o);
  }
  __syncthreads();
  double* gp = gpart + (size_t)blockIdx.x * NGRAD;
  if (threadIdx.x == 0) {
    gp[0] = dg;
    gp[1] = dx;
  }
  for (int q = threadIdx.x; q < 64; q += blockDim.x) gp[2 + q] = lds_de[q];
}

template <bool ANCH>
__global__ __launch_bounds__(64) void k_grad32(BatchDev b, PlanDev p, const QTab<float>* __restrict__ gtab,
                                               const GradTab* __restrict__ gt, int smax, float* __restrict__ ckall,
                                               size_t ck_stride, const double* __restrict__ weights,
                                               double* __restrict__ out, double* __restrict__ gpart) {
  extern __shared__ __align__(16) unsigned char lds[];
  double* lds_de = reinterpret_cast<double*>(lds);
  QTab<float>* tab = reinterpret_cast<QTab<float>*>(lds + 64 * sizeof(double));
  uint16_t* codes =
      reinterpret_cast<uint16_t*>(lds + 64 * sizeof(double) + ((sizeof(QTab<float>) + 15) & ~size_t(15)));
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(gtab);
    uint32_t* dst = reinterpret_cast<uint32_t*>(tab);
    for (int i = threadIdx.x; i < (int)(sizeof(QTab<float>) / 4); i += blockDim.x) dst[i] = src[i];
    for (int q = threadIdx.x; q < 64; q += blockDim.x) lds_de[q] = 0.0;
  }
  __syncthreads();
  float* ck = ckall + (size_t)blockIdx.x * ck_stride;
  const int lane = threadIdx.x & 63;
  double dg = 0.0, dx = 0.0;
  const unsigned long long total = p.ctr[CTR_NUNITS];
  for (;;) {
    unsigned long long u = 0;
    if (lane == 0) u = atomicAdd(&p.ctr[CTR_UNIT32], 1ull);
    u = __shfl(u, 0);
    if (u >= total) break;
    GroupPair gp;
    const int c = decode_unit32(b, p, u, gp, out);
    if (c < 0) continue;
    switch (c) {
#define VD_CASE(RR)                                                                                        \
  case RR:                                                                                                 \
    unit_grad<float, RR, ANCH>(b, p, tab, gt, codes, smax, gp, ck, weights, out, dg, dx, lds_de, false); \
    break;
      VD_CASE(1) VD_CASE(2) VD_CASE(3) VD_CASE(4) VD_CASE(5) VD_CASE(6) VD_CASE(7) VD_CASE(8)
      VD_CASE(9) VD_CASE(10) VD_CASE(11) VD_CASE(12) VD_CASE(13) VD_CASE(14) VD_CASE(15) VD_CASE(16)
#undef VD_CASE
      default: break;
    }
  }
  grad_flush(dg, dx, lds_de, gpart);
}

// FP64 gradients of the pairs the FP32 pass flagged (reads of at most 256
// rows: the host rejects longer ones for gradients), 4 listed pairs per unit.
template <bool ANCH>
__global__ __launch_bounds__(64) void k_grad64(BatchDev b, PlanDev p, const QTab<double>* __restrict__ gtab,
                                               const GradTab* __restrict__ gt, int smax, double* __restrict__ ckall,
                                               size_t ck_stride, const double* __restrict__ weights,
                                               double* __restrict__ out, double* __restrict__ gpart) {
  extern __shared__ __align__(16) unsigned char lds[];
  double* lds_de = reinterpret_cast<double*>(lds);
  QTab<double>* tab = reinterpret_cast<QTab<double>*>(lds + 64 * sizeof(double));
  uint16_t* codes =
      reinterpret_cast<uint16_t*>(lds + 64 * sizeof(double) + ((sizeof(QTab<double>) + 15) & ~size_t(15)));
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(gtab);
    uint32_t* dst = reinterpret_cast<uint32_t*>(tab);
    for (int i = threadIdx.x; i < (int)(sizeof(QTab<double>) / 4); i += blockDim.x) dst[i] = src[i];
    for (int q = threadIdx.x; q < 64; q += blockDim.x) lds_de[q] = 0.0;
  }
  __syncthreads();
  double* ck = ckall + (size_t)blockIdx.x * ck_stride;
  const int lane = threadIdx.x & 63, g = lane >> 4;
  double dg = 0.0, dx = 0.0;
  unsigned long long count = p.ctr[CTR_F64];
  if ((int64_t)count > p.f64_cap) count = (unsigned long long)p.f64_cap;
  const unsigned long long units = (count + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT;
  for (;;) {
    unsigned long long u = 0;
    if (lane == 0) u = atomicAdd(&p.ctr[CTR_UNIT64], 1ull);
    u = __shfl(u, 0);
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
    if (gp.valid && (gp.m == 0 || gp.n + (ANCH ? 1 : 0) > GROUP * MAX_R32)) gp.valid = 0;
    int mx = gp.valid ? gp.m : 0, mn = gp.valid ? gp.m : 0x7fffffff;
    mx = max(mx, __shfl_xor(mx, 16));
    mx = max(mx, __shfl_xor(mx, 32));
    mn = min(mn, __shfl_xor(mn, 16));
    mn = min(mn, __shfl_xor(mn, 32));
    if (mx == 0) continue;
    gp.m_max = mx;
    gp.m_min = mn;
    if (!gp.valid) gp.m = mx;
    unit_grad<double, MAX_R32, ANCH>(b, p, tab, gt, codes, smax, gp, ck, weights, out, dg, dx, lds_de, true);
  }
  grad_flush(dg, dx, lds_de, gpart);
}

// Sum of the per-block partial gradients in block order (deterministic for a
// given grid).
__global__ void k_grad_reduce(const double* __restrict__ gpart, int nparts, double* __restrict__ grad) {
  const int i = threadIdx.x;
  if (i < NGRAD) {
    double s = 0.0;
    for (int k = 0; k < nparts; ++k) s += gpart[(size_t)k * NGRAD + i];
    grad[i] = s;
  }
}

size_t lds_grad32(int smax) {
  return 64 * sizeof(double) + ((sizeof(QTab<float>) + 15) & ~size_t(15)) +
         (size_t)PAIRS_PER_UNIT * smax * sizeof(uint16_t);
}
size_t lds_grad64(int smax) {
  return 64 * sizeof(double) + ((sizeof(QTab<double>) + 15) & ~size_t(15)) +
         (size_t)PAIRS_PER_UNIT * smax * sizeof(uint16_t);
}

size_t grad_ck_stride(int smax, int fp64) {
  const size_t blocks = (size_t)(smax / GK + 2);
  (void)fp64;
  return blocks * CKW_MAX * 64;
}

hipError_t grad_occupancy(int anchored, int fp64, int smax, int* blocks_per_cu) {
  if (fp64)
    return anchored
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_grad64<true>, 64, lds_grad64(smax))
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_grad64<false>, 64, lds_grad64(smax));
  return anchored ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_grad32<true>, 64, lds_grad32(smax))
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_grad32<false>, 64, lds_grad32(smax));
}

hipError_t launch_grad(const BatchDev& b, const PlanDev& p, const int32_t* qE, const QTab<float>* tab32,
                       const QTab<double>* tab64, const GradTab* gt, int anchored, int smax, int blocks32,
                       int blocks64, float* ck32, double* ck64, const double* weights, double* out, double* gpart,
                       double* grad, hipStream_t st, hipEvent_t* ev) {
  if (ev) (void)hipEventRecord(ev[0], st);
  launch_plan(b, p, qE, anchored, 0, out, st);
  if (ev) (void)hipEventRecord(ev[1], st);
  const size_t s32 = grad_ck_stride(smax, 0), s64 = grad_ck_stride(smax, 1);
  if (anchored)
    hipLaunchKernelGGL(k_grad32<true>, dim3(blocks32), dim3(64), lds_grad32(smax), st, b, p, tab32, gt, smax, ck32,
                       s32, weights, out, gpart);
  else
    hipLaunchKernelGGL(k_grad32<false>, dim3(blocks32), dim3(64), lds_grad32(smax), st, b, p, tab32, gt, smax, ck32,
                       s32, weights, out, gpart);
  if (ev) (void)hipEventRecord(ev[2], st);
  double* gpart64 = gpart + (size_t)blocks32 * NGRAD;
  if (anchored)
    hipLaunchKernelGGL(k_grad64<true>, dim3(blocks64), dim3(64), lds_grad64(smax), st, b, p, tab64, gt, smax, ck64,
                       s64, weights, out, gpart64);
  else
    hipLaunchKernelGGL(k_grad64<false>, dim3(blocks64), dim3(64), lds_grad64(smax), st, b, p, tab64, gt, smax, ck64,
                       s64, weights, out, gpart64);
  hipLaunchKernelGGL(k_grad_reduce, dim3(1), dim3(128), 0, st, (const double*)gpart, blocks32 + blocks64, grad);
  if (ev) (void)hipEventRecord(ev[3], st);
  return hipGetLastError();
}

}  // namespace vd

// Adjoint (reverse-mode) kernels: gradients of sum_p w_p ln L_p with respect
// to the gap open g, the gap continuation x and the 64-bin error table.
//
// There is no counterpart in the reference; the model is the one the forward
// kernels evaluate (batch_hmm.cpp:31-70).  A unit is the forward kernel's
// unit (4 pairs of one region and row class, 16 lanes per pair, R rows per
// lane), processed by one wave in three sweeps:
//
// 1. forward over all anti-diagonal steps, writing the lane state to a
//    per-wave checkpoint area every GK * NSUB steps: for each of its R rows
//    x D^ + M^ (the next step's D^, the only use of the pair) and Q, plus the
//    two DPP carries; and computing each pair's terminal;
// 2. backward over the steps in reverse, checkpoint block by block, each in
//    NSUB register blocks of GK steps taken last to first: a register block
//    is recomputed from the checkpoint (the steps before it without storing)
//    with the per-step values the adjoint needs (each row's diagonal source
//    Q_{i-1}(j-1), its D^, the lane's injected P) kept in registers, then swept
//    backwards.  The checkpoint traffic is (2R + 2) values per lane every
//    GK * NSUB steps, at the price of recomputing (NSUB + 1) / 2 times;
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
// w / terminal, which makes every sum below a derivative of w ln L.  Per row
// the emission sum E_i = sum_j Mb Q_{i-1}(j-1) s (s = -1 on a match, 1/3 on a
// mismatch, 0 past the haplotype end) gives d/de of that row's quality bin
// (times a/rho); the other sums are per lane, the row constants being folded
// in through closed forms of their derivatives (mu = x/(4 rho),
// beta = (1-x) g/(4 a rho), cfin = g/(4 rho), gamma = (1-x) g/a):
//   U = sum mu Pb I^, V = sum beta Qb I^ (real rows but the read's last)
//                               -> d/dx: U/x - V/(1-x), d/dg: V/(g a)
//   W = sum cfin (Pb + Qb) I^ (the read's last row) -> d/dg: W/g
//   Tg = sum Qb D^ (gamma rows) -> d gamma/dg, d gamma/dx;  Tl (last row,
//   anchored gamma_last = g) -> d/dg;  Tx = sum Db_i(j+1) D^_i(j) -> d/dx.
// The emissions' d/dg, -2/a sum_rows sum_j Mb M^ (the M occupancy), comes from
// the row-cut identity: every path emits each read base once, in M or in I,
// so the M occupancy of the n real rows is n w - (U + V + W).
#include "vdhmm_device.h"

namespace vd {

constexpr int GK32 = 2;      // register block (anti-diagonal steps), FP32 (default variant)
constexpr int GK32B = 4;     // the alternative FP32 variant the autotuner may pick
constexpr int GK64 = 2;      // FP64 (twice the registers per value)
constexpr int NSUB = 2;      // register blocks per checkpoint
constexpr int CKW_MAX = 2 * MAX_R32 + 2;
constexpr int NBIN64 = 8;    // FP64 gradient row bins: R = 2, 4, ..., 16

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
// D^ (dd).  Emission is zero past the haplotype end (TAIL steps: codes past
// an end only occur at steps >= m_min + 2, as in the forward kernels).  FIRST:
// the state was just restored from a checkpoint (M = x D^ + M^, D = 0), so
// the new D^ is M itself.
template <typename T, int R, bool TAIL = true>
__device__ __forceinline__ T gstep(T (&M)[R], T (&D)[R], T (&Q)[R], const T (&gd)[R], const T (&bd)[R],
                                   const T (&mu)[R], const T (&be)[R], const uint32_t (&rc)[R], T gam, T gam_last,
                                   T xx, T Pin, T Qin, uint32_t code, T (&qu)[R], T (&dd)[R], bool first = false) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    D[r] = first ? M[r] : fmaT(xx, D[r], M[r]);
    dd[r] = D[r];
  }
  const bool post = TAIL && code == SENT_POST;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const T qup = r == 0 ? Qin : Q[r - 1];
    T e = rc[r] == code ? gd[r] : bd[r];
    if (TAIL) e = post ? T(0) : e;
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

// The adjoint's per-lane sums (see the top of this file).
template <typename T>
struct GradSums {
  // U, UL: sums of (mu Pb + be Qb) I over the interior rows and the read's
  // last row; V, VL: their be Qb I parts; Tg, Tl, Tx: the gamma and
  // extension terms
  T U, V, UL, VL, Tg, Tl, Tx;
};

// One backward anti-diagonal step: the transposed recurrence over the
// lane's R rows, bottom to top, from the step's stored diagonal sources (sQ),
// D^ (sD) and injected P (sPin), with the gradient sums.  The reverse carries
// come from lane l + 1; lane 15 takes pold into its P carry and qold into the
// carry that becomes its next step's Qb (the seeds).  TAIL: the step may see
// columns past a haplotype end (code SENT_POST: emission and emission-sum
// weight zero).
template <typename T, int R, bool TAIL>
__device__ __forceinline__ void bwd_step(uint32_t code, T pold, T qold, const T (&sQ)[R], const T (&sD)[R], T sPin,
                                         const T (&gd)[R], const T (&bd)[R], const T (&mu)[R], const T (&be)[R],
                                         const uint32_t (&rc)[R], const bool (&pass)[R], T gam, T gam_last, T xx,
                                         T (&EMp)[R], T (&Dbr)[R], T (&Er)[R], GradSums<T>& gs, T& Ilast, T& Eprev) {
#pragma clang fp contract(off)
  const bool post = TAIL && code == SENT_POST;
  const T smis = post ? T(0) : T(1.0 / 3.0);  // emission-sum weight of a mismatch
  const T Pbin = dpp_shl1(pold, Ilast);
  const T Qbin = Eprev;
  const T En = dpp_shl1(qold, EMp[0]);
  T Qb[R], e[R], sw[R], Iv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) Qb[r] = r == R - 1 ? Qbin : EMp[r + 1];
  T Ih = sPin;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool mt = rc[r] == code;
    e[r] = mt ? gd[r] : bd[r];
    if (TAIL) e[r] = post ? T(0) : e[r];
    sw[r] = mt ? T(-1) : smis;
    Iv[r] = pass[r] ? T(0) : Ih;
    Ih = fmaT(mu[r], Ih, e[r] * sQ[r]);
  }
  T Pb = Pbin;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const T gr = r == R - 1 ? gam_last : gam;
    const T Dbar = fmaT(xx, Dbr[r], gr * Qb[r]);
    const T Mbar = Pb + Qb[r] + Dbr[r];
    const T EM = e[r] * Mbar;
    Er[r] = fmaT(Mbar * sQ[r], sw[r], Er[r]);
    const T v = be[r] * Qb[r];
    if (r == R - 1) {
      gs.VL = fmaT(v, Iv[r], gs.VL);
      gs.Tl = fmaT(Qb[r], sD[r], gs.Tl);
    } else {
      gs.V = fmaT(v, Iv[r], gs.V);
      gs.Tg = fmaT(Qb[r], sD[r], gs.Tg);
    }
    gs.Tx = fmaT(Dbr[r], sD[r], gs.Tx);
    Dbr[r] = Dbar;
    EMp[r] = EM;
    Pb = fmaT(mu[r], Pb, v);  // one dependent operation per row on the carry chain
    // the new carry is mu Pb + be Qb: U (UL) sums it times Iv, so the mu
    // part alone is U - V, with no product of its own
    if (r == R - 1) gs.UL = fmaT(Pb, Iv[r], gs.UL);
    else gs.U = fmaT(Pb, Iv[r], gs.U);
  }
  Ilast = Pb;
  Eprev = En;
}

template <typename T, int R, bool ANCH, int GK>
__device__ __forceinline__ void unit_grad(const BatchDev& b, const PlanDev& p, const QTab<T>* __restrict__ tab,
                                          const GradTab* __restrict__ gt, uint16_t* codes, int smax,
                                          const GroupPair gp, T* __restrict__ ck, const double* __restrict__ weights,
                                          double* __restrict__ out, double* acc, double* lds_de, bool rescue,
                                          T* ckl = nullptr) {
  // only the written fmas fuse: every instance and variant rounds alike, so
  // every launch choice gives the same gradients
#pragma clang fp contract(off)
  constexpr int CKW = 2 * R + 2;
  const int lane = threadIdx.x & 63, l = lane & 15;
  const int tend = gp.m_max + 16;
  // codes up to the padded last step of the last checkpoint block
  fill_codes(b, gp, codes, smax, tend + 16 + GK * NSUB);
  T gd[R], bd[R], mu[R], be[R];
  uint32_t rc[R];
  bool pass[R];  // pass-through row above the read (structural mu = beta = 1)
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
    pass[r] = rel < 0;
  }
  const T gam = tab->gamma, xx = tab->ext;
  const T gam_last = l == GROUP - 1 ? (ANCH ? tab->gap : T(0)) : gam;
  const T C = (T)ldexp(1.0, GradScale<T>::CEXP);
  const uint16_t* cl = codes + (lane >> 4) * smax + 16 - l;
  __syncthreads();

  // ---- sweep 1: forward with checkpoints.  Steps before tmain see no column
  // past any haplotype end; they run in aligned groups of CK_STEPS without the
  // end test, the rest with it.
  T M[R], D[R], Q[R], qu[R], dd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) M[r] = D[r] = Q[r] = 0;
  T Plast = 0, Xprev = 0, term_acc = 0, rmax = 0, latch = 0;
  constexpr int CK_STEPS = GK * NSUB;
  const int tmain = gp.m_min + 2;
  auto ck_store = [&](int t) {
    T* c = ck + (size_t)(t / CK_STEPS) * CKW * 64 + lane;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      c[(size_t)r * 64] = fmaT(xx, D[r], M[r]);
      c[(size_t)(R + r) * 64] = Q[r];
    }
    c[(size_t)(2 * R) * 64] = Plast;
    c[(size_t)(2 * R + 1) * 64] = Xprev;
  };
  auto fwd_step = [&](int t, auto tail_c) {
    constexpr bool TL = decltype(tail_c)::value;
    T pinj, qinj;
    injections<T, ANCH>(t, gp.m, C, pinj, qinj);
    const T Pin = dpp_shr1(pinj, Plast);
    const T Xn = dpp_shr1(qinj, Q[R - 1]);
    Plast = gstep<T, R, TL>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code_reg(cl[t]), qu, dd);
    Xprev = Xn;
    if (ANCH) {
      rmax = fmax(rmax, Q[R - 1]);
      // lane 15 reaches column m at step m + 16 >= tmain
      if (TL && t - l - 1 == gp.m) latch = Q[R - 1];
    } else {
      term_acc += Plast;
    }
  };
  int t1 = 0;
  for (; t1 + CK_STEPS <= tmain && t1 + CK_STEPS - 1 <= tend; t1 += CK_STEPS) {
    ck_store(t1);
#pragma unroll
    for (int k = 0; k < CK_STEPS; ++k) fwd_step(t1 + k, std::false_type{});
  }
  for (; t1 <= tend; ++t1) {
    if (t1 % CK_STEPS == 0) ck_store(t1);
    fwd_step(t1, std::true_type{});
  }

  // ---- terminal, flag and seed (lane 15 of each group holds the pair's terminal)
  const T term = ANCH ? latch : term_acc;
  bool flag = false;
  double seed = 0.0, wpair = 0.0;
  if (l == GROUP - 1 && gp.valid) {
    const double lnk = sizeof(T) == 4 ? p.read_ln32[gp.read] - GradScale<T>::CEXP * LN2
                                      : (double)(p.read_exp[gp.read] - GradScale<T>::CEXP) * LN2;
    double v;
    if (ANCH) {
      const double lc = log((double)rmax) + LN_1EM300;
      const double ll = log((double)latch);
      v = (ll > lc ? ll : lc) + lnk;
    } else {
      v = log((double)term_acc) + lnk - log((double)(gp.m + 1));
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
    wpair = weights ? weights[gp.out] : 1.0;
    seed = flag ? 0.0 : wpair / (double)term;
  }
  {  // the group's lane 15 holds the terminal: broadcast by readlane (scalar)
    const int g = lane >> 4;
    const int64_t sb = __builtin_bit_cast(int64_t, seed);
    int64_t sv[PAIRS_PER_UNIT];
    int fv[PAIRS_PER_UNIT];
#pragma unroll
    for (int q = 0; q < PAIRS_PER_UNIT; ++q) {
      const int lo = __builtin_amdgcn_readlane((int)(uint32_t)sb, 16 * q + 15);
      const int hi = __builtin_amdgcn_readlane((int)(sb >> 32), 16 * q + 15);
      sv[q] = (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
      fv[q] = __builtin_amdgcn_readlane((int)flag, 16 * q + 15);
    }
    const int64_t pick = g == 0 ? sv[0] : g == 1 ? sv[1] : g == 2 ? sv[2] : sv[3];
    seed = __builtin_bit_cast(double, pick);
    flag = (g == 0 ? fv[0] : g == 1 ? fv[1] : g == 2 ? fv[2] : fv[3]) != 0;
    // a unit whose four seeds are all zero (weights 0, flagged or invalid
    // pairs) contributes nothing to the gradient: its outputs are written,
    // the backward sweeps are skipped (the decision is the wave's: scalars)
    if ((sv[0] | sv[1] | sv[2] | sv[3]) == 0) {
      __syncthreads();  // codes reused by the next unit
      return;
    }
  }
  const T tseed = (T)seed;

  // ---- sweep 2: backward, block by block.  FP32: every block runs its
  // CK_STEPS steps (the last one's steps past tend see columns past every
  // haplotype end, where the forward values on the P chain and every gradient
  // term are zero), so its register blocks are straight-line code without
  // step guards, and the blocks below tmain skip the end test.  FP64 keeps
  // guarded steps: the straight-line form needs more than its 512 registers.
  T EMp[R], Dbr[R], Er[R];
#pragma unroll
  for (int r = 0; r < R; ++r) EMp[r] = Dbr[r] = Er[r] = 0;
  GradSums<T> gs{0, 0, 0, 0, 0, 0, 0};
  T Ilast = 0;
  const T Pseed = ANCH ? T(0) : tseed;
  const int nblk = tend / CK_STEPS + 1;
  if constexpr (sizeof(T) == 4) {
    const int tlast = nblk * CK_STEPS - 1;
    // the anchored terminal enters lane 15 through the reverse carry at step
    // m + 17; a pair with m + 16 = tlast has it before the first backward step
    T Eprev = (ANCH && l == GROUP - 1 && tlast - 16 == gp.m) ? tseed : T(0);
    // checkpoint blocks are staged through LDS (ckl): the memory unit copies
    // block blk - 1's checkpoint straight into LDS while block blk runs, so
    // the wave never waits on HBM at a block start
    T cur[CKW];
    typedef __attribute__((address_space(3))) void* lds_ptr;
    auto stage = [&](int blk) {
      const T* c = ck + (size_t)blk * CKW * 64 + lane;
#pragma unroll
      for (int w = 0; w < CKW; ++w)
        __builtin_amdgcn_global_load_lds(c + (size_t)w * 64, (lds_ptr)(ckl + w * 64), 4, 0, 0);
    };
    stage(nblk - 1);
    // one checkpoint block, its NSUB register blocks last to first
    auto bwd_block = [&](int tb, auto tail_c) {
      constexpr bool TL = decltype(tail_c)::value;
#pragma unroll
      for (int sub = NSUB - 1; sub >= 0; --sub) {
        // the register blocks are independent until their backward steps: keep
        // the scheduler from overlapping them (that doubles the live values)
        __builtin_amdgcn_sched_barrier(0);
        const int t0 = tb + sub * GK;
        T Mr[R], Dr[R], Qr[R], qx[R], dx[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          Mr[r] = cur[r];
          Dr[r] = 0;
          Qr[r] = cur[R + r];
        }
        T Pl = cur[2 * R], Xp = cur[2 * R + 1];
#pragma unroll
        for (int k = 0; k < sub * GK; ++k) {  // steps before the register block, not stored
          const int t = tb + k;
          T pinj, qinj;
          injections<T, ANCH>(t, gp.m, C, pinj, qinj);
          const T Pin = dpp_shr1(pinj, Pl);
          const T Xn = dpp_shr1(qinj, Qr[R - 1]);
          Pl = gstep<T, R, TL>(Mr, Dr, Qr, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xp, code_reg(cl[t]), qx, dx,
                               k == 0);
          Xp = Xn;
        }
        T sQ[GK][R], sD[GK][R], sPin[GK];
#pragma unroll
        for (int s = 0; s < GK; ++s) {
          const int t = t0 + s;
          T pinj, qinj;
          injections<T, ANCH>(t, gp.m, C, pinj, qinj);
          const T Pin = dpp_shr1(pinj, Pl);
          const T Xn = dpp_shr1(qinj, Qr[R - 1]);
          Pl = gstep<T, R, TL>(Mr, Dr, Qr, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xp, code_reg(cl[t]), sQ[s],
                               sD[s], sub == 0 && s == 0);
          Xp = Xn;
          sPin[s] = Pin;
        }
#pragma unroll
        for (int s = GK - 1; s >= 0; --s) {
          const int t = t0 + s;
          // the anchored terminal enters lane 15 at step m + 17 (past the main blocks)
          const T qold = (TL && ANCH && t - 17 == gp.m) ? tseed : T(0);
          bwd_step<T, R, TL>(code_reg(cl[t]), Pseed, qold, sQ[s], sD[s], sPin[s], gd, bd, mu, be, rc, pass, gam,
                             gam_last, xx, EMp, Dbr, Er, gs, Ilast, Eprev);
        }
      }
    };
    auto load_block = [&](int blk) {
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the staged copy has landed
#pragma unroll
      for (int w = 0; w < CKW; ++w) cur[w] = ckl[w * 64 + lane];
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): read before the next copy overwrites it
      if (blk > 0) stage(blk - 1);
    };
    // blocks that reach tmain or beyond test for columns past an end, the
    // blocks below run without the test (two loops, so the two variants of the
    // block never merge their registers)
    const int nmain = tmain / CK_STEPS;
    int blk = nblk - 1;
    for (; blk >= nmain; --blk) {
      load_block(blk);
      bwd_block(blk * CK_STEPS, std::true_type{});
    }
    for (; blk >= 0; --blk) {
      load_block(blk);
      bwd_block(blk * CK_STEPS, std::false_type{});
    }
  } else {
    // the anchored terminal of a pair whose haplotype is the unit's longest
    // enters before the first backward step
    T Eprev = (ANCH && l == GROUP - 1 && tend - 16 == gp.m) ? tseed : T(0);
    constexpr bool PREF = sizeof(T) == 4 ? false : R >= 8;
    T cur[CKW], ckn[PREF ? CKW : 1];
    if (PREF) {
      const T* c = ck + (size_t)(nblk - 1) * CKW * 64 + lane;
#pragma unroll
      for (int w = 0; w < (PREF ? CKW : 1); ++w) ckn[w] = c[(size_t)w * 64];
    }
    for (int blk = nblk - 1; blk >= 0; --blk) {
      if (PREF) {
#pragma unroll
        for (int w = 0; w < CKW; ++w) cur[w] = ckn[w % (PREF ? CKW : 1)];
        if (blk > 0) {
          const T* c = ck + (size_t)(blk - 1) * CKW * 64 + lane;
#pragma unroll
          for (int w = 0; w < (PREF ? CKW : 1); ++w) ckn[w] = c[(size_t)w * 64];
        }
      } else {
        const T* c = ck + (size_t)blk * CKW * 64 + lane;
#pragma unroll
        for (int w = 0; w < CKW; ++w) cur[w] = c[(size_t)w * 64];
      }
      const int tb = blk * CK_STEPS;
      for (int sub = NSUB - 1; sub >= 0; --sub) {
        const int t0 = tb + sub * GK;
        if (t0 > tend) continue;
        const int cnt = min(GK, tend + 1 - t0);
        // state before step tb: M = x D^ + M^ and D = 0 make the first step's
        // D update reproduce D^(tb) exactly
#pragma unroll
        for (int r = 0; r < R; ++r) {
          M[r] = cur[r];
          D[r] = 0;
          Q[r] = cur[R + r];
        }
        Plast = cur[2 * R];
        Xprev = cur[2 * R + 1];
        for (int t = tb; t < t0; ++t) {  // steps before the register block, not stored
          T pinj, qinj;
          injections<T, ANCH>(t, gp.m, C, pinj, qinj);
          const T Pin = dpp_shr1(pinj, Plast);
          const T Xn = dpp_shr1(qinj, Q[R - 1]);
          Plast = gstep<T, R>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code_reg(cl[t]), qu, dd);
          Xprev = Xn;
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
            Plast = gstep<T, R>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code_reg(cl[t]), sQ[s],
                                sD[s]);
            Xprev = Xn;
            sPin[s] = Pin;
          }
        }
#pragma unroll
        for (int s = GK - 1; s >= 0; --s) {
          if (s < cnt) {
            const int t = t0 + s;
            const T qold = (ANCH && t - 17 == gp.m) ? tseed : T(0);
            bwd_step<T, R, true>(code_reg(cl[t]), Pseed, qold, sQ[s], sD[s], sPin[s], gd, bd, mu, be, rc, pass, gam,
                                 gam_last, xx, EMp, Dbr, Er, gs, Ilast, Eprev);
          }
        }
      }
    }
  }

  // ---- sweep 3: gradient terms of the pair
  if (gp.valid && !flag) {
    const double g = (double)tab->gap, x = (double)tab->ext, a = 1.0 - 2.0 * g;
    const bool last_lane = l == GROUP - 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int kk = l * R + r - pad - (ANCH ? 1 : 0);
      if (kk >= 0) {
        int q = b.qual_bytes[rb + kk];
        q = q < 63 ? q : 63;
        // d good/de = -a/rho for the precision's own normaliser
        atomicAdd(&lds_de[q], (double)Er[r] * (sizeof(T) == 4 ? gt->dgood_de32[q] : gt->dgood_de[q]) * -1.0);
      }
    }
    const double uv = (double)gs.U + (last_lane ? 0.0 : (double)gs.UL);
    const double v = (double)gs.V + (last_lane ? 0.0 : (double)gs.VL);
    const double u = uv - v;
    const double w = last_lane ? (double)gs.UL : 0.0;
    const double tg = (double)gs.Tg + (last_lane ? 0.0 : (double)gs.Tl);
    const double tl = last_lane ? (double)gs.Tl : 0.0;
    const double mocc = last_lane ? (double)gp.n * wpair : 0.0;  // n w, once per pair
    acc[0] += -2.0 / a * (mocc - (u + v + w)) + v / (g * a) + w / g + tg * gt->dgam_dg + (ANCH ? tl : 0.0);
    acc[1] += u / x - v / (1.0 - x) + tg * gt->dgam_dx + (double)gs.Tx;
  }
  __syncthreads();  // codes reused by the next unit
}

// Adds the block's gradient (lane sums of d/dg, d/dx and the LDS per-bin
// d/de) into its partial row; the per-class kernels run one after another on
// one stream, so each adds to the rows the previous ones left.
__device__ __forceinline__ void grad_flush(double dg, double dx, const double* lds_de, double* __restrict__ gpart) {
  for (int o = 32; o > 0; o >>= 1) {
    dg += __shfl_xor(dg, o);
    dx += __shfl_xor(dx, o);
  }
  __syncthreads();
  double* gp = gpart + (size_t)blockIdx.x * NGRAD;
  if (threadIdx.x == 0) {
    gp[0] += dg;
    gp[1] += dx;
  }
  for (int q = threadIdx.x; q < 64; q += blockDim.x) gp[2 + q] += lds_de[q];
}

template <typename T>
__device__ __forceinline__ void grad_lds_init(const void* gtab, size_t tab_bytes, unsigned char* lds) {
  double* lds_de = reinterpret_cast<double*>(lds);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(gtab);
  uint32_t* dst = reinterpret_cast<uint32_t*>(lds + 64 * sizeof(double));
  for (int i = threadIdx.x; i < (int)(tab_bytes / 4); i += blockDim.x) dst[i] = src[i];
  for (int q = threadIdx.x; q < 64; q += blockDim.x) lds_de[q] = 0.0;
  __syncthreads();
}

// FP32 adjoint of the units of row class R: the contiguous unit range of the
// class in the class-major plan, taken from a per-class counter.
template <bool ANCH, int R, int GK>
__device__ __forceinline__ void grad32_body(const BatchDev& b, const PlanDev& p, const QTab<float>* __restrict__ gtab,
                                            const GradTab* __restrict__ gt, int smax, float* __restrict__ ckall,
                                            size_t ck_stride, const double* __restrict__ weights,
                                            double* __restrict__ out, double* __restrict__ gpart) {
  // the class's units: [lo, hi) of the unit sequence, or, with packed pairs,
  // the pairs [lo, hi) in units of 4 that may span regions
  const unsigned long long lo = (unsigned long long)p.unit_cnt[bucket_of(R, 0, b.nregions, 1)];
  const unsigned long long hi = (unsigned long long)p.unit_cnt[bucket_of(R - 1, 0, b.nregions, 1)];
  if (lo >= hi) return;  // no read of this class: the block's partial row stays as it is
  const unsigned long long nunits = p.pack_pairs ? (hi - lo + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT : hi - lo;
  extern __shared__ __align__(16) unsigned char lds[];
  grad_lds_init<float>(gtab, qtab_lds_bytes<float>(), lds);
  double* lds_de = reinterpret_cast<double*>(lds);
  const QTab<float>* tab = reinterpret_cast<const QTab<float>*>(lds + 64 * sizeof(double));
  uint16_t* codes =
      reinterpret_cast<uint16_t*>(lds + 64 * sizeof(double) + ((qtab_lds_bytes<float>() + 15) & ~size_t(15)));
  float* ckl = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(codes) + lds_codes_bytes(smax));
  float* ck = ckall + (size_t)blockIdx.x * ck_stride;
  const int lane = threadIdx.x & 63;
  double acc[2] = {0.0, 0.0};
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
    unit_grad<float, R, ANCH, GK>(b, p, tab, gt, codes, smax, gp, ck, weights, out, acc, lds_de, false, ckl);
  }
  grad_flush(acc[0], acc[1], lds_de, gpart);
}

// Two variants per class: k_grad32w (register blocks of GK32 = 2 steps,
// compiled for two waves per SIMD, for classes up to GRAD_W2_MAX rows per lane
// where that fits without heavy spills) and k_grad32 (GK32B = 4 steps, the
// occupancy the registers allow).  The autotuner picks one per class; the
// default is k_grad32w where it exists.
constexpr int GRAD_W2_MAX = 10;

template <bool ANCH, int R>
__global__ __launch_bounds__(64) void k_grad32(BatchDev b, PlanDev p, const QTab<float>* __restrict__ gtab,
                                               const GradTab* __restrict__ gt, int smax, float* __restrict__ ckall,
                                               size_t ck_stride, const double* __restrict__ weights,
                                               double* __restrict__ out, double* __restrict__ gpart) {
  grad32_body<ANCH, R, GK32B>(b, p, gtab, gt, smax, ckall, ck_stride, weights, out, gpart);
}

template <bool ANCH, int R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_grad32w(
    BatchDev b, PlanDev p, const QTab<float>* __restrict__ gtab, const GradTab* __restrict__ gt, int smax,
    float* __restrict__ ckall, size_t ck_stride, const double* __restrict__ weights, double* __restrict__ out,
    double* __restrict__ gpart) {
  grad32_body<ANCH, R, GK32>(b, p, gtab, gt, smax, ckall, ck_stride, weights, out, gpart);
}

// Bins the FP64 gradient list (the pairs the FP32 pass flagged, and the reads
// beyond 256 rows) by rows per lane, rounded up to an even count, so that each
// FP64 kernel runs with the register footprint of its bin; reads beyond 16
// rows per lane go to bin NBIN64, the tiled adjoint's.
template <bool ANCH>
__global__ __launch_bounds__(256) void k_bin64(BatchDev b, PlanDev p, PairRef* __restrict__ bins, int64_t bin_cap) {
  __shared__ unsigned int cnt[NBIN64 + 1];
  __shared__ unsigned long long base[NBIN64 + 1];
  unsigned long long count = p.ctr[CTR_F64];
  if ((int64_t)count > p.f64_cap) count = (unsigned long long)p.f64_cap;
  for (unsigned long long i0 = (unsigned long long)blockIdx.x * blockDim.x; i0 < count;
       i0 += (unsigned long long)gridDim.x * blockDim.x) {
    if (threadIdx.x <= NBIN64) cnt[threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long i = i0 + threadIdx.x;
    int bin = -1;
    unsigned int slot = 0;
    PairRef pr{0, 0, 0};
    if (i < count) {
      pr = p.f64_list[i];
      const int np = (int)(b.read_off[pr.read + 1] - b.read_off[pr.read]) + (ANCH ? 1 : 0);
      const int rows = (np + GROUP - 1) / GROUP;
      bin = rows <= MAX_R32 ? (rows + 1) / 2 - 1 : NBIN64;
      slot = atomicAdd(&cnt[bin], 1u);
    }
    __syncthreads();
    if (threadIdx.x <= NBIN64) {
      const int c = threadIdx.x < NBIN64 ? CTR_BIN64 + (int)threadIdx.x : CTR_BIN64T;
      base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&p.ctr[c], (unsigned long long)cnt[threadIdx.x]) : 0ull;
    }
    __syncthreads();
    if (bin >= 0) {
      const unsigned long long k = base[bin] + slot;
      if ((int64_t)k < bin_cap) bins[(size_t)bin * bin_cap + k] = pr;
    }
    __syncthreads();
  }
}

// FP64 adjoint of one row bin of the flagged pairs, 4 listed pairs per unit.
template <bool ANCH, int R>
__global__ __launch_bounds__(64) void k_grad64(BatchDev b, PlanDev p, const QTab<double>* __restrict__ gtab,
                                               const GradTab* __restrict__ gt, int smax, double* __restrict__ ckall,
                                               size_t ck_stride, const PairRef* __restrict__ bins, int64_t bin_cap,
                                               const double* __restrict__ weights, double* __restrict__ out,
                                               double* __restrict__ gpart) {
  constexpr int BIN = R / 2 - 1;
  unsigned long long count = p.ctr[CTR_BIN64 + BIN];
  // an empty bin adds nothing: the partial rows were zeroed at the call's start
  if (count == 0) return;
  if ((int64_t)count > bin_cap) count = (unsigned long long)bin_cap;
  extern __shared__ __align__(16) unsigned char lds[];
  grad_lds_init<double>(gtab, qtab_lds_bytes<double>(), lds);
  double* lds_de = reinterpret_cast<double*>(lds);
  const QTab<double>* tab = reinterpret_cast<const QTab<double>*>(lds + 64 * sizeof(double));
  uint16_t* codes =
      reinterpret_cast<uint16_t*>(lds + 64 * sizeof(double) + ((qtab_lds_bytes<double>() + 15) & ~size_t(15)));
  double* ck = ckall + (size_t)blockIdx.x * ck_stride;
  const int lane = threadIdx.x & 63, g = lane >> 4;
  double acc[2] = {0.0, 0.0};
  const PairRef* list = bins + (size_t)BIN * bin_cap;
  const unsigned long long units = (count + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT;
  for (;;) {
    unsigned long long u = 0;
    if (lane == 0) u = atomicAdd(&p.ctr[CTR_UNIT64], 1ull);
    u = first_lane(u);
    if (u >= units) break;
    const unsigned long long idx = u * PAIRS_PER_UNIT + g;
    GroupPair gp;
    gp.valid = idx < count;
    const PairRef pr = list[gp.valid ? idx : u * PAIRS_PER_UNIT];
    gp.read = pr.read;
    gp.hap = pr.hap;
    gp.out = pr.out;
    gp.n = (int)(b.read_off[gp.read + 1] - b.read_off[gp.read]);
    gp.m = (int)(b.hap_off[gp.hap + 1] - b.hap_off[gp.hap]);
    if (gp.valid && (gp.m == 0 || gp.n + (ANCH ? 1 : 0) > GROUP * R)) gp.valid = 0;
    const int mx0 = gp.valid ? gp.m : 0, mn0 = gp.valid ? gp.m : 0x7fffffff;
    const int mx = max(max(lane16(mx0, 0), lane16(mx0, 1)), max(lane16(mx0, 2), lane16(mx0, 3)));
    const int mn = min(min(lane16(mn0, 0), lane16(mn0, 1)), min(lane16(mn0, 2), lane16(mn0, 3)));
    if (mx == 0) continue;
    gp.m_max = mx;
    gp.m_min = mn;
    if (!gp.valid) gp.m = mx;
    unit_grad<double, R, ANCH, GK64>(b, p, tab, gt, codes, smax, gp, ck, weights, out, acc, lds_de, true);
  }
  grad_flush(acc[0], acc[1], lds_de, gpart);
}

// ---------------------------------------------------------------------------
// Tiled FP64 adjoint for reads beyond 256 rows (no counterpart in the
// reference, like every adjoint here).  A unit is 4 listed pairs of the long
// bin, 16 lanes per pair, R64 = 4 rows per lane: the FP64 forward's 64-row
// tiles.
//
// Phase A is the tiled FP64 forward (tile64, as k_fwd64 runs it), keeping
// every tile's last row (the P and Q streams lane 15 hands down) and the
// power-of-two factors its successor's injection is scaled by, and the
// pair's terminal, ln L and seed.
//
// Phase B takes the tiles last to first.  A tile runs unit_grad's three
// sweeps on its 64 rows: forward from its injection (the previous tile's
// stored row times its factors, or the model's start for the first tile)
// with checkpoints, then backward block by block.  Its backward seeds at the
// bottom are the terminal seeds for the last tile and otherwise the adjoints
// of the tile below's injection, per column; its own injection's adjoints
// (lane 0's outgoing Pb and EM, per column) are written out, times the
// injection's factors, for the tile above.  Every gradient sum is a product
// of a forward value and an adjoint of the same tile, so the tile scales
// cancel in it; the sums of the tiles add up.

constexpr int GKT = 2;  // register block (steps) of the tiled adjoint
constexpr double C64T = 0x1p500;  // tile64's start scale

// Tile workspace (doubles) of one block, for tiles of at most ntmax per pair:
//   stored forward rows: tile k < ntmax - 1, stream s (P, Q), group g at ((k * 2 + s) * 4 + g) * smax
//   adjoint rows (two buffers): A + ((buf * 2 + s) * 4 + g) * smax, A = (ntmax - 1) * 8 * smax
//   injection factors of tile k: F + (k * 4 + g) * 2 + {0, 1}, F = A + 16 * smax
size_t grad_tile_ws(int smax, int ntmax) {
  return ((size_t)(ntmax > 1 ? ntmax - 1 : 0) * 8 + 16) * (size_t)smax + (size_t)ntmax * 8;
}

template <bool ANCH, bool FIRST, bool LAST>
__device__ __forceinline__ void tile_grad64(const GroupPair& gp, const uint16_t* cl, int tend, const double (&gd)[R64],
                                            const double (&bd)[R64], const double (&mu)[R64],
                                            const double (&be)[R64], const uint32_t (&rc)[R64],
                                            const bool (&pass)[R64], double gam, double gam_last, double xx,
                                            const double* __restrict__ inP, const double* __restrict__ inQ,
                                            double sc1, double sc2, const double* __restrict__ bP,
                                            const double* __restrict__ bQ, double* __restrict__ tP,
                                            double* __restrict__ tQ, double tseed, double* __restrict__ ck,
                                            double (&Er)[R64], GradSums<double>& gs) {
  constexpr int R = R64, CKW = 2 * R + 2, CK_STEPS = GKT * NSUB;
  const int lane = threadIdx.x & 63, l = lane & 15;
  // this tile's injection at step t (lane 0 takes it through DPP's old operand)
  auto inj = [&](int t, double& pinj, double& qinj) {
    if (FIRST) {
      injections<double, ANCH>(t, gp.m, C64T, pinj, qinj);
    } else {
      pinj = qinj = 0.0;
      if (t - 1 <= gp.m_max) {  // column t - 1, slot t + 16: the slots tile64 loads
        pinj = inP[t + 16] * sc1 * sc2;
        qinj = inQ[t + 16] * sc1 * sc2;
      }
    }
  };
  // ---- sweep 1: forward with checkpoints
  double M[R], D[R], Q[R], qu[R], dd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) M[r] = D[r] = Q[r] = 0.0;
  double Plast = 0.0, Xprev = 0.0;
  for (int t = 0; t <= tend; ++t) {
    if (t % CK_STEPS == 0) {
      double* c = ck + (size_t)(t / CK_STEPS) * CKW * 64 + lane;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        c[(size_t)r * 64] = fmaT(xx, D[r], M[r]);
        c[(size_t)(R + r) * 64] = Q[r];
      }
      c[(size_t)(2 * R) * 64] = Plast;
      c[(size_t)(2 * R + 1) * 64] = Xprev;
    }
    double pinj, qinj;
    inj(t, pinj, qinj);
    const double Pin = dpp_shr1(pinj, Plast);
    const double Xn = dpp_shr1(qinj, Q[R - 1]);
    Plast = gstep<double, R>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code_reg(cl[t]), qu, dd);
    Xprev = Xn;
  }
  // ---- sweep 2: backward, block by block
  // lane 15's seeds from below: the terminal's for the last tile, otherwise
  // the tile below's injection adjoints (already times its factors):
  // Pb of column t - 16 at step t (slot t + 1), and Qb of column t - 17
  // handed in at step t for step t - 1 (slot t)
  auto seed_p = [&](int t) -> double { return LAST ? (ANCH ? 0.0 : tseed) : bP[t + 1]; };
  auto seed_q = [&](int t) -> double {
    return LAST ? ((ANCH && t - 17 == gp.m) ? tseed : 0.0) : bQ[t];
  };
  double EMp[R], Dbr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) EMp[r] = Dbr[r] = 0.0;
  double Ilast = 0.0;
  double Eprev = l == GROUP - 1 ? (LAST ? ((ANCH && tend - 16 == gp.m) ? tseed : 0.0) : bQ[tend + 1]) : 0.0;
  const int nblk = tend / CK_STEPS + 1;
  double cur[CKW];
  for (int blk = nblk - 1; blk >= 0; --blk) {
    {
      const double* c = ck + (size_t)blk * CKW * 64 + lane;
#pragma unroll
      for (int w = 0; w < CKW; ++w) cur[w] = c[(size_t)w * 64];
    }
    const int tb = blk * CK_STEPS;
    for (int sub = NSUB - 1; sub >= 0; --sub) {
      const int t0 = tb + sub * GKT;
      if (t0 > tend) continue;
      const int cnt = min(GKT, tend + 1 - t0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        M[r] = cur[r];
        D[r] = 0.0;
        Q[r] = cur[R + r];
      }
      Plast = cur[2 * R];
      Xprev = cur[2 * R + 1];
      for (int t = tb; t < t0; ++t) {
        double pinj, qinj;
        inj(t, pinj, qinj);
        const double Pin = dpp_shr1(pinj, Plast);
        const double Xn = dpp_shr1(qinj, Q[R - 1]);
        Plast = gstep<double, R>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code_reg(cl[t]), qu, dd);
        Xprev = Xn;
      }
      double sQ[GKT][R], sD[GKT][R], sPin[GKT];
#pragma unroll
      for (int s = 0; s < GKT; ++s) {
        if (s < cnt) {
          const int t = t0 + s;
          double pinj, qinj;
          inj(t, pinj, qinj);
          const double Pin = dpp_shr1(pinj, Plast);
          const double Xn = dpp_shr1(qinj, Q[R - 1]);
          Plast = gstep<double, R>(M, D, Q, gd, bd, mu, be, rc, gam, gam_last, xx, Pin, Xprev, code_reg(cl[t]), sQ[s],
                                   sD[s]);
          Xprev = Xn;
          sPin[s] = Pin;
        }
      }
#pragma unroll
      for (int s = GKT - 1; s >= 0; --s) {
        if (s < cnt) {
          const int t = t0 + s;
          bwd_step<double, R, true>(code_reg(cl[t]), seed_p(t), seed_q(t), sQ[s], sD[s], sPin[s], gd, bd, mu, be, rc,
                                    pass, gam, gam_last, xx, EMp, Dbr, Er, gs, Ilast, Eprev);
          // lane 0: adjoints of this step's injection, Pin (column t - 1) and
          // the diagonal source Qin (column t - 2), times the injection's
          // factors for the tile above
          if (!FIRST && l == 0) {
            tP[t + 16] = Ilast * sc1 * sc2;
            tQ[t + 15] = EMp[0] * sc1 * sc2;
          }
        }
      }
    }
  }
}

template <bool ANCH>
__global__ __launch_bounds__(64) void k_grad64t(BatchDev b, PlanDev p, const QTab<double>* __restrict__ gtab,
                                                const GradTab* __restrict__ gt, int smax, int ntmax,
                                                double* __restrict__ ckall, size_t ck_stride,
                                                double* __restrict__ wsall, size_t ws_stride,
                                                const PairRef* __restrict__ bins, int64_t bin_cap,
                                                const double* __restrict__ weights, double* __restrict__ out,
                                                double* __restrict__ gpart) {
  constexpr int R = R64;
  // no long read: nothing to add (the partial rows were zeroed at the call's start)
  if (p.ctr[CTR_BIN64T] == 0) return;
  extern __shared__ __align__(16) unsigned char lds[];
  grad_lds_init<double>(gtab, qtab_lds_bytes<double>(), lds);
  double* lds_de = reinterpret_cast<double*>(lds);
  const QTab<double>* tab = reinterpret_cast<const QTab<double>*>(lds + 64 * sizeof(double));
  uint16_t* codes =
      reinterpret_cast<uint16_t*>(lds + 64 * sizeof(double) + ((qtab_lds_bytes<double>() + 15) & ~size_t(15)));
  double* ck = ckall + (size_t)blockIdx.x * ck_stride;
  double* ws = wsall + (size_t)blockIdx.x * ws_stride;
  const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
  double* const adjb = ws + (size_t)(ntmax > 1 ? ntmax - 1 : 0) * 8 * smax;
  double* const fac = adjb + (size_t)16 * smax;
  double acc[2] = {0.0, 0.0};
  unsigned long long count = p.ctr[CTR_BIN64T];
  if ((int64_t)count > bin_cap) count = (unsigned long long)bin_cap;
  const PairRef* list = bins + (size_t)NBIN64 * bin_cap;
  const unsigned long long units = (count + PAIRS_PER_UNIT - 1) / PAIRS_PER_UNIT;
  for (;;) {
    unsigned long long u = 0;
    if (lane == 0) u = atomicAdd(&p.ctr[CTR_UNIT64], 1ull);
    u = first_lane(u);
    if (u >= units) break;
    const unsigned long long idx = u * PAIRS_PER_UNIT + g;
    GroupPair gp;
    gp.valid = idx < count;
    const PairRef pr = list[gp.valid ? idx : u * PAIRS_PER_UNIT];
    gp.read = pr.read;
    gp.hap = pr.hap;
    gp.out = pr.out;
    gp.n = (int)(b.read_off[gp.read + 1] - b.read_off[gp.read]);
    gp.m = (int)(b.hap_off[gp.hap + 1] - b.hap_off[gp.hap]);
    const int np0 = gp.n + (ANCH ? 1 : 0);
    if (gp.valid && (gp.m == 0 || np0 > ntmax * GROUP * R)) gp.valid = 0;
    const int mx0 = gp.valid ? gp.m : 0, mn0 = gp.valid ? gp.m : 0x7fffffff;
    const int mx = max(max(lane16(mx0, 0), lane16(mx0, 1)), max(lane16(mx0, 2), lane16(mx0, 3)));
    const int mn = min(min(lane16(mn0, 0), lane16(mn0, 1)), min(lane16(mn0, 2), lane16(mn0, 3)));
    if (mx == 0) continue;
    gp.m_max = mx;
    gp.m_min = mn;
    if (!gp.valid) gp.m = mx;
    const int tend = gp.m_max + 16;
    const int tmain = gp.m_min + 2;
    fill_codes(b, gp, codes, smax, tend + 18);
    const int np = gp.n + (ANCH ? 1 : 0);
    const int ntile_g = gp.valid ? (np + GROUP * R - 1) / (GROUP * R) : 0;
    const int ntiles = max(max(lane16(ntile_g, 0), lane16(ntile_g, 1)), max(lane16(ntile_g, 2), lane16(ntile_g, 3)));
    const int pad = ntiles * GROUP * R - np;
    const int64_t rb = b.read_off[gp.read];
    const double gam = tab->gamma, xx = tab->ext;
    const uint16_t* cl = codes + g * smax + 16 - l;
    auto rowP = [&](int k) { return ws + ((size_t)(k * 2 + 0) * PAIRS_PER_UNIT + g) * smax; };
    auto rowQ = [&](int k) { return ws + ((size_t)(k * 2 + 1) * PAIRS_PER_UNIT + g) * smax; };
    auto adjP = [&](int k) { return adjb + ((size_t)((k & 1) * 2 + 0) * PAIRS_PER_UNIT + g) * smax; };
    auto adjQ = [&](int k) { return adjb + ((size_t)((k & 1) * 2 + 1) * PAIRS_PER_UNIT + g) * smax; };
    // row constants of tile `tile` (as unit64)
    double gd[R], bd[R], mu[R], be[R];
    uint32_t rc[R];
    bool pass[R];
    int qrow[R];
    auto rows = [&](int tile) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int rel = tile * GROUP * R + l * R + r - pad;
        double vg = 0.0, vb = 0.0, vm = 0.0, ve = 0.0;
        uint32_t vc = NOMATCH;
        int qq = -1;
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
            qq = q;
          }
        }
        gd[r] = vg; bd[r] = vb; mu[r] = vm; be[r] = ve; rc[r] = vc;
        pass[r] = rel < 0;
        qrow[r] = qq;
      }
    };
    // ---- phase A: the tiled forward, keeping every tile's last row and factors
    int shift_total = 0;
    double sc1 = 1.0, sc2 = 1.0, tacc = 0.0, rmax = 0.0, latch = 0.0;
    if (l == 0) {
      fac[(0 * PAIRS_PER_UNIT + g) * 2 + 0] = 1.0;
      fac[(0 * PAIRS_PER_UNIT + g) * 2 + 1] = 1.0;
    }
    __syncthreads();
    for (int tile = 0; tile < ntiles; ++tile) {
      const bool first = tile == 0, last = tile == ntiles - 1;
      const double gam_last = (l == GROUP - 1 && last) ? (ANCH ? tab->gap : 0.0) : gam;
      rows(tile);
      const double* inP = first ? nullptr : rowP(tile - 1);
      const double* inQ = first ? nullptr : rowQ(tile - 1);
      double* oP = last ? nullptr : rowP(tile);
      double* oQ = last ? nullptr : rowQ(tile);
      double bmax = 0.0;
#define VD_TILE(F, L)                                                                                             \
  tile64<ANCH, F, L>(gp, cl, tmain, tend, gd, bd, mu, be, rc, gam, gam_last, xx, sc1, sc2, inP, inQ, oP, oQ, tacc, \
                     rmax, latch, bmax)
      if (first && last) VD_TILE(true, true);
      else if (first) VD_TILE(true, false);
      else if (last) VD_TILE(false, true);
      else VD_TILE(false, false);
#undef VD_TILE
      if (!last) {
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
        if (l == 0) {
          fac[((tile + 1) * PAIRS_PER_UNIT + g) * 2 + 0] = sc1;
          fac[((tile + 1) * PAIRS_PER_UNIT + g) * 2 + 1] = sc2;
        }
        __threadfence();
        __syncthreads();
      }
    }
    // terminal, ln L, flag and seed (lane 15 of each group), then the seed to the group
    bool flag = false;
    double seed = 0.0, wpair = 0.0;
    if (l == GROUP - 1 && gp.valid) {
      const double term = ANCH ? latch : tacc;
      const double lnk = (double)(p.read_exp[gp.read] - CEXP64 + shift_total) * LN2;
      double v;
      if (ANCH) {
        const double lc = log(rmax) + LN_1EM300;
        const double ll = log(latch);
        v = (ll > lc ? ll : lc) + lnk;
      } else {
        v = log(tacc) + lnk - log((double)(gp.m + 1));
      }
      flag = !(term > 0.0) || !(term < __builtin_inf());
      if (flag) atomicAdd(&p.ctr[CTR_F64BAD], 1ull);
      out[gp.out] = v;
      wpair = weights ? weights[gp.out] : 1.0;
      seed = flag ? 0.0 : wpair / term;
    }
    seed = __shfl(seed, g * GROUP + GROUP - 1);
    flag = __shfl((int)flag, g * GROUP + GROUP - 1) != 0;
    if (!__builtin_amdgcn_ballot_w64(seed != 0.0)) {  // no pair of the unit contributes: outputs only
      __syncthreads();
      continue;
    }
    // ---- phase B: the tiles last to first
    for (int k = lane; k < 4 * smax; k += 64) {  // both adjoint buffers of this group start at zero
      const int s_ = k / smax, slot = k % smax;
      adjb[((size_t)s_ * PAIRS_PER_UNIT + g) * smax + slot] = 0.0;
    }
    __threadfence();
    __syncthreads();
    GradSums<double> gs{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double Ufin = 0.0, Tfin = 0.0;  // the read's last row: lane 15, register R-1 of the last tile
    for (int tile = ntiles - 1; tile >= 0; --tile) {
      const bool first = tile == 0, last = tile == ntiles - 1;
      const double gam_last = (l == GROUP - 1 && last) ? (ANCH ? tab->gap : 0.0) : gam;
      rows(tile);
      const double* inP = first ? nullptr : rowP(tile - 1);
      const double* inQ = first ? nullptr : rowQ(tile - 1);
      const double f1 = fac[(tile * PAIRS_PER_UNIT + g) * 2 + 0], f2 = fac[(tile * PAIRS_PER_UNIT + g) * 2 + 1];
      const double* bPp = last ? nullptr : adjP(tile + 1);
      const double* bQp = last ? nullptr : adjQ(tile + 1);
      double* tPp = first ? nullptr : adjP(tile);
      double* tQp = first ? nullptr : adjQ(tile);
      double Er[R];
#pragma unroll
      for (int r = 0; r < R; ++r) Er[r] = 0.0;
      const double ts = flag ? 0.0 : seed;
#define VD_TG(F, L)                                                                                                 \
  tile_grad64<ANCH, F, L>(gp, cl, tend, gd, bd, mu, be, rc, pass, gam, gam_last, xx, inP, inQ, f1, f2, bPp, bQp, tPp, \
                          tQp, ts, ck, Er, gs)
      if (first && last) VD_TG(true, true);
      else if (first) VD_TG(true, false);
      else if (last) VD_TG(false, true);
      else VD_TG(false, false);
#undef VD_TG
      // in every tile but the last, lane 15's last register row is an interior row
      if (last && l == GROUP - 1) {
        Ufin = gs.UL;
        Tfin = gs.Tl;
      } else {
        gs.U += gs.UL;
        gs.V += gs.VL;
        gs.Tg += gs.Tl;
      }
      gs.UL = gs.VL = gs.Tl = 0.0;
      if (gp.valid && !flag) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (qrow[r] >= 0) atomicAdd(&lds_de[qrow[r]], Er[r] * gt->dgood_de[qrow[r]] * -1.0);
      }
      __threadfence();
      __syncthreads();
    }
    if (gp.valid && !flag) {
      const double gg = (double)tab->gap, x = (double)tab->ext, a = 1.0 - 2.0 * gg;
      const bool last_lane = l == GROUP - 1;
      const double ww = Ufin;  // the last row's (mu Pb + be Qb) I sum; zero off lane 15
      const double mocc = last_lane ? (double)gp.n * wpair : 0.0;
      acc[0] += -2.0 / a * (mocc - (gs.U + ww)) + gs.V / (gg * a) + ww / gg + gs.Tg * gt->dgam_dg +
                (ANCH ? Tfin : 0.0);
      acc[1] += (gs.U - gs.V) / x - gs.V / (1.0 - x) + gs.Tg * gt->dgam_dx + gs.Tx;
    }
    __syncthreads();  // codes reused by the next unit
  }
  grad_flush(acc[0], acc[1], lds_de, gpart);
}

__global__ void k_zero_f64(double* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

__global__ void k_reset_unit64(unsigned long long* ctr) {
  if (threadIdx.x == 0) ctr[CTR_UNIT64] = 0ull;
}

// Sum of the per-block partial gradients, one block per component (a fixed
// reduction order for a given grid).
__global__ __launch_bounds__(256) void k_grad_reduce(const double* __restrict__ gpart, int nparts,
                                                    double* __restrict__ grad) {
  __shared__ double part[256];
  const int i = blockIdx.x;  // gradient component
  double v = 0.0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) v += gpart[(size_t)k * NGRAD + i];
  part[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[i] = part[0];
}

// FP32: per-row tables, hap codes, then the staged checkpoint block of a
// kernel with r rows per lane (r = 0: the largest class)
size_t lds_grad32(int smax, int r) {
  const int ckw = r > 0 ? 2 * r + 2 : CKW_MAX;
  return 64 * sizeof(double) + ((qtab_lds_bytes<float>() + 15) & ~size_t(15)) + lds_codes_bytes(smax) +
         (size_t)ckw * 64 * sizeof(float);
}
size_t lds_grad64(int smax) {
  return 64 * sizeof(double) + ((qtab_lds_bytes<double>() + 15) & ~size_t(15)) +
         (size_t)PAIRS_PER_UNIT * smax * sizeof(uint16_t);
}

// checkpoint area of one block, in values of the kernel's precision (sized for
// the variant with the shortest checkpoint interval)
size_t grad_ck_stride(int smax, int fp64) {
  const int steps = (fp64 ? GK64 : (GK32 < GK32B ? GK32 : GK32B)) * NSUB;
  return (size_t)(smax / steps + 2) * CKW_MAX * 64;
}

template <bool ANCH>
static hipError_t occ_grad_r(int r, int fp64, int variant, int smax, int* n) {
  switch (fp64 ? r / 2 + 16 : r) {
#define VD_OCC32(RR)                                                                                        \
  case RR:                                                                                                  \
    return (RR <= GRAD_W2_MAX && variant == 0)                                                              \
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_grad32w<ANCH, RR>, 64, lds_grad32(smax, RR)) \
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_grad32<ANCH, RR>, 64, lds_grad32(smax, RR));
    VD_OCC32(1) VD_OCC32(2) VD_OCC32(3) VD_OCC32(4) VD_OCC32(5) VD_OCC32(6) VD_OCC32(7) VD_OCC32(8)
    VD_OCC32(9) VD_OCC32(10) VD_OCC32(11) VD_OCC32(12) VD_OCC32(13) VD_OCC32(14) VD_OCC32(15) VD_OCC32(16)
#undef VD_OCC32
#define VD_OCC64(RR) \
  case 16 + RR / 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_grad64<ANCH, RR>, 64, lds_grad64(smax));
    VD_OCC64(2) VD_OCC64(4) VD_OCC64(6) VD_OCC64(8) VD_OCC64(10) VD_OCC64(12) VD_OCC64(14) VD_OCC64(16)
#undef VD_OCC64
    default: *n = 1; return hipSuccess;
  }
}

// blocks per CU of the FP32 kernel of class r (1..16) in variant `variant`
// (0: two-step register blocks, 1: four-step), or of the FP64 kernel of the bin
// with r rows per lane (r = 2, 4, ..., 16) when fp64
hipError_t grad_occupancy(int anchored, int r, int fp64, int variant, int smax, int* blocks_per_cu) {
  return anchored ? occ_grad_r<true>(r, fp64, variant, smax, blocks_per_cu)
                  : occ_grad_r<false>(r, fp64, variant, smax, blocks_per_cu);
}

template <bool ANCH>
static void launch_grad32_r(int r, int variant, dim3 grid, size_t lds, hipStream_t st, const BatchDev& b,
                            const PlanDev& p,
                            const QTab<float>* tab32, const GradTab* gt, int smax, float* ck, size_t s32,
                            const double* w, double* out, double* gpart) {
  switch (r) {
#define VD_L32(RR)                                                                                          \
  case RR:                                                                                                  \
    if (RR <= GRAD_W2_MAX && variant == 0)                                                                  \
      hipLaunchKernelGGL((k_grad32w<ANCH, RR>), grid, dim3(64), lds, st, b, p, tab32, gt, smax, ck, s32, w, out, \
                         gpart);                                                                            \
    else                                                                                                    \
      hipLaunchKernelGGL((k_grad32<ANCH, RR>), grid, dim3(64), lds, st, b, p, tab32, gt, smax, ck, s32, w, out, \
                         gpart);                                                                            \
    break;
    VD_L32(1) VD_L32(2) VD_L32(3) VD_L32(4) VD_L32(5) VD_L32(6) VD_L32(7) VD_L32(8)
    VD_L32(9) VD_L32(10) VD_L32(11) VD_L32(12) VD_L32(13) VD_L32(14) VD_L32(15) VD_L32(16)
#undef VD_L32
    default: break;
  }
}

template <bool ANCH>
static void launch_grad64_bin(int bin, dim3 grid, size_t lds, hipStream_t st, const BatchDev& b, const PlanDev& p,
                              const QTab<double>* tab64, const GradTab* gt, int smax, double* ck, size_t s64,
                              const PairRef* bins, int64_t bin_cap, const double* w, double* out, double* gpart) {
  switch (bin) {
#define VD_L64(BB)                                                                                                  \
  case BB:                                                                                                          \
    hipLaunchKernelGGL((k_grad64<ANCH, 2 * (BB + 1)>), grid, dim3(64), lds, st, b, p, tab64, gt, smax, ck, s64, bins, \
                       bin_cap, w, out, gpart);                                                                     \
    break;
    VD_L64(0) VD_L64(1) VD_L64(2) VD_L64(3) VD_L64(4) VD_L64(5) VD_L64(6) VD_L64(7)
#undef VD_L64
    default: break;
  }
}

hipError_t grad_tiled_occupancy(int anchored, int smax, int* blocks_per_cu) {
  return anchored ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_grad64t<true>, 64, lds_grad64(smax))
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_grad64t<false>, 64, lds_grad64(smax));
}

hipError_t launch_grad(const BatchDev& b, const PlanDev& p, const int32_t* qE, const QTab<float>* tab32,
                       const QTab<double>* tab64, const GradTab* gt, int anchored, int smax,
                       const int* blocks32, const int* variant32, const int* blocks64, int max_blocks, float* ck32,
                       double* ck64, PairRef* bins, int64_t bin_cap, const double* weights, double* out,
                       double* gpart, double* grad, hipStream_t st, hipEvent_t* ev, hipEvent_t* ev_cls,
                       const GradTiled* tiled) {
  if (ev) (void)hipEventRecord(ev[0], st);
  launch_plan(b, p, qE, tab32, anchored, 0, out, st);
  hipLaunchKernelGGL(k_zero_f64, dim3(64), dim3(256), 0, st, gpart, (int64_t)max_blocks * NGRAD);
  if (ev) (void)hipEventRecord(ev[1], st);
  const size_t s32 = grad_ck_stride(smax, 0), s64 = grad_ck_stride(smax, 1);
  for (int r = MAX_R32; r >= 1; --r) {
    if (ev_cls) (void)hipEventRecord(ev_cls[r], st);
    if (anchored)
      launch_grad32_r<true>(r, variant32[r], dim3(blocks32[r]), lds_grad32(smax, r), st, b, p, tab32, gt, smax, ck32,
                            s32, weights, out, gpart);
    else
      launch_grad32_r<false>(r, variant32[r], dim3(blocks32[r]), lds_grad32(smax, r), st, b, p, tab32, gt, smax, ck32,
                             s32, weights, out, gpart);
  }
  if (ev_cls) (void)hipEventRecord(ev_cls[0], st);
  if (ev) (void)hipEventRecord(ev[2], st);
  if (anchored)
    hipLaunchKernelGGL(k_bin64<true>, dim3(256), dim3(256), 0, st, b, p, bins, bin_cap);
  else
    hipLaunchKernelGGL(k_bin64<false>, dim3(256), dim3(256), 0, st, b, p, bins, bin_cap);
  for (int bin = 0; bin < NBIN64; ++bin) {
    hipLaunchKernelGGL(k_reset_unit64, dim3(1), dim3(64), 0, st, p.ctr);
    if (anchored)
      launch_grad64_bin<true>(bin, dim3(blocks64[bin]), lds_grad64(smax), st, b, p, tab64, gt, smax, ck64, s64, bins,
                              bin_cap, weights, out, gpart);
    else
      launch_grad64_bin<false>(bin, dim3(blocks64[bin]), lds_grad64(smax), st, b, p, tab64, gt, smax, ck64, s64,
                               bins, bin_cap, weights, out, gpart);
  }
  if (tiled && tiled->ntmax > 0) {
    hipLaunchKernelGGL(k_reset_unit64, dim3(1), dim3(64), 0, st, p.ctr);
    if (anchored)
      hipLaunchKernelGGL(k_grad64t<true>, dim3(tiled->blocks), dim3(64), lds_grad64(smax), st, b, p, tab64, gt, smax,
                         tiled->ntmax, ck64, s64, tiled->ws, tiled->ws_stride, (const PairRef*)bins, bin_cap, weights,
                         out, gpart);
    else
      hipLaunchKernelGGL(k_grad64t<false>, dim3(tiled->blocks), dim3(64), lds_grad64(smax), st, b, p, tab64, gt,
                         smax, tiled->ntmax, ck64, s64, tiled->ws, tiled->ws_stride, (const PairRef*)bins, bin_cap,
                         weights, out, gpart);
  }
  hipLaunchKernelGGL(k_grad_reduce, dim3(NGRAD), dim3(256), 0, st, (const double*)gpart, max_blocks, grad);
  if (ev) (void)hipEventRecord(ev[3], st);
  return hipGetLastError();
}

}  // namespace vd

// The plain GPU forward used as a baseline: one read x haplotype pair per
// kernel launch, one workgroup, one thread per read row, one barrier per
// anti-diagonal (the textbook GPU port of the recurrence, without batching,
// persistence or register-resident rows).  Same rescaled recurrence and
// tables as the batched kernels; FP32 with no rescue (it is a timing
// baseline: pairs outside the FP32 window come out as NaN or -inf).
#include "vdhmm_device.h"

namespace vd {

constexpr int PLAIN_THREADS = 256;  // rows per launch (reads up to 256 bases, 255 anchored)

template <bool ANCH>
__global__ __launch_bounds__(PLAIN_THREADS) void k_plain_pair(const uint8_t* __restrict__ read,
                                                              const uint8_t* __restrict__ qual, int n,
                                                              const uint8_t* __restrict__ hap, int m,
                                                              const QTab<float>* __restrict__ tab, double lnk,
                                                              double* __restrict__ out) {
  __shared__ float Pr[2][PLAIN_THREADS], Qr[3][PLAIN_THREADS];
  const int i = threadIdx.x;  // row (anchored: row 0 is the virtual init row)
  const int np = n + (ANCH ? 1 : 0);
  const bool row = i < np;
  float gd = 0.f, bd = 0.f, mu = 0.f, be = 0.f, gam = tab->gamma;
  uint32_t rc = NOMATCH;
  if (row) {
    if (ANCH && i == 0) {
      gd = bd = 1.f;
    } else {
      const int k = i - (ANCH ? 1 : 0);
      int q = qual[k];
      q = q < 63 ? q : 63;
      rc = read[k];
      gd = tab->good[q];
      bd = tab->bad[q];
      mu = tab->mu[q];
      be = tab->beta[q];
      if (k == n - 1) {
        mu = be = tab->cfin[q];
        gam = ANCH ? tab->gap : 0.f;
      }
    }
  }
  for (int k = 0; k < 2; ++k) Pr[k][i] = 0.f;
  for (int k = 0; k < 3; ++k) Qr[k][i] = 0.f;
  __syncthreads();
  const float x = tab->ext, C = 0x1p108f;
  float Ml = 0.f, Dl = 0.f, acc = 0.f, rmax = 0.f, latch = 0.f;
  const int ndiag = np + m + 2;
  for (int d = 0; d < ndiag; ++d) {
    const int c = d - i - 1;  // column of this row on diagonal d (-1 .. m)
    float P = 0.f, Q = 0.f;
    if (row && c >= -1 && c <= m) {
      float pup, qdiag;
      if (i == 0) {  // boundary row above the first row
        if (ANCH) {
          pup = c == -1 ? C : 0.f;
          qdiag = c - 1 == -1 ? C : 0.f;
        } else {
          pup = c >= 0 ? C : 0.f;
          qdiag = c - 1 >= 0 ? C : 0.f;
        }
      } else {
        pup = Pr[(d - 1) & 1][i - 1];
        qdiag = Qr[(d + 1) % 3][i - 1];  // slot of diagonal d-2
      }
      const float D = fmaf(x, Dl, Ml);
      const float e = (c >= 1 && hap[c - 1] == rc) ? gd : bd;
      const float M = e * qdiag;
      P = fmaf(mu, pup, M);
      Q = fmaf(gam, D, fmaf(be, pup, M));
      Ml = M;
      Dl = D;
      if (i == np - 1) {
        if (ANCH) {
          rmax = fmaxf(rmax, Q);
          if (c == m) latch = Q;
        } else {
          acc += P;
        }
      }
    }
    Pr[d & 1][i] = P;
    Qr[d % 3][i] = Q;
    __syncthreads();
  }
  if (i == np - 1) {
    double v;
    if (ANCH) {
      const double lc = log((double)rmax) + LN_1EM300;
      const double ll = log((double)latch);
      v = (ll > lc ? ll : lc) + lnk;
    } else {
      v = log((double)acc) + lnk - log((double)(m + 1));
    }
    *out = v;
  }
}

hipError_t launch_plain_pair(const uint8_t* read, const uint8_t* qual, int n, const uint8_t* hap, int m,
                             const QTab<float>* tab, int anchored, double lnk, double* out, hipStream_t st) {
  if (anchored)
    hipLaunchKernelGGL(k_plain_pair<true>, dim3(1), dim3(PLAIN_THREADS), 0, st, read, qual, n, hap, m, tab, lnk, out);
  else
    hipLaunchKernelGGL(k_plain_pair<false>, dim3(1), dim3(PLAIN_THREADS), 0, st, read, qual, n, hap, m, tab, lnk,
                       out);
  return hipGetLastError();
}

}  // namespace vd

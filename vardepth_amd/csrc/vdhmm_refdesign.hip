// A measurement baseline: the reference's own GPU design, batched, on this
// GPU.  The reference's default CUDA strategy (pairhmm_cuda.cu, dispatched by
// accelerator.py:80-159 with VARDEPTH_CUDA_PRECISION=mixed) gives each
// read x haplotype pair one thread that sweeps the read rows in order, with
// the pair's DP row (M, I, D per haplotype column) in global memory, columns
// interleaved across the pairs of a launch (word j * count + p) so that the
// threads of a wave touch consecutive words; FP32 row values with an FP64 log
// scale; a non-anchored row rescaled by 2^64 when its maximum drops below
// 2^-40, an anchored row normalised by its maximum; a terminal below 2^-40
// reported as NaN (its signal to recompute the pair in FP64).  Pairs are
// grouped by (read, haplotype) length in 32-base buckets and each bucket runs
// in launches bounded by a workspace budget.  The transition and emission
// constants are this library's parameters (the reference's kernel fixes
// g = .001, x = .1).  Not a production path: bench.py times it beside the
// engine on the same pairs.
#include "vdhmm_device.h"

namespace vd {

struct RefParams {
  float err[64];
  float gap, ext;
};

template <bool ANCH>
__global__ __launch_bounds__(256) void k_ref_pairs(BatchDev b, const PairRef* __restrict__ pairs, int count,
                                                   int cols, RefParams prm, float* __restrict__ ws,
                                                   double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  const PairRef pr = pairs[p];
  const int64_t r0 = b.read_off[pr.read], h0 = b.hap_off[pr.hap];
  const int n = (int)(b.read_off[pr.read + 1] - r0), m = (int)(b.hap_off[pr.hap + 1] - h0);
  if (n <= 0 || m <= 0) {
    out[pr.out] = -__builtin_inf();
    return;
  }
  const float g = prm.gap, x = prm.ext, a = 1.f - 2.f * g, c = 1.f - x;
  const float gi = 0.25f * g, xi = 0.25f * x;  // into I: the transition times the insertion emission
  const size_t stride = (size_t)(cols + 1) * count;
  float* pm = ws + p;
  float* pi = pm + stride;
  float* pd = pi + stride;
  const uint8_t* hap = b.hap_bytes + h0;
  {
    float dinit = g;
    for (int j = 0; j <= m; ++j) {
      const size_t k = (size_t)j * count;
      pm[k] = ANCH ? (j == 0 ? 1.f : 0.f) : 1.f / (float)(m + 1);
      pi[k] = 0.f;
      pd[k] = (ANCH && j > 0) ? dinit : 0.f;
      if (ANCH && j > 0) dinit *= x;
    }
  }
  double logs = 0.0;
  for (int i = 0; i < n; ++i) {
    float dm = pm[0], di = pi[0], dd = pd[0];
    const float i0 = gi * dm + xi * di;
    pm[0] = 0.f;
    pi[0] = i0;
    pd[0] = 0.f;
    float curm = 0.f, curd = 0.f, scale = i0;
    int q = b.qual_bytes[r0 + i];
    q = q < 63 ? q : 63;
    const float e = prm.err[q], match = 1.f - e, mis = e / 3.f;
    const uint8_t base = b.read_bytes[r0 + i];
    for (int j = 1; j <= m; ++j) {
      const size_t k = (size_t)j * count;
      const float om = pm[k], oi = pi[k], od = pd[k];
      const float em = base == hap[j - 1] ? match : mis;
      const float nm = em * (a * dm + c * (di + dd));
      const float ni = gi * om + xi * oi;
      const float nd = g * curm + x * curd;
      pm[k] = nm;
      pi[k] = ni;
      pd[k] = nd;
      dm = om;
      di = oi;
      dd = od;
      curm = nm;
      curd = nd;
      const float mass = nm + ni + nd;
      scale = scale > mass ? scale : mass;
    }
    if (!(scale > 0.f)) {
      out[pr.out] = -__builtin_inf();
      return;
    }
    if (ANCH) {
      const float inv = 1.f / scale;
      for (int j = 0; j <= m; ++j) {
        const size_t k = (size_t)j * count;
        pm[k] *= inv;
        pi[k] *= inv;
        pd[k] *= inv;
      }
      logs += log((double)scale);
    } else if (scale < 0x1p-40f) {
      for (int j = 0; j <= m; ++j) {
        const size_t k = (size_t)j * count;
        pm[k] *= 0x1p64f;
        pi[k] *= 0x1p64f;
        pd[k] *= 0x1p64f;
      }
      logs -= 64.0 * LN2;
    }
  }
  double z = 0.0;
  if (ANCH) {
    const size_t k = (size_t)m * count;
    z = (double)pm[k] + (double)pi[k] + (double)pd[k];
  } else {
    for (int j = 0; j <= m; ++j) {
      const size_t k = (size_t)j * count;
      z += (double)pm[k] + (double)pi[k];
    }
  }
  out[pr.out] = z < 0x1p-40 ? __builtin_nan("") : log(z > 1e-300 ? z : 1e-300) + logs;
}

hipError_t launch_ref_pairs(const BatchDev& b, const PairRef* pairs, int count, int cols, const float* err,
                            float gap, float ext, int anchored, float* ws, double* out, hipStream_t st) {
  RefParams prm;
  for (int q = 0; q < 64; ++q) prm.err[q] = err[q];
  prm.gap = gap;
  prm.ext = ext;
  const dim3 grid((count + 255) / 256), block(256);
  if (anchored)
    hipLaunchKernelGGL(k_ref_pairs<true>, grid, block, 0, st, b, pairs, count, cols, prm, ws, out);
  else
    hipLaunchKernelGGL(k_ref_pairs<false>, grid, block, 0, st, b, pairs, count, cols, prm, ws, out);
  return hipGetLastError();
}

}  // namespace vd

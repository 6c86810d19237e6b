// Host side of the MI355X PairHMM engine: the C ABI of include/vdhmm.h.
//
// One process-wide engine (device, stream, grow-only device buffers) behind a
// mutex.  vd_forward_device enqueues the whole batch -- planning, the FP32
// persistent class kernels, FP64 rescue -- on the caller's stream with no host
// round trip (consecutive class kernels alternate with a second stream that
// forks from and joins back into it), so a caller can capture it in a
// hipGraph once its buffers exist.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vdhmm.h"
#include "vdhmm_internal.h"

namespace {

std::mutex g_mu;
std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 2;
}

// Workspace epochs.  A call captured into a HIP graph bakes the addresses of
// the engine's workspaces (plan, trie, suffix, FP64, gradient and genotype
// scratch, counters) and parameter tables into that graph.  The first eager
// call after one or more captured calls advances the epoch: every buffer that
// exists then belongs to the graphs captured so far and is never freed or
// rewritten again, and eager calls allocate buffers of the new epoch (the
// held ones stay allocated until the process ends).  Calls captured back to
// back -- several calls in one graph, or graphs captured with no eager call
// between them -- share one epoch.  A captured call allocates and uploads
// nothing: it runs on the buffers and parameter tables the last eager call
// left, and fails with an error when it would need more.
unsigned g_epoch = 1;            // epoch of the buffers eager calls use
bool g_epoch_captured = false;   // a captured call used the current epoch
bool g_capturing = false;        // the current call is being captured
bool g_capture_miss = false;     // ... and needed a buffer it may not allocate
std::vector<void*> g_held;       // buffers of earlier epochs (held by graphs)

const char* const CAPTURE_MISS =
    "vdhmm: a call under stream capture needs workspaces or parameter tables that the last eager call did not "
    "leave (a larger batch, other parameters, or buffers an earlier capture holds): make the same call eagerly, "
    "uncaptured, right before capturing it";

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      if (g_capture_miss) return fail(CAPTURE_MISS);                                       \
      return fail(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    }                                                                                      \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  unsigned epoch = 0;
  bool current() const { return p && epoch == g_epoch; }
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && current()) return hipSuccess;
    if (g_capturing) {
      g_capture_miss = true;
      return hipErrorStreamCaptureUnsupported;
    }
    if (p) {
      if (epoch != g_epoch) {
        g_held.push_back(p);  // a captured graph may replay into it
      } else {
        hipError_t e = hipFree(p);
        if (e != hipSuccess) return e;
      }
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max<size_t>(bytes + bytes / 4, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) {
      cap = want;
      epoch = g_epoch;
    }
    return e;
  }
  template <typename T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

void open_eager_epoch() {
  if (g_epoch_captured) {
    ++g_epoch;
    g_epoch_captured = false;
  }
}

// The engine mutex for entry points that never run under capture (host
// buffers, synchronous): an eager call after captured ones opens a new epoch.
struct EagerLock {
  std::lock_guard<std::mutex> lk{g_mu};
  EagerLock() { open_eager_epoch(); }
};

// Set for the duration of one device entry point: whether `st` is capturing.
// An eager call after captured ones opens a new epoch first.
struct CaptureScope {
  explicit CaptureScope(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    g_capturing = hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
    g_capture_miss = false;
    if (g_capturing)
      g_epoch_captured = true;
    else
      open_eager_epoch();
  }
  ~CaptureScope() {
    g_capturing = false;
    g_capture_miss = false;
  }
  CaptureScope(const CaptureScope&) = delete;
  CaptureScope& operator=(const CaptureScope&) = delete;
};

struct Engine {
  int dev = -1;
  int cus = 0;
  size_t lds_per_block = 0;  // LDS one workgroup may allocate (device property)
  hipStream_t own_stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // staged inputs / outputs of the host entry points
  DevBuf in_read, in_qual, in_roff, in_hap, in_hoff, in_rreg, in_hreg, in_ooff, in_out;
  // plan scratch
  DevBuf read_cls, read_exp, read_ln32, class_off, read_perm, unit_cnt, unit_bucket, f64_list, ctr, ws64;
  DevBuf f64_sorted, f64_bins;  // the FP64 list binned by shape, and its bin counters
  // parameter tables
  DevBuf qE, tab32, tab64, gtab;
  // gradient scratch: checkpoints, per-block partial gradients, staged weights / gradient
  DevBuf ck32, ck64, gpart, bins64, in_w, in_grad, ws64t;
  // scalar drop-ins: one packed upload (inputs, pair list, exponent sum) and
  // the outputs, staged through one pinned host buffer
  DevBuf in_fast;
  void* pinned = nullptr;
  size_t pinned_cap = 0;
  int32_t E_host[64] = {};  // FP64 exponents of the parameters on the device
  vd_params params_on_device;
  bool have_params = false;
  int occ32[2][17][2] = {}, occ64[2] = {0, 0};  // [anchored][class][unroll 2 / 4]
  int occ_smax = -1;
  int gocc32[2][17][2] = {}, gocc64[2][8] = {};  // [anchored][class][variant]
  vd_tuning tune;
  bool tune_init = false;
  bool class_timing = false;
  hipEvent_t ev_cls[17] = {};
  hipEvent_t ev_last = nullptr;  // end of the previous call's device work
  bool have_last = false;
  double fwd_cls_ms[17] = {}, grad_cls_ms[17] = {};
  int gocc_smax = -1;
  // haplotype prefix sharing of the FP32 forward
  bool share = true;
  bool suffix = true;
  int occ32t[2][17][2] = {};  // [anchored][class][suffix continuation]
  int occt_smax = -1;
  DevBuf trie_order, trie_start, trie_fork, trie_nfork, trie_ws, trie_sfx, sfx_ws;
  // consecutive FP32 class kernels alternate between the caller's stream and
  // this one (vd_class_overlap; VARDEPTH_CLASS_OVERLAP=0 at start-up keeps one)
  bool overlap = true;
  hipStream_t side_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

Engine g;

int engine_init() {
  if (g.dev >= 0) return 0;
  const char* env = std::getenv("VARDEPTH_CUDA_DEVICE");
  int dev = env ? std::atoi(env) : 0;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0)
    return fail(std::string("vdhmm: no HIP device available (") + hipGetErrorString(e) + ")");
  if (dev < 0 || dev >= ndev) return fail("vdhmm: VARDEPTH_CUDA_DEVICE out of range");
  HIPCHK(hipSetDevice(dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  g.cus = prop.multiProcessorCount;
  g.lds_per_block = prop.sharedMemPerBlock;
  HIPCHK(hipStreamCreateWithFlags(&g.own_stream, hipStreamNonBlocking));
  for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&g.ev[i]));
  for (int i = 0; i < 17; ++i) HIPCHK(hipEventCreate(&g.ev_cls[i]));
  HIPCHK(hipEventCreateWithFlags(&g.ev_last, hipEventDisableTiming));
  HIPCHK(hipStreamCreateWithFlags(&g.side_stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&g.ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&g.ev_join, hipEventDisableTiming));
  const char* ov = std::getenv("VARDEPTH_CLASS_OVERLAP");
  if (ov) g.overlap = ov[0] != '0';
  g.dev = dev;
  return 0;
}

// The forward and gradient kernels keep each group's haplotype codes
// (columns -17 .. m+30, 2 bytes each, 4 groups) and the quality tables in LDS,
// so one workgroup's LDS bounds the longest haplotype of a batch.  gfx950
// gives a workgroup 160 KiB; the device's own figure replaces it once the
// engine is initialised.
constexpr size_t LDS_GFX950 = 163840;

int max_hap_len_for(int gradients) {
  const size_t limit = g.lds_per_block ? g.lds_per_block : LDS_GFX950;
  // largest slot count s (a multiple of 8) whose LDS fits, then m = s - 48
  size_t lo = 0, hi = limit;  // lds(s) grows with s
  while (lo < hi) {
    const size_t mid = (lo + hi + 1) / 2;
    const int s = (int)(mid * 8);
    const size_t need = gradients ? std::max(vd::lds_grad32(s), vd::lds_grad64(s))
                                  : std::max(vd::lds_fwd32(s), vd::lds_fwd64(s));
    if (need <= limit) lo = mid; else hi = mid - 1;
  }
  return (int)std::max<int64_t>(0, (int64_t)lo * 8 - 48);
}

int check_hap_len(int max_hap_len, int gradients) {
  const int lim = max_hap_len_for(gradients);
  if (max_hap_len > lim)
    return fail("vdhmm: a haplotype of " + std::to_string(max_hap_len) + " bases exceeds the " +
                std::to_string(lim) + " bases the " + (gradients ? "gradient" : "forward") +
                " kernels hold in one workgroup's LDS; score longer haplotypes in pieces or elsewhere");
  return 0;
}

// Normaliser rho_q of quality bin q (the row constants' comment in
// vdhmm_internal.h): for FP64 the power of two 2^E >= max(1-e, e/3)(1-2g)
// (E is returned either way), for FP32 the largest scaled row factor itself.
double normaliser(const vd_params& p, int q, bool fp32, int* E) {
  const double a = 1 - 2 * p.gap, b = 1 - p.ext, c = 0.25 * p.gap, d = 0.25 * p.ext;
  const double e = p.err[q];
  const double best = std::max(1 - e, e / 3) * a;  // >= a/4 > 0
  int ex = 0;
  std::frexp(best, &ex);  // best = f 2^ex, f in [0.5, 1)
  if (E) *E = ex;
  if (!fp32) return std::ldexp(1.0, ex);
  return std::max(std::max(best, d), std::max(c, b * c / a));
}

template <typename T>
void build_tab(const vd_params& p, vd::QTab<T>& t, int32_t* E) {
  const double a = 1 - 2 * p.gap, b = 1 - p.ext, c = 0.25 * p.gap, d = 0.25 * p.ext;
  for (int q = 0; q < 64; ++q) {
    const double e = p.err[q];
    const double gd = 1 - e, bd = e / 3;
    int ex = 0;
    const double rho = normaliser(p, q, sizeof(T) == 4, &ex);
    t.good[q] = (T)(gd * a / rho);
    t.bad[q] = (T)(bd * a / rho);
    t.mu[q] = (T)(d / rho);
    t.beta[q] = (T)((b * c / a) / rho);
    t.cfin[q] = (T)(c / rho);
    t.lnrho[q] = std::log(rho);
    E[q] = ex;  // the FP64 exponent, whichever precision's table this is
  }
  t.gamma = (T)(b * p.gap / a);
  t.gap = (T)p.gap;
  t.ext = (T)p.ext;
}

// Successive calls share the engine's device workspaces (plan, trie, suffix,
// FP64 and genotype buffers, counters, parameter tables), and a call may run
// on another stream than the previous one: its stream first waits for the
// work the previous call enqueued.  Not under stream capture, where the
// captured work is ordered by the graph and a wait on an event recorded
// outside the capture is not allowed.
struct StreamOrder {
  hipStream_t st;
  bool on = false;
  explicit StreamOrder(hipStream_t s) : st(s) {
    if (!g.ev_last || g_capturing) return;
    on = true;
    if (g.have_last) (void)hipStreamWaitEvent(st, g.ev_last, 0);
  }
  ~StreamOrder() {
    if (on && hipEventRecord(g.ev_last, st) == hipSuccess) g.have_last = true;
  }
  StreamOrder(const StreamOrder&) = delete;
  StreamOrder& operator=(const StreamOrder&) = delete;
};

int check_params(const vd_params& p) {
  if (!(p.gap > 0 && p.gap < 0.5)) return fail("vdhmm: gap open must be in (0, 0.5)");
  if (!(p.ext >= 0 && p.ext < 1)) return fail("vdhmm: gap continuation must be in [0, 1)");
  for (int q = 0; q < 64; ++q)
    if (!(p.err[q] >= 0 && p.err[q] <= 1)) return fail("vdhmm: error probabilities must be in [0, 1]");
  return 0;
}

void build_grad_tab(const vd_params& p, vd::GradTab& t);

// The parameter tables of `p` on the device (scaled FP32 / FP64 emission and
// transition tables, exponents, the gradient's derivative table).  Uploaded
// only when the parameters or the table buffers changed; under capture the
// tables of the last eager call must already hold `p`.
int upload_params(const vd_params& p, hipStream_t st) {
  const bool same = g.have_params && std::memcmp(&p, &g.params_on_device, sizeof(p)) == 0;
  if (same && g.qE.current() && g.tab32.current() && g.tab64.current() && g.gtab.current()) return 0;
  if (g_capturing) return fail(CAPTURE_MISS);
  static vd::QTab<float> t32;
  static vd::QTab<double> t64;
  static vd::GradTab gt;
  static int32_t E[64];
  build_tab(p, t32, E);
  build_tab(p, t64, E);
  build_grad_tab(p, gt);
  std::memcpy(g.E_host, E, sizeof(E));
  HIPCHK(g.qE.ensure(sizeof(E)));
  HIPCHK(g.tab32.ensure(sizeof(t32)));
  HIPCHK(g.tab64.ensure(sizeof(t64)));
  HIPCHK(g.gtab.ensure(sizeof(gt)));
  HIPCHK(hipMemcpyAsync(g.qE.p, E, sizeof(E), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.tab32.p, &t32, sizeof(t32), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.tab64.p, &t64, sizeof(t64), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.gtab.p, &gt, sizeof(gt), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the static host tables are rewritten by the next upload
  g.params_on_device = p;
  g.have_params = true;
  return 0;
}

void tune_defaults() {
  if (g.tune_init) return;
  for (int r = 0; r < 17; ++r) {
    g.tune.fwd_unroll[r] = 2;
    g.tune.fwd_blocks_per_cu[r] = 0;
    g.tune.grad_variant[r] = 0;
    g.tune.grad_blocks_per_cu[r] = 0;
    g.tune.fwd_share[r] = 2;
    g.tune.fwd_route64[r] = 0;
  }
  g.tune.share_chunk = 4;
  g.tune.fwd_pack = 0;
  g.tune_init = true;
}

// Plan routing mask (k_prep_reads): bit 0 sends every read to the FP64 path
// (precision FP64), bit c the reads of class c (vd_tuning.fwd_route64).
int route_mask(int force64) {
  if (force64) return 1;
  int m = 0;
  for (int r = 1; r <= vd::MAX_R32; ++r) m |= g.tune.fwd_route64[r] ? (1 << r) : 0;
  return m;
}

int plan_buffers(const vd_batch* db, int64_t npairs) {
  const int64_t nreads = std::max<int64_t>(db->nreads, 1), nreg = std::max<int64_t>(db->nregions, 1);
  const int64_t nb = db->nregions * vd::NCLS;
  const int64_t max_units = npairs / vd::PAIRS_PER_UNIT + nb + 1;
  HIPCHK(g.read_cls.ensure(nreads));
  HIPCHK(g.read_exp.ensure(nreads * 4));
  HIPCHK(g.read_ln32.ensure(nreads * 8));
  HIPCHK(g.class_off.ensure(nreg * (vd::NCLS + 1) * 4));
  HIPCHK(g.read_perm.ensure(nreads * 8));
  HIPCHK(g.unit_cnt.ensure(vd::unit_cnt_len(nb) * 8));
  // a unit map, or with packed pairs (the gradients) a bucket per FP32 pair
  HIPCHK(g.unit_bucket.ensure(std::max<int64_t>(max_units, npairs + 1) * 4));
  HIPCHK(g.f64_list.ensure(std::max<int64_t>(npairs, 1) * sizeof(vd::PairRef)));
  HIPCHK(g.ctr.ensure(vd::CTR_COUNT * 8));
  return 0;
}

vd::BatchDev batch_dev(const vd_batch* db) {
  return vd::BatchDev{db->read_bytes, db->qual_bytes, db->read_off, db->hap_bytes, db->hap_off,
                      db->region_reads, db->region_haps, db->out_off, db->nregions, db->nreads, db->nhaps};
}

vd::PlanDev plan_dev(int64_t npairs, int class_major, int trie = 0) {
  return vd::PlanDev{g.read_cls.as<int8_t>(), g.read_exp.as<int32_t>(), g.read_ln32.as<double>(),
                     g.class_off.as<int32_t>(),
                     g.read_perm.as<int64_t>(), g.unit_cnt.as<int64_t>(), g.unit_bucket.as<int32_t>(),
                     g.f64_list.as<vd::PairRef>(), g.ctr.as<unsigned long long>(), std::max<int64_t>(npairs, 1),
                     class_major, trie, g.tune.share_chunk, g.trie_order.as<int32_t>(), g.trie_start.as<int32_t>(),
                     g.trie_fork.as<int32_t>(), g.trie_nfork.as<int32_t>(), g.trie_ws.as<float>(),
                     0, 0, g.trie_sfx.as<int32_t>(), g.sfx_ws.as<float>()};
}

int forward_device_locked(const vd_batch* db, const vd_params* params, int mode, int precision, int64_t npairs,
                          int max_hap_len, double* dout, hipStream_t st, vd_stats* stats) {
  if (mode < 0 || mode > 2) return fail("vdhmm: mode must be 0, 1 or 2");
  if (precision != VD_PREC_MIXED && precision != VD_PREC_FP64) return fail("vdhmm: bad precision");
  if (db->nregions < 0 || db->nreads < 0 || db->nhaps < 0 || npairs < 0 || max_hap_len < 0)
    return fail("vdhmm: negative batch size");
  if (int rc = check_params(*params)) return rc;
  if (int rc = engine_init()) return rc;
  if (int rc = check_hap_len(max_hap_len, 0)) return rc;
  StreamOrder order(st);
  if (int rc = upload_params(*params, st)) return rc;
  const int anchored = mode == VD_MODE_ANCHORED ? 1 : 0;
  const int force64 = precision == VD_PREC_FP64 ? 1 : 0;
  // hap-code / boundary slots per group: columns -17 .. m+16
  const int smax = ((max_hap_len + 48) + 7) & ~7;
  if (int rc = plan_buffers(db, npairs)) return rc;
  if (g.occ_smax != smax) {
    for (int a = 0; a < 2; ++a) {
      for (int r = 1; r <= vd::MAX_R32; ++r) {
        HIPCHK(vd::occupancy(a, r, 2, smax, &g.occ32[a][r][0]));
        HIPCHK(vd::occupancy(a, r, 4, smax, &g.occ32[a][r][1]));
      }
      HIPCHK(vd::occupancy(a, 0, 2, smax, &g.occ64[a]));
    }
    g.occ_smax = smax;
  }
  tune_defaults();
  int blocks32[17] = {0}, unroll32[17] = {0};
  for (int r = 1; r <= vd::MAX_R32; ++r) {
    unroll32[r] = g.tune.fwd_unroll[r] == 4 ? 4 : 2;
    const int occ = std::max(1, g.occ32[anchored][r][unroll32[r] == 4 ? 1 : 0]);
    const int want = g.tune.fwd_blocks_per_cu[r];
    blocks32[r] = g.cus * (want > 0 ? std::min(want, occ) : occ);
  }
  const int blocks64 = std::max(1, g.cus * std::max(1, std::min(g.occ64[anchored], 16)));
  HIPCHK(g.ws64.ensure((size_t)blocks64 * 4 * vd::PAIRS_PER_UNIT * smax * sizeof(double)));
  HIPCHK(g.f64_sorted.ensure(std::max<int64_t>(npairs, 1) * sizeof(vd::PairRef)));
  HIPCHK(g.f64_bins.ensure(2 * vd::F64_BINS_TOTAL * sizeof(unsigned long long)));
  // packed pair units (fwd_pack): every class, no prefix sharing
  const int pack = g.tune.fwd_pack && !force64 ? 1 : 0;
  int trie = 0;  // class mask of the read units with prefix sharing
  if (g.share && !force64 && !pack)
    for (int r = 1; r <= vd::MAX_R32; ++r) trie |= g.tune.fwd_share[r] ? (1 << r) : 0;
  int sfx = 0;    // class mask of the read units with the suffix continuation
  int split = 0;  // of those, classes split between the suffix and the prefix-only kernel
  int blocks32s[17] = {0};  // grid of the suffix kernel of a split class
  vd::SideLane lane2{g.side_stream, g.ev_fork, g.ev_join, 0, 0};
  const bool two = g.overlap && g.side_stream;
  const size_t halves = two ? 2 : 1;  // the side stream's kernels use the second half
  if (trie) {
    if (g.occt_smax != smax) {
      for (int a = 0; a < 2; ++a)
        for (int r = 1; r <= vd::MAX_R32; ++r)
          for (int x = 0; x < 2; ++x) HIPCHK(vd::occupancy_trie(a, r, x, smax, &g.occ32t[a][r][x]));
      g.occt_smax = smax;
    }
    // fwd_share 2 (default): the suffix kernel takes every read unit of the
    // class; 3: it takes the units of regions with a shared suffix and the
    // prefix-only kernel, which holds more waves in some classes, the others
    // (each kernel passes over the class's units and skips the other's)
    if (g.suffix)
      for (int r = 1; r <= vd::MAX_R32; ++r) {
        const int fs = g.tune.fwd_share[r];
        sfx |= fs >= 2 ? (1 << r) : 0;
        split |= fs == 3 ? (1 << r) : 0;
      }
    size_t ws_floats = 0, sfx_floats = 0;
    for (int r = 1; r <= vd::MAX_R32; ++r) {
      if (!((trie >> r) & 1)) continue;
      const bool s_only = ((sfx >> r) & 1) && !((split >> r) & 1);
      const int want = g.tune.fwd_blocks_per_cu[r];
      const int occ = std::max(1, g.occ32t[anchored][r][s_only ? 1 : 0]);
      blocks32[r] = g.cus * (want > 0 ? std::min(want, occ) : occ);
      const int occs = std::max(1, g.occ32t[anchored][r][1]);
      blocks32s[r] = (split >> r) & 1 ? g.cus * (want > 0 ? std::min(want, occs) : occs) : 0;
      const int maxb = std::max(blocks32[r], blocks32s[r]);
      ws_floats = std::max(ws_floats, (size_t)maxb * vd::PAIRS_PER_UNIT * vd::TRIE_SLOTS * vd::GROUP *
                                          ((3 * r + 3 + 3) & ~3));
      sfx_floats = std::max(sfx_floats, (size_t)(s_only ? blocks32[r] : blocks32s[r]) * vd::SFX_WS);
    }
    const int64_t nh = std::max<int64_t>(db->nhaps, 1), nreg = std::max<int64_t>(db->nregions, 1);
    HIPCHK(g.trie_order.ensure(nh * 4));
    HIPCHK(g.trie_start.ensure(nh * 4));
    HIPCHK(g.trie_fork.ensure(nh * 4));
    HIPCHK(g.trie_nfork.ensure(nreg * 4));
    HIPCHK(g.trie_ws.ensure(halves * ws_floats * sizeof(float)));
    HIPCHK(g.trie_sfx.ensure(nreg * 4));
    if (sfx) HIPCHK(g.sfx_ws.ensure(halves * sfx_floats * sizeof(float)));
    lane2.trie_half = ws_floats;
    lane2.sfx_half = sfx ? sfx_floats : 0;
  }
  const vd::BatchDev b = batch_dev(db);
  vd::PlanDev p = plan_dev(npairs, 1, trie);
  p.sfx = sfx;
  p.split = split;
  p.pack_pairs = pack;
  const bool timing = stats && g.class_timing;
  HIPCHK(vd::launch_pipeline(b, p, g.qE.as<int32_t>(), g.tab32.as<vd::QTab<float>>(),
                             g.tab64.as<vd::QTab<double>>(), anchored, route_mask(force64), smax, blocks32, blocks32s,
                             unroll32,
                             blocks64, g.ws64.as<double>(), g.f64_sorted.as<vd::PairRef>(),
                             g.f64_bins.as<unsigned long long>(), dout, st, stats ? g.ev : nullptr,
                             timing ? g.ev_cls : nullptr, two ? &lane2 : nullptr));
  if (timing && !force64) {
    HIPCHK(hipEventSynchronize(g.ev[3]));
    for (int r = 1; r <= vd::MAX_R32; ++r) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, g.ev_cls[r], g.ev_cls[r - 1]));
      g.fwd_cls_ms[r] = ms;
    }
  }
  if (stats) {
    HIPCHK(hipEventSynchronize(g.ev[3]));
    unsigned long long c[vd::CTR_COUNT];
    HIPCHK(hipMemcpy(c, g.ctr.p, sizeof(c), hipMemcpyDeviceToHost));
    float ms = 0, a = 0, b2 = 0, c3 = 0;
    HIPCHK(hipEventElapsedTime(&ms, g.ev[0], g.ev[3]));
    HIPCHK(hipEventElapsedTime(&a, g.ev[0], g.ev[1]));
    HIPCHK(hipEventElapsedTime(&b2, g.ev[1], g.ev[2]));
    HIPCHK(hipEventElapsedTime(&c3, g.ev[2], g.ev[3]));
    stats->plan_ms = a;
    stats->fwd32_ms = b2;
    stats->fwd64_ms = c3;
    stats->pairs = (uint64_t)npairs;
    // packed plans count pairs
    stats->fp32_units = force64 ? 0 : pack ? (c[vd::CTR_NUNITS] + vd::PAIRS_PER_UNIT - 1) / vd::PAIRS_PER_UNIT
                                           : c[vd::CTR_NUNITS];
    stats->flagged = c[vd::CTR_FLAGGED];
    stats->fp64_pairs = std::min<unsigned long long>(c[vd::CTR_F64], (unsigned long long)std::max<int64_t>(npairs, 1));
    stats->fp64_bad = c[vd::CTR_F64BAD];
    stats->overflow = c[vd::CTR_OVERFLOW];
    stats->device_ms = ms;
    stats->fp32_cells_swept = force64 ? 0 : c[vd::CTR_SWEPT];
    if (c[vd::CTR_OVERFLOW]) return fail("vdhmm: FP64 work list overflow");
  }
  return 0;
}

// The forward of vd_forward_device with every scratch buffer carved from the
// library's static device array (vdhmm_noalloc.hip): plan scratch, FP64 work
// list, counters, parameter tables and the FP64 tile rows.  No allocation, no
// host-to-device copy (the tables go as kernel arguments), no prefix sharing
// (its wavefront workspace scales with the grid) and one stream; the FP64
// rescue grid shrinks to the scratch left after the plan.
int forward_noalloc_locked(const vd_batch* db, const vd_params* params, int mode, int precision, int64_t npairs,
                           int max_hap_len, double* dout, hipStream_t st) {
  if (mode < 0 || mode > 2) return fail("vdhmm: mode must be 0, 1 or 2");
  if (precision != VD_PREC_MIXED && precision != VD_PREC_FP64) return fail("vdhmm: bad precision");
  if (db->nregions < 0 || db->nreads < 0 || db->nhaps < 0 || npairs < 0 || max_hap_len < 0)
    return fail("vdhmm: negative batch size");
  if (int rc = check_params(*params)) return rc;
  if (int rc = engine_init()) return rc;
  if (int rc = check_hap_len(max_hap_len, 0)) return rc;
  void* base = nullptr;
  size_t cap = 0;
  HIPCHK(vd::noalloc_scratch(g.dev, &base, &cap));
  const int anchored = mode == VD_MODE_ANCHORED ? 1 : 0;
  const int force64 = precision == VD_PREC_FP64 ? 1 : 0;
  const int smax = ((max_hap_len + 48) + 7) & ~7;
  const int64_t nreads = std::max<int64_t>(db->nreads, 1), nreg = std::max<int64_t>(db->nregions, 1);
  const int64_t nb = db->nregions * vd::NCLS;
  const int64_t max_units = npairs / vd::PAIRS_PER_UNIT + nb + 1;
  // layout: 256-byte aligned slices, in order
  size_t off = 0;
  auto slice = [&](size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) & ~size_t(255);
    return at;
  };
  const size_t o_qE = slice(64 * sizeof(int32_t)), o_t32 = slice(sizeof(vd::QTab<float>)),
               o_t64 = slice(sizeof(vd::QTab<double>)), o_cls = slice(nreads), o_exp = slice(nreads * 4),
               o_ln32 = slice(nreads * 8),
               o_coff = slice(nreg * (vd::NCLS + 1) * 4), o_perm = slice(nreads * 8),
               o_ucnt = slice(vd::unit_cnt_len(nb) * 8), o_ubkt = slice(max_units * 4), o_f64 = slice(std::max<int64_t>(npairs, 1) * sizeof(vd::PairRef)),
               o_ctr = slice(vd::CTR_COUNT * 8);
  const size_t ws_block = (size_t)4 * vd::PAIRS_PER_UNIT * smax * sizeof(double);
  if (off + ws_block > cap)
    return fail("vdhmm: this batch needs " + std::to_string(off + ws_block) +
                " bytes of scratch (21 per read, 17 per pair, 204 per region and one FP64 tile workspace of 128 per "
                "haplotype slot); the no-allocation forward holds " + std::to_string(cap));
  if (g.occ_smax != smax) {
    for (int a = 0; a < 2; ++a) {
      for (int r = 1; r <= vd::MAX_R32; ++r) {
        HIPCHK(vd::occupancy(a, r, 2, smax, &g.occ32[a][r][0]));
        HIPCHK(vd::occupancy(a, r, 4, smax, &g.occ32[a][r][1]));
      }
      HIPCHK(vd::occupancy(a, 0, 2, smax, &g.occ64[a]));
    }
    g.occ_smax = smax;
  }
  tune_defaults();
  int blocks32[17] = {0}, unroll32[17] = {0};
  for (int r = 1; r <= vd::MAX_R32; ++r) {
    unroll32[r] = g.tune.fwd_unroll[r] == 4 ? 4 : 2;
    const int occ = std::max(1, g.occ32[anchored][r][unroll32[r] == 4 ? 1 : 0]);
    const int want = g.tune.fwd_blocks_per_cu[r];
    blocks32[r] = g.cus * (want > 0 ? std::min(want, occ) : occ);
  }
  const int64_t fit = (int64_t)((cap - off) / ws_block);
  const int blocks64 = (int)std::max<int64_t>(
      1, std::min<int64_t>(fit, (int64_t)g.cus * std::max(1, std::min(g.occ64[anchored], 16))));
  const size_t o_ws = slice((size_t)blocks64 * ws_block);
  char* B = static_cast<char*>(base);
  static vd::QTab<float> t32;
  static vd::QTab<double> t64;
  static int32_t E[64];
  build_tab(*params, t32, E);
  build_tab(*params, t64, E);
  auto* dE = reinterpret_cast<int32_t*>(B + o_qE);
  auto* d32 = reinterpret_cast<vd::QTab<float>*>(B + o_t32);
  auto* d64 = reinterpret_cast<vd::QTab<double>*>(B + o_t64);
  HIPCHK(vd::launch_put_tables(t32, t64, E, d32, d64, dE, st));
  vd::PlanDev p{};
  p.read_cls = reinterpret_cast<int8_t*>(B + o_cls);
  p.read_exp = reinterpret_cast<int32_t*>(B + o_exp);
  p.read_ln32 = reinterpret_cast<double*>(B + o_ln32);
  p.class_off = reinterpret_cast<int32_t*>(B + o_coff);
  p.read_perm = reinterpret_cast<int64_t*>(B + o_perm);
  p.unit_cnt = reinterpret_cast<int64_t*>(B + o_ucnt);
  p.unit_bucket = reinterpret_cast<int32_t*>(B + o_ubkt);
  p.f64_list = reinterpret_cast<vd::PairRef*>(B + o_f64);
  p.ctr = reinterpret_cast<unsigned long long*>(B + o_ctr);
  p.f64_cap = std::max<int64_t>(npairs, 1);
  p.class_major = 1;
  p.trie = 0;
  p.trie_chunk = g.tune.share_chunk;
  HIPCHK(vd::launch_pipeline(batch_dev(db), p, dE, d32, d64, anchored, route_mask(force64), smax, blocks32, blocks32,
                             unroll32,
                             blocks64, reinterpret_cast<double*>(B + o_ws), nullptr, nullptr, dout, st, nullptr,
                             nullptr, nullptr));
  return 0;
}

int validate_host_batch(const vd_batch* b, int64_t* npairs, int* max_hap, uint64_t* cells) {
  if (b->nregions < 0 || b->nreads < 0 || b->nhaps < 0) return fail("vdhmm: negative batch size");
  if (b->read_off[0] != 0 || b->hap_off[0] != 0 || b->region_reads[0] != 0 || b->region_haps[0] != 0 ||
      b->out_off[0] != 0)
    return fail("vdhmm: offsets must start at 0");
  for (int64_t i = 0; i < b->nreads; ++i)
    if (b->read_off[i + 1] < b->read_off[i]) return fail("vdhmm: read offsets must be non-decreasing");
  int mh = 0;
  for (int64_t i = 0; i < b->nhaps; ++i) {
    int64_t len = b->hap_off[i + 1] - b->hap_off[i];
    if (len < 0) return fail("vdhmm: haplotype offsets must be non-decreasing");
    if (len > (1 << 20)) return fail("vdhmm: haplotype longer than 2^20 bases");
    mh = std::max<int>(mh, (int)len);
  }
  uint64_t cl = 0;
  for (int64_t k = 0; k < b->nregions; ++k) {
    int64_t r0 = b->region_reads[k], r1 = b->region_reads[k + 1];
    int64_t h0 = b->region_haps[k], h1 = b->region_haps[k + 1];
    if (r1 < r0 || h1 < h0 || r1 > b->nreads || h1 > b->nhaps) return fail("vdhmm: bad region ranges");
    if (b->out_off[k + 1] - b->out_off[k] != (r1 - r0) * (h1 - h0))
      return fail("vdhmm: out_off does not match region sizes");
    cl += (uint64_t)(b->read_off[r1] - b->read_off[r0]) * (uint64_t)(b->hap_off[h1] - b->hap_off[h0]);
  }
  if (b->region_reads[b->nregions] != b->nreads || b->region_haps[b->nregions] != b->nhaps)
    return fail("vdhmm: regions must cover all reads and haplotypes");
  *npairs = b->out_off[b->nregions];
  *max_hap = mh;
  *cells = cl;
  return 0;
}

int stage_host_batch(const vd_batch* hb, int64_t npairs, hipStream_t st) {
  const int64_t nrb = hb->read_off[hb->nreads], nhb = hb->hap_off[hb->nhaps];
  HIPCHK(g.in_read.ensure(nrb + 1));
  HIPCHK(g.in_qual.ensure(nrb + 1));
  HIPCHK(g.in_roff.ensure((hb->nreads + 1) * 8));
  HIPCHK(g.in_hap.ensure(nhb + 1));
  HIPCHK(g.in_hoff.ensure((hb->nhaps + 1) * 8));
  HIPCHK(g.in_rreg.ensure((hb->nregions + 1) * 8));
  HIPCHK(g.in_hreg.ensure((hb->nregions + 1) * 8));
  HIPCHK(g.in_ooff.ensure((hb->nregions + 1) * 8));
  HIPCHK(g.in_out.ensure(std::max<int64_t>(npairs, 1) * 8));
  if (nrb) {
    HIPCHK(hipMemcpyAsync(g.in_read.p, hb->read_bytes, nrb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(g.in_qual.p, hb->qual_bytes, nrb, hipMemcpyHostToDevice, st));
  }
  if (nhb) HIPCHK(hipMemcpyAsync(g.in_hap.p, hb->hap_bytes, nhb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.in_roff.p, hb->read_off, (hb->nreads + 1) * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.in_hoff.p, hb->hap_off, (hb->nhaps + 1) * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.in_rreg.p, hb->region_reads, (hb->nregions + 1) * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.in_hreg.p, hb->region_haps, (hb->nregions + 1) * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g.in_ooff.p, hb->out_off, (hb->nregions + 1) * 8, hipMemcpyHostToDevice, st));
  return 0;
}

vd_batch staged_view(const vd_batch* hb) {
  return vd_batch{g.in_read.as<uint8_t>(), g.in_qual.as<uint8_t>(), g.in_roff.as<int64_t>(), g.in_hap.as<uint8_t>(),
                  g.in_hoff.as<int64_t>(), g.in_rreg.as<int64_t>(), g.in_hreg.as<int64_t>(),
                  g.in_ooff.as<int64_t>(), hb->nregions, hb->nreads, hb->nhaps};
}

int forward_host_locked(const vd_batch* hb, const vd_params* params, int mode, int precision, double* out,
                        vd_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  int64_t npairs = 0;
  int max_hap = 0;
  uint64_t cells = 0;
  if (int rc = validate_host_batch(hb, &npairs, &max_hap, &cells)) return rc;
  if (int rc = check_hap_len(max_hap, 0)) return rc;
  if (int rc = engine_init()) return rc;
  hipStream_t st = g.own_stream;
  StreamOrder order(st);
  if (int rc = stage_host_batch(hb, npairs, st)) return rc;
  vd_batch db = staged_view(hb);
  vd_stats local;
  vd_stats* s = stats ? stats : &local;
  std::memset(s, 0, sizeof(*s));
  if (int rc = forward_device_locked(&db, params, mode, precision, npairs, max_hap, g.in_out.as<double>(), st, s))
    return rc;
  if (npairs) HIPCHK(hipMemcpyAsync(out, g.in_out.p, npairs * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  s->cells = cells;
  s->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (s->fp64_bad) return fail("vdhmm: FP64 pass produced non-finite likelihoods");
  return 0;
}

// ---- gradients -------------------------------------------------------------

void build_grad_tab(const vd_params& p, vd::GradTab& t) {
  const double g = p.gap, x = p.ext, a = 1 - 2 * g;
  for (int q = 0; q < 64; ++q) {
    const double e = p.err[q];
    const double rho = normaliser(p, q, false, nullptr);
    t.dgood_de32[q] = -a / normaliser(p, q, true, nullptr);
    t.dgood_dg[q] = -2 * (1 - e) / rho;
    t.dbad_dg[q] = -2 * (e / 3) / rho;
    t.dgood_de[q] = -a / rho;
    t.dbad_de[q] = a / (3 * rho);
    t.dmu_dx[q] = 0.25 / rho;
    t.dbeta_dx[q] = -g / (4 * a * rho);
    t.dbeta_dg[q] = (1 - x) / (4 * rho * a * a);
    t.dcfin_dg[q] = 0.25 / rho;
  }
  t.dgam_dg = (1 - x) / (a * a);
  t.dgam_dx = -g / a;
}

int grad_device_locked(const vd_batch* db, const vd_params* params, int mode, int64_t npairs, int max_hap_len,
                       int max_read_len, const double* dweights, double* dout, double* dgrad, hipStream_t st,
                       vd_stats* stats) {
  if (mode < 0 || mode > 2) return fail("vdhmm: mode must be 0, 1 or 2");
  if (db->nregions < 0 || db->nreads < 0 || db->nhaps < 0 || npairs < 0 || max_hap_len < 0 || max_read_len < 0)
    return fail("vdhmm: negative batch size");
  const int anchored = mode == VD_MODE_ANCHORED ? 1 : 0;
  if (int rc = check_params(*params)) return rc;
  if (!(params->ext > 0)) return fail("vdhmm: gradients need a gap continuation in (0, 1)");
  if (int rc = engine_init()) return rc;
  if (int rc = check_hap_len(max_hap_len, 1)) return rc;
  StreamOrder order(st);
  if (int rc = upload_params(*params, st)) return rc;
  const int smax = ((max_hap_len + 48) + 7) & ~7;
  if (int rc = plan_buffers(db, npairs)) return rc;
  if (g.gocc_smax != smax) {
    for (int a = 0; a < 2; ++a) {
      for (int r = 1; r <= vd::MAX_R32; ++r)
        for (int v = 0; v < 2; ++v) HIPCHK(vd::grad_occupancy(a, r, 0, v, smax, &g.gocc32[a][r][v]));
      for (int k = 0; k < 8; ++k) HIPCHK(vd::grad_occupancy(a, 2 * (k + 1), 1, 0, smax, &g.gocc64[a][k]));
    }
    g.gocc_smax = smax;
  }
  tune_defaults();
  // per-class persistent grids; the checkpoint and partial-gradient areas are
  // sized for the widest of them
  int blocks32[17] = {0}, variant32[17] = {0}, blocks64[8] = {0}, max_blocks = 1;
  for (int r = 1; r <= vd::MAX_R32; ++r) {
    variant32[r] = g.tune.grad_variant[r] == 1 ? 1 : 0;
    const int occ = std::max(1, g.gocc32[anchored][r][variant32[r]]);
    const int want = g.tune.grad_blocks_per_cu[r];
    blocks32[r] = g.cus * (want > 0 ? std::min(want, occ) : occ);
    max_blocks = std::max(max_blocks, blocks32[r]);
  }
  int max64 = 1;
  for (int k = 0; k < 8; ++k) {
    blocks64[k] = std::max(1, g.cus * std::max(1, std::min(g.gocc64[anchored][k], 16)));
    max64 = std::max(max64, blocks64[k]);
  }
  max_blocks = std::max(max_blocks, max64);
  const int64_t bin_cap = std::max<int64_t>(npairs, 1);
  // reads beyond 256 rows: the tiled FP64 adjoint, 64-row tiles, its
  // persistent grid bounded by the FP64 kernels' (checkpoint areas) and by
  // a 2 GiB tile workspace
  vd::GradTiled tiled{0, 0, nullptr, 0};
  if (max_read_len + anchored > vd::GROUP * vd::MAX_R32) {
    tiled.ntmax = (max_read_len + anchored + vd::GROUP * vd::R64 - 1) / (vd::GROUP * vd::R64);
    tiled.ws_stride = vd::grad_tile_ws(smax, tiled.ntmax);
    int occ = 1;
    HIPCHK(vd::grad_tiled_occupancy(anchored, smax, &occ));
    const size_t cap_ws = ((size_t)2 << 30) / (tiled.ws_stride * sizeof(double));
    tiled.blocks = (int)std::max<size_t>(
        1, std::min<size_t>({(size_t)g.cus * std::max(1, std::min(occ, 16)), (size_t)max64, cap_ws}));
    HIPCHK(g.ws64t.ensure((size_t)tiled.blocks * tiled.ws_stride * sizeof(double)));
    tiled.ws = g.ws64t.as<double>();
  }
  HIPCHK(g.ck32.ensure((size_t)max_blocks * vd::grad_ck_stride(smax, 0) * sizeof(float)));
  HIPCHK(g.ck64.ensure((size_t)max64 * vd::grad_ck_stride(smax, 1) * sizeof(double)));
  HIPCHK(g.gpart.ensure((size_t)max_blocks * vd::NGRAD * sizeof(double)));
  HIPCHK(g.bins64.ensure((size_t)9 * bin_cap * sizeof(vd::PairRef)));  // 8 row bins + the tiled bin
  const vd::BatchDev b = batch_dev(db);
  vd::PlanDev p = plan_dev(npairs, 1);
  p.pack_pairs = 1;  // pair units may span regions of a class: no idle lanes for small regions
  HIPCHK(vd::launch_grad(b, p, g.qE.as<int32_t>(), g.tab32.as<vd::QTab<float>>(), g.tab64.as<vd::QTab<double>>(),
                         g.gtab.as<vd::GradTab>(), anchored, smax, blocks32, variant32, blocks64, max_blocks,
                         g.ck32.as<float>(), g.ck64.as<double>(), g.bins64.as<vd::PairRef>(), bin_cap, dweights,
                         dout, g.gpart.as<double>(), dgrad, st, stats ? g.ev : nullptr,
                         stats && g.class_timing ? g.ev_cls : nullptr, &tiled));
  if (stats && g.class_timing) {
    HIPCHK(hipEventSynchronize(g.ev[3]));
    for (int r = 1; r <= vd::MAX_R32; ++r) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, g.ev_cls[r], g.ev_cls[r - 1]));
      g.grad_cls_ms[r] = ms;
    }
  }
  if (stats) {
    HIPCHK(hipEventSynchronize(g.ev[3]));
    unsigned long long c[vd::CTR_COUNT];
    HIPCHK(hipMemcpy(c, g.ctr.p, sizeof(c), hipMemcpyDeviceToHost));
    float ms = 0, a = 0, b2 = 0, c3 = 0;
    HIPCHK(hipEventElapsedTime(&ms, g.ev[0], g.ev[3]));
    HIPCHK(hipEventElapsedTime(&a, g.ev[0], g.ev[1]));
    HIPCHK(hipEventElapsedTime(&b2, g.ev[1], g.ev[2]));
    HIPCHK(hipEventElapsedTime(&c3, g.ev[2], g.ev[3]));
    stats->plan_ms = a;
    stats->fwd32_ms = b2;
    stats->fwd64_ms = c3;
    stats->pairs = (uint64_t)npairs;
    stats->fp32_units = (c[vd::CTR_NUNITS] + vd::PAIRS_PER_UNIT - 1) / vd::PAIRS_PER_UNIT;  // packed: FP32 pairs / 4
    stats->flagged = c[vd::CTR_FLAGGED];
    stats->fp64_pairs = std::min<unsigned long long>(c[vd::CTR_F64], (unsigned long long)std::max<int64_t>(npairs, 1));
    stats->fp64_bad = c[vd::CTR_F64BAD];
    stats->overflow = c[vd::CTR_OVERFLOW];
    stats->device_ms = ms;
    stats->fp32_cells_swept = c[vd::CTR_SWEPT];
    if (c[vd::CTR_OVERFLOW]) return fail("vdhmm: FP64 work list overflow");
  }
  return 0;
}

int grad_host_locked(const vd_batch* hb, const vd_params* params, int mode, const double* weights, double* out,
                     double* grad, vd_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  int64_t npairs = 0;
  int max_hap = 0;
  uint64_t cells = 0;
  if (int rc = validate_host_batch(hb, &npairs, &max_hap, &cells)) return rc;
  int max_read = 0;
  for (int64_t i = 0; i < hb->nreads; ++i)
    max_read = std::max<int>(max_read, (int)std::min<int64_t>(hb->read_off[i + 1] - hb->read_off[i], 1 << 30));
  if (mode < 0 || mode > 2) return fail("vdhmm: mode must be 0, 1 or 2");
  if (int rc = check_params(*params)) return rc;
  if (!(params->ext > 0)) return fail("vdhmm: gradients need a gap continuation in (0, 1)");
  if (int rc = check_hap_len(max_hap, 1)) return rc;
  if (int rc = engine_init()) return rc;
  hipStream_t st = g.own_stream;
  StreamOrder order(st);
  if (int rc = stage_host_batch(hb, npairs, st)) return rc;
  HIPCHK(g.in_grad.ensure(vd::NGRAD * sizeof(double)));
  const double* dw = nullptr;
  if (weights) {
    HIPCHK(g.in_w.ensure(std::max<int64_t>(npairs, 1) * 8));
    if (npairs) HIPCHK(hipMemcpyAsync(g.in_w.p, weights, npairs * 8, hipMemcpyHostToDevice, st));
    dw = g.in_w.as<double>();
  }
  vd_batch db = staged_view(hb);
  vd_stats local;
  vd_stats* s = stats ? stats : &local;
  std::memset(s, 0, sizeof(*s));
  if (int rc = grad_device_locked(&db, params, mode, npairs, max_hap, max_read, dw, g.in_out.as<double>(),
                                  g.in_grad.as<double>(), st, s))
    return rc;
  if (out && npairs) HIPCHK(hipMemcpyAsync(out, g.in_out.p, npairs * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(grad, g.in_grad.p, vd::NGRAD * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  s->cells = cells;
  s->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (s->fp64_bad) return fail("vdhmm: FP64 gradient pass met non-finite likelihoods");
  return 0;
}

// ---- plain baseline: one pair per launch ---------------------------------

int plain_host_locked(const vd_batch* hb, const vd_params* params, int mode, int64_t max_pairs, double* out,
                      double* gpu_ms) {
  int64_t npairs = 0;
  int max_hap = 0;
  uint64_t cells = 0;
  if (int rc = validate_host_batch(hb, &npairs, &max_hap, &cells)) return rc;
  if (mode < 0 || mode > 2) return fail("vdhmm: mode must be 0, 1 or 2");
  const int anchored = mode == VD_MODE_ANCHORED ? 1 : 0;
  if (int rc = check_params(*params)) return rc;
  if (int rc = engine_init()) return rc;
  hipStream_t st = g.own_stream;
  StreamOrder order(st);
  if (int rc = upload_params(*params, st)) return rc;
  if (int rc = stage_host_batch(hb, npairs, st)) return rc;
  static vd::QTab<float> t32;
  static int32_t E[64];
  build_tab(*params, t32, E);
  const int64_t todo = std::min<int64_t>(std::max<int64_t>(max_pairs, 0), npairs);
  std::vector<double> lnk(hb->nreads);
  for (int64_t r = 0; r < hb->nreads; ++r) {
    double e = 0;
    for (int64_t k = hb->read_off[r]; k < hb->read_off[r + 1]; ++k) e += t32.lnrho[std::min<int>(hb->qual_bytes[k], 63)];
    lnk[r] = e - vd::CEXP32 * 0.69314718055994530942;
  }
  const uint8_t* dread = g.in_read.as<uint8_t>();
  const uint8_t* dqual = g.in_qual.as<uint8_t>();
  const uint8_t* dhap = g.in_hap.as<uint8_t>();
  double* dout = g.in_out.as<double>();
  std::vector<char> state(std::max<int64_t>(npairs, 1), 0);  // 1 = launched, 2 = empty pair
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipEventRecord(g.ev[0], st));
  int64_t done = 0;
  for (int64_t k = 0; k < hb->nregions && done < todo; ++k) {
    const int64_t r0 = hb->region_reads[k], r1 = hb->region_reads[k + 1];
    const int64_t h0 = hb->region_haps[k], h1 = hb->region_haps[k + 1];
    for (int64_t r = r0; r < r1 && done < todo; ++r)
      for (int64_t h = h0; h < h1 && done < todo; ++h, ++done) {
        const int n = (int)(hb->read_off[r + 1] - hb->read_off[r]);
        const int m = (int)(hb->hap_off[h + 1] - hb->hap_off[h]);
        const int64_t o = hb->out_off[k] + (r - r0) * (h1 - h0) + (h - h0);
        if (n == 0 || m == 0) {
          state[o] = 2;
          continue;
        }
        if (n + anchored > 256) return fail("vdhmm: the plain baseline handles reads of at most 256 rows");
        HIPCHK(vd::launch_plain_pair(dread + hb->read_off[r], dqual + hb->read_off[r], n, dhap + hb->hap_off[h], m,
                                     g.tab32.as<vd::QTab<float>>(), anchored, lnk[r], dout + o, st));
        state[o] = 1;
      }
  }
  HIPCHK(hipEventRecord(g.ev[1], st));
  HIPCHK(hipEventSynchronize(g.ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, g.ev[0], g.ev[1]));
  if (gpu_ms) *gpu_ms = ms;
  std::vector<double> tmp(std::max<int64_t>(npairs, 1));
  if (npairs) HIPCHK(hipMemcpy(tmp.data(), dout, npairs * 8, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < npairs; ++i) out[i] = state[i] == 1 ? tmp[i] : (state[i] == 2 ? -INFINITY : NAN);
  return 0;
}

// ---- reference-design baseline: one thread per pair, rows in global memory --

int refdesign_host_locked(const vd_batch* hb, const vd_params* params, int mode, int64_t budget_mb, double* out,
                          double* gpu_ms, int64_t* launches) {
  int64_t npairs = 0;
  int max_hap = 0;
  uint64_t cells = 0;
  if (int rc = validate_host_batch(hb, &npairs, &max_hap, &cells)) return rc;
  if (mode < 0 || mode > 2) return fail("vdhmm: mode must be 0, 1 or 2");
  if (budget_mb < 8) return fail("vdhmm: the workspace budget must be at least 8 MiB");
  const int anchored = mode == VD_MODE_ANCHORED ? 1 : 0;
  if (int rc = check_params(*params)) return rc;
  if (int rc = engine_init()) return rc;
  hipStream_t st = g.own_stream;
  StreamOrder order(st);
  if (int rc = stage_host_batch(hb, npairs, st)) return rc;
  // pairs grouped by (read, haplotype) length in 32-base buckets, as the
  // reference's dispatcher does for its one-thread-per-pair kernel
  std::vector<vd::PairRef> list;
  list.reserve(npairs);
  std::vector<std::pair<uint64_t, int64_t>> key;  // (bucket, position in list)
  key.reserve(npairs);
  for (int64_t k = 0; k < hb->nregions; ++k) {
    const int64_t r0 = hb->region_reads[k], r1 = hb->region_reads[k + 1];
    const int64_t h0 = hb->region_haps[k], h1 = hb->region_haps[k + 1];
    for (int64_t r = r0; r < r1; ++r)
      for (int64_t h = h0; h < h1; ++h) {
        const int64_t n = hb->read_off[r + 1] - hb->read_off[r], m = hb->hap_off[h + 1] - hb->hap_off[h];
        const uint64_t b = ((uint64_t)((n + 31) / 32) << 32) | (uint64_t)((m + 31) / 32);
        key.emplace_back(b, (int64_t)list.size());
        list.push_back(vd::PairRef{r, h, hb->out_off[k] + (r - r0) * (h1 - h0) + (h - h0)});
      }
  }
  std::stable_sort(key.begin(), key.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<vd::PairRef> sorted(list.size());
  for (size_t i = 0; i < key.size(); ++i) sorted[i] = list[key[i].second];
  struct Owned {  // the baseline's own buffers, freed on every return
    void* p = nullptr;
    ~Owned() {
      if (p) (void)hipFree(p);
    }
  } dlist, dws;
  HIPCHK(hipMalloc(&dlist.p, std::max<size_t>(sorted.size(), 1) * sizeof(vd::PairRef)));
  if (!sorted.empty())
    HIPCHK(hipMemcpyAsync(dlist.p, sorted.data(), sorted.size() * sizeof(vd::PairRef), hipMemcpyHostToDevice, st));
  const size_t budget = (size_t)budget_mb << 20;
  HIPCHK(hipMalloc(&dws.p, budget));
  float err[64];
  for (int q = 0; q < 64; ++q) err[q] = (float)params->err[q];
  const vd_batch dv = staged_view(hb);
  const vd::BatchDev b = batch_dev(&dv);
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipEventRecord(g.ev[0], st));
  int64_t nl = 0;
  for (size_t i = 0; i < key.size();) {
    size_t j = i;
    while (j < key.size() && key[j].first == key[i].first) ++j;
    // the bucket's widest haplotype sets the row stride of its launches
    int cols = 0;
    for (size_t t = i; t < j; ++t) {
      const vd::PairRef& pr = sorted[t];
      cols = std::max<int>(cols, (int)(hb->hap_off[pr.hap + 1] - hb->hap_off[pr.hap]));
    }
    const size_t per_pair = (size_t)3 * (cols + 1) * sizeof(float);
    if (per_pair > budget) return fail("vdhmm: one pair exceeds the baseline's workspace budget");
    const size_t step = std::max<size_t>(1, std::min<size_t>(budget / per_pair, (size_t)1 << 30));
    for (size_t s0 = i; s0 < j; s0 += step) {
      const int cnt = (int)std::min(step, j - s0);
      HIPCHK(vd::launch_ref_pairs(b, static_cast<vd::PairRef*>(dlist.p) + s0, cnt, cols, err, (float)params->gap,
                                  (float)params->ext, anchored, static_cast<float*>(dws.p), g.in_out.as<double>(), st));
      ++nl;
    }
    i = j;
  }
  HIPCHK(hipEventRecord(g.ev[1], st));
  HIPCHK(hipEventSynchronize(g.ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, g.ev[0], g.ev[1]));
  if (gpu_ms) *gpu_ms = ms;
  if (launches) *launches = nl;
  if (npairs) HIPCHK(hipMemcpy(out, g.in_out.p, npairs * 8, hipMemcpyDeviceToHost));
  return 0;
}

// ---- diploid genotype likelihood ------------------------------------------

struct GenoBufs {
  DevBuf mol, keep, units, map, ctr, mq_err, kidx, nkept, status;
  DevBuf in_off, in_lh, in_gl, in_mq, in_vals, in_cert;  // host entry points
  bool mq_err_ready = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
};
GenoBufs gg;

int geno_init(hipStream_t st) {
  if (!gg.ev[0])
    for (int i = 0; i < 3; ++i) HIPCHK(hipEventCreate(&gg.ev[i]));
  if (!gg.mq_err_ready || !gg.mq_err.current()) {
    if (g_capturing) return fail(CAPTURE_MISS);
    // molecule_rows.py:4: _MQ_ERROR[mq] = 10**(-min(mq,60)/10)
    static double t[256];
    for (int q = 0; q < 256; ++q) t[q] = std::pow(10., -std::min(q, 60) / 10.);
    HIPCHK(gg.mq_err.ensure(sizeof(t)));
    HIPCHK(hipMemcpyAsync(gg.mq_err.p, t, sizeof(t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    gg.mq_err_ready = true;
  }
  HIPCHK(gg.ctr.ensure(sizeof(unsigned long long) * vd::GCTR_COUNT));
  return 0;
}

int check_geno_batch(const vd_geno_batch* b) {
  if (!b) return fail("vdhmm: null genotype batch");
  if (b->nregions < 0 || b->nrows < 0 || b->mat_size < 0 || b->gl_size < 0)
    return fail("vdhmm: negative genotype batch size");
  if (b->nregions > 0 && (!b->row_off || !b->col_off || !b->mat_off || !b->gl_off))
    return fail("vdhmm: null genotype batch offsets");
  if (b->nregions > INT32_MAX) return fail("vdhmm: more than 2^31-1 genotype regions");
  return 0;
}

int check_prior(double prior, bool allow_none) {
  if (std::isnan(prior) || prior >= 1 || (!allow_none && prior < 0))
    return fail("vdhmm: invalid mixture prior (0 <= prior < 1, or negative for none)");
  return 0;
}

vd::GenoDev geno_dev(const vd_geno_batch* b) {
  vd::GenoDev d;
  d.row_off = b->row_off;
  d.col_off = b->col_off;
  d.mat_off = b->mat_off;
  d.gl_off = b->gl_off;
  d.nregions = b->nregions;
  d.nrows = b->nrows;
  d.mat_size = b->mat_size;
  d.gl_size = b->gl_size;
  d.unit_cap = b->gl_size / 64 + 2;  // ceil(sum of pairs / 64) when blocks do not overlap
  return d;
}

int geno_zero_ctr(hipStream_t st) {
  // a fill kernel rather than a memset node, so a captured graph replays it
  HIPCHK(vd::launch_zero_u64(gg.ctr.as<unsigned long long>(), vd::GCTR_COUNT, st));
  return 0;
}

int geno_stats(hipStream_t st, vd_geno_stats* s, bool molecule) {
  if (!s) return 0;
  HIPCHK(hipStreamSynchronize(st));
  unsigned long long c[vd::GCTR_COUNT];
  HIPCHK(hipMemcpy(c, gg.ctr.p, sizeof(c), hipMemcpyDeviceToHost));
  s->units = (int64_t)c[vd::GCTR_UNITS];
  s->invalid = (int64_t)c[vd::GCTR_INVALID];
  s->kept_rows = (int64_t)c[vd::GCTR_KEPT];
  s->skipped = (int64_t)c[vd::GCTR_LAYOUT];
  float ms = 0;
  if (molecule) {
    HIPCHK(hipEventElapsedTime(&ms, gg.ev[0], gg.ev[1]));
    s->molecule_ms = ms;
  }
  HIPCHK(hipEventElapsedTime(&ms, gg.ev[1], gg.ev[2]));
  s->mixture_ms = ms;
  return 0;
}

// the mixture's workspaces (sized by the batch; see vd::MixScratch)
int mix_scratch(const vd_geno_batch* b, const vd::GenoDev& d, vd::MixScratch* w) {
  const int64_t nreg = std::max<int64_t>(b->nregions, 1);
  HIPCHK(gg.units.ensure(sizeof(int64_t) * 2 * (b->nregions + 1)));
  HIPCHK(gg.map.ensure(sizeof(int32_t) * d.unit_cap));
  HIPCHK(gg.kidx.ensure(sizeof(int32_t) * std::max<int64_t>(b->nrows, 1)));
  HIPCHK(gg.nkept.ensure(sizeof(int32_t) * nreg));
  HIPCHK(gg.status.ensure(nreg));
  *w = vd::MixScratch{gg.units.as<int64_t>(), gg.map.as<int32_t>(), gg.kidx.as<int32_t>(), gg.nkept.as<int32_t>(),
                      gg.status.as<int8_t>()};
  return 0;
}

int mixture_device_locked(const vd_geno_batch* b, const double* lh, const int8_t* keep, double prior, int tri,
                          double* gl, hipStream_t st, vd_geno_stats* s) {
  if (int rc = check_geno_batch(b)) return rc;
  if (int rc = check_prior(prior, true)) return rc;
  if (b->nregions > 0 && ((!lh && b->mat_size > 0) || (!gl && b->gl_size > 0)))
    return fail("vdhmm: null genotype matrix or output");
  if (int rc = engine_init()) return rc;
  StreamOrder order(st);
  if (int rc = geno_init(st)) return rc;
  const vd::GenoDev d = geno_dev(b);
  vd::MixScratch w;
  if (int rc = mix_scratch(b, d, &w)) return rc;
  if (int rc = geno_zero_ctr(st)) return rc;
  if (s) HIPCHK(hipEventRecord(gg.ev[1], st));
  HIPCHK(vd::launch_mixture(d, lh, keep, prior, tri, w, gl, gg.ctr.as<unsigned long long>(), g.cus, st));
  if (s) HIPCHK(hipEventRecord(gg.ev[2], st));
  return geno_stats(st, s, false);
}

int genotype_device_locked(const vd_geno_batch* b, const double* ll, const uint8_t* mq, double min_cert,
                           double prior, double* gl, double* values, double* cert, hipStream_t st,
                           vd_geno_stats* s) {
  if (int rc = check_geno_batch(b)) return rc;
  if (int rc = check_prior(prior, true)) return rc;
  if (std::isnan(min_cert)) return fail("vdhmm: min_certainty is NaN");
  if (b->nregions > 0 && ((!ll && b->mat_size > 0) || (!mq && b->nrows > 0) || (!gl && b->gl_size > 0)))
    return fail("vdhmm: null genotype input or output");
  if (int rc = engine_init()) return rc;
  StreamOrder order(st);
  if (int rc = geno_init(st)) return rc;
  const vd::GenoDev d = geno_dev(b);
  HIPCHK(gg.mol.ensure(sizeof(double) * std::max<int64_t>(b->mat_size, 1)));
  HIPCHK(gg.keep.ensure(std::max<int64_t>(b->nrows, 1)));
  vd::MixScratch w;
  if (int rc = mix_scratch(b, d, &w)) return rc;
  if (int rc = geno_zero_ctr(st)) return rc;
  if (s) HIPCHK(hipEventRecord(gg.ev[0], st));
  HIPCHK(vd::launch_molecule(d, ll, mq, gg.mq_err.as<double>(), min_cert, values, cert, gg.mol.as<double>(),
                             gg.keep.as<int8_t>(), gg.ctr.as<unsigned long long>(), g.cus, st));
  if (s) HIPCHK(hipEventRecord(gg.ev[1], st));
  HIPCHK(vd::launch_mixture(d, gg.mol.as<double>(), gg.keep.as<int8_t>(), prior, 1, w, gl,
                            gg.ctr.as<unsigned long long>(), g.cus, st));
  if (s) HIPCHK(hipEventRecord(gg.ev[2], st));
  return geno_stats(st, s, true);
}

int guard_device_locked(const vd_geno_batch* b, const double* gl, const double* cert, const double* af,
                        double min_cert, double margin_scale, int8_t* flags, double* gaps, hipStream_t st,
                        int64_t* nflagged) {
  if (int rc = check_geno_batch(b)) return rc;
  if (std::isnan(min_cert) || !(margin_scale >= 0)) return fail("vdhmm: invalid guard margin or certainty bound");
  if (b->nregions > 0 && (!gl || !cert || !flags || !gaps)) return fail("vdhmm: null guard argument");
  if (int rc = engine_init()) return rc;
  StreamOrder order(st);
  if (int rc = geno_init(st)) return rc;
  if (int rc = geno_zero_ctr(st)) return rc;
  HIPCHK(vd::launch_guard(geno_dev(b), gl, cert, af, min_cert, margin_scale, flags, gaps,
                          gg.ctr.as<unsigned long long>(), st));
  if (nflagged) {
    HIPCHK(hipStreamSynchronize(st));
    unsigned long long c[vd::GCTR_COUNT];
    HIPCHK(hipMemcpy(c, gg.ctr.p, sizeof(c), hipMemcpyDeviceToHost));
    *nflagged = (int64_t)c[vd::GCTR_GUARD];
  }
  return 0;
}

// one region of rows x n staged from host memory
int stage_one_region(int rows, int n, int64_t pairs, vd_geno_batch* b, hipStream_t st) {
  const int64_t off[8] = {0, rows, 0, n, 0, 0, 0, 0};  // row_off[2], col_off[2], mat_off, gl_off
  HIPCHK(gg.in_off.ensure(sizeof(off)));
  HIPCHK(hipMemcpyAsync(gg.in_off.p, off, sizeof(off), hipMemcpyHostToDevice, st));
  const int64_t* o = gg.in_off.as<int64_t>();
  b->row_off = o;
  b->col_off = o + 2;
  b->mat_off = o + 4;
  b->gl_off = o + 5;
  b->nregions = 1;
  b->nrows = rows;
  b->mat_size = (int64_t)rows * n;
  b->gl_size = pairs;
  return 0;
}

int check_lh(int rows, int n, const double* lh) {  // mixture_cpu.py:14,25
  if (n < 1 || rows < 0) return fail("vdhmm: invalid mixture likelihood shape");
  if (rows > 0 && !lh) return fail("vdhmm: null mixture likelihood");
  const int64_t cnt = (int64_t)rows * n;
  for (int64_t i = 0; i < cnt; ++i)
    if (!std::isfinite(lh[i]) || lh[i] < 0) return fail("vdhmm: invalid mixture likelihood (non-finite or negative)");
  return 0;
}

int mixture_host_locked(int rows, int n, const double* lh, double prior, int tri, double* out) {
  if (int rc = check_lh(rows, n, lh)) return rc;
  if (int rc = check_prior(prior, true)) return rc;
  if (!out) return fail("vdhmm: null output");
  if (int rc = engine_init()) return rc;
  hipStream_t st = g.own_stream;
  StreamOrder order(st);
  if (int rc = geno_init(st)) return rc;
  const int64_t pairs = tri ? (int64_t)n * (n + 1) / 2 : (int64_t)n * n;
  const int64_t cnt = (int64_t)rows * n;
  vd_geno_batch b;
  if (int rc = stage_one_region(rows, n, pairs, &b, st)) return rc;
  HIPCHK(gg.in_lh.ensure(sizeof(double) * std::max<int64_t>(cnt, 1)));
  HIPCHK(gg.in_gl.ensure(sizeof(double) * pairs));
  if (cnt) HIPCHK(hipMemcpyAsync(gg.in_lh.p, lh, sizeof(double) * cnt, hipMemcpyHostToDevice, st));
  if (int rc = mixture_device_locked(&b, gg.in_lh.as<double>(), nullptr, prior, tri, gg.in_gl.as<double>(), st,
                                     nullptr))
    return rc;
  HIPCHK(hipMemcpyAsync(out, gg.in_gl.p, sizeof(double) * pairs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int normalized_host_locked(int rows, int n, const double* ll, const uint8_t* mq, double* values, double* cert) {
  if (rows < 0 || n < 0) return fail("vdhmm: negative shape");
  if (rows == 0) return 0;
  if (n < 1) return fail("vdhmm: normalized rows need at least one column");
  if (!ll || !mq || !values || !cert) return fail("vdhmm: null argument");
  if (int rc = engine_init()) return rc;
  hipStream_t st = g.own_stream;
  StreamOrder order(st);
  if (int rc = geno_init(st)) return rc;
  const int64_t cnt = (int64_t)rows * n;
  vd_geno_batch b;
  if (int rc = stage_one_region(rows, n, 0, &b, st)) return rc;
  HIPCHK(gg.in_lh.ensure(sizeof(double) * cnt));
  HIPCHK(gg.in_mq.ensure(rows));
  HIPCHK(gg.in_vals.ensure(sizeof(double) * cnt));
  HIPCHK(gg.in_cert.ensure(sizeof(double) * rows));
  HIPCHK(gg.mol.ensure(sizeof(double) * cnt));
  HIPCHK(gg.keep.ensure(rows));
  HIPCHK(hipMemcpyAsync(gg.in_lh.p, ll, sizeof(double) * cnt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(gg.in_mq.p, mq, rows, hipMemcpyHostToDevice, st));
  const vd::GenoDev d = geno_dev(&b);
  HIPCHK(vd::launch_molecule(d, gg.in_lh.as<double>(), gg.in_mq.as<uint8_t>(), gg.mq_err.as<double>(), -INFINITY,
                             gg.in_vals.as<double>(), gg.in_cert.as<double>(), gg.mol.as<double>(),
                             gg.keep.as<int8_t>(), nullptr, g.cus, st));
  HIPCHK(hipMemcpyAsync(values, gg.in_vals.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(cert, gg.in_cert.p, sizeof(double) * rows, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

// ---- helpers for the drop-ins ---------------------------------------------

struct HostBatch {
  std::vector<uint8_t> rb, qb, hb;
  std::vector<int64_t> ro{0}, ho{0}, rr{0}, rh{0}, oo{0};
  void add_read(const char* s, const unsigned char* q, int n) {
    if (n > 0) {
      rb.insert(rb.end(), (const uint8_t*)s, (const uint8_t*)s + n);
      qb.insert(qb.end(), q, q + n);
    }
    ro.push_back(ro.back() + std::max(n, 0));
  }
  void add_hap(const char* h, int m) {
    if (m > 0) hb.insert(hb.end(), (const uint8_t*)h, (const uint8_t*)h + m);
    ho.push_back(ho.back() + std::max(m, 0));
  }
  void close_region() {
    int64_t R = (int64_t)ro.size() - 1 - rr.back(), H = (int64_t)ho.size() - 1 - rh.back();
    rr.push_back((int64_t)ro.size() - 1);
    rh.push_back((int64_t)ho.size() - 1);
    oo.push_back(oo.back() + R * H);
  }
  vd_batch view() const {
    return vd_batch{rb.data(), qb.data(), ro.data(), hb.data(), ho.data(), rr.data(), rh.data(), oo.data(),
                    (int64_t)rr.size() - 1, (int64_t)ro.size() - 1, (int64_t)ho.size() - 1};
  }
};

int run_host(const HostBatch& hbat, double gap, double ext, int mode, int precision, double* out) {
  vd_params p;
  vd_params_default(&p, gap, ext);
  vd_batch v = hbat.view();
  return forward_host_locked(&v, &p, mode, precision, out, nullptr);
}

int env_precision() {
  const char* e = std::getenv("VARDEPTH_CUDA_PRECISION");
  if (e && std::strcmp(e, "float64") == 0) return VD_PREC_FP64;
  return VD_PREC_MIXED;
}

// trie_vector.cpp:198-205,271: evaluated-cell counter of the reference's
// prefix-sharing kernel (lane width 8/4/2 as dispatched on an AVX-512 host).
uint64_t trie_cells(const int* rn, int nr, const char* const* haps, const int* hn, int nh, int lanes) {
  int suffix = 0;
  if (nh > 1) {
    int minimum = *std::min_element(hn, hn + nh);
    while (suffix < minimum) {
      char c = haps[0][hn[0] - suffix - 1];
      bool same = true;
      for (int h = 1; h < nh; ++h)
        if (haps[h][hn[h] - suffix - 1] != c) {
          same = false;
          break;
        }
      if (!same) break;
      ++suffix;
    }
    if (suffix < 16) suffix = 0;
  }
  uint64_t columns = 0;
  for (int h = 0; h < nh; ++h) {
    int shared = 0;
    if (h) {
      int lim = std::min(hn[h], hn[h - 1]) - suffix;
      while (shared < lim && haps[h][shared] == haps[h - 1][shared]) ++shared;
    }
    int length = hn[h] - suffix;
    if (length > shared) columns += (uint64_t)(length - shared);
  }
  const int L = (lanes == 0 || lanes == 8) ? 8 : (lanes == 4 ? 4 : 2);
  uint64_t cells = 0;
  for (int first = 0; first < nr; first += L) {
    int take = std::min(L, nr - first), n = 0;
    for (int a = 0; a < take; ++a) n = std::max(n, rn[first + a]);
    cells += (uint64_t)n * (uint64_t)take * columns;
  }
  return cells;
}

// One read against `count` haplotypes in FP64 without the batch planner: the
// read, its qualities, the haplotypes, their offsets, the pair list and the
// read's exponent sum go up in one copy from a pinned buffer, one launch of
// k_scalar64 (a unit of 4 pairs per block) computes them, and the outputs come
// back in one copy.  The values equal the batched FP64 path's bit for bit.
int one_read_fast_locked(const char* read, const unsigned char* qual, int n, const char* const* haps,
                         const int* lengths, int count, double gap, double ext, int mode, double* out) {
  if (count <= 0) return 0;
  if (mode < 0 || mode > 2) return fail("vdhmm: mode must be 0, 1 or 2");
  if (n <= 0) {
    for (int k = 0; k < count; ++k) out[k] = -INFINITY;
    return 0;
  }
  vd_params p;
  vd_params_default(&p, gap, ext);
  if (int rc = check_params(p)) return rc;
  int max_m = 0;
  size_t hap_bytes = 0;
  for (int k = 0; k < count; ++k) {  // a negative length is an empty haplotype (-inf), as in the batched path
    max_m = std::max(max_m, lengths[k]);
    hap_bytes += (size_t)std::max(lengths[k], 0);
  }
  if (int rc = engine_init()) return rc;
  if (int rc = check_hap_len(max_m, 0)) return rc;
  hipStream_t st = g.own_stream;
  StreamOrder order(st);
  if (int rc = upload_params(p, st)) return rc;
  // payload: [read_off 2][hap_off count+1][pairs count][read_exp][read n][qual n][haps]
  auto up8 = [](size_t x) { return (x + 15) & ~size_t(15); };
  const size_t o_roff = 0, o_hoff = up8(o_roff + 2 * 8), o_pairs = up8(o_hoff + (size_t)(count + 1) * 8),
               o_exp = up8(o_pairs + (size_t)count * sizeof(vd::PairRef)), o_read = up8(o_exp + 4),
               o_qual = up8(o_read + n), o_hap = up8(o_qual + n), bytes = up8(o_hap + hap_bytes + 1),
               o_out = bytes, total = bytes + (size_t)count * 8;
  if (g.pinned_cap < total) {
    if (g.pinned) HIPCHK(hipHostFree(g.pinned));
    g.pinned = nullptr;
    g.pinned_cap = 0;
    const size_t want = std::max<size_t>(total + total / 2, 1 << 16);
    HIPCHK(hipHostMalloc(&g.pinned, want, hipHostMallocDefault));
    g.pinned_cap = want;
  }
  char* h = static_cast<char*>(g.pinned);
  int64_t* roff = reinterpret_cast<int64_t*>(h + o_roff);
  roff[0] = 0;
  roff[1] = n;
  int64_t* hoff = reinterpret_cast<int64_t*>(h + o_hoff);
  hoff[0] = 0;
  vd::PairRef* pairs = reinterpret_cast<vd::PairRef*>(h + o_pairs);
  char* hp = h + o_hap;
  for (int k = 0; k < count; ++k) {
    if (lengths[k] > 0) std::memcpy(hp + hoff[k], haps[k], (size_t)lengths[k]);
    hoff[k + 1] = hoff[k] + std::max(lengths[k], 0);
    pairs[k] = vd::PairRef{0, k, k};
  }
  int32_t e = 0;
  for (int i = 0; i < n; ++i) e += g.E_host[std::min<int>(qual[i], 63)];
  *reinterpret_cast<int32_t*>(h + o_exp) = e;
  std::memcpy(h + o_read, read, (size_t)n);
  std::memcpy(h + o_qual, qual, (size_t)n);
  HIPCHK(g.in_fast.ensure(total));
  char* d = g.in_fast.as<char>();
  HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
  const int smax = ((max_m + 48) + 7) & ~7;
  const int blocks = (count + vd::PAIRS_PER_UNIT - 1) / vd::PAIRS_PER_UNIT;
  HIPCHK(g.ws64.ensure((size_t)blocks * 4 * vd::PAIRS_PER_UNIT * smax * sizeof(double)));
  HIPCHK(g.ctr.ensure(vd::CTR_COUNT * 8));
  const vd::BatchDev b{reinterpret_cast<const uint8_t*>(d + o_read), reinterpret_cast<const uint8_t*>(d + o_qual),
                       reinterpret_cast<const int64_t*>(d + o_roff), reinterpret_cast<const uint8_t*>(d + o_hap),
                       reinterpret_cast<const int64_t*>(d + o_hoff), nullptr, nullptr, nullptr, 1, 1, count};
  vd::PlanDev pd{};
  pd.read_exp = reinterpret_cast<int32_t*>(d + o_exp);
  pd.ctr = g.ctr.as<unsigned long long>();
  double* dout = reinterpret_cast<double*>(d + o_out);
  HIPCHK(vd::launch_scalar64(b, pd, reinterpret_cast<const vd::PairRef*>(d + o_pairs), count,
                             g.tab64.as<vd::QTab<double>>(), mode == VD_MODE_ANCHORED ? 1 : 0, smax, g.ws64.as<double>(),
                             dout, st));
  HIPCHK(hipMemcpyAsync(h + o_out, dout, (size_t)count * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::memcpy(out, h + o_out, (size_t)count * 8);
  return 0;
}

double single(const char* read, const unsigned char* qual, int n, const char* hap, int m, double gap, double ext,
              int mode) {
  if (n <= 0 || m <= 0) return -INFINITY;
  double v = NAN;
  EagerLock lk;
  if (one_read_fast_locked(read, qual, n, &hap, &m, 1, gap, ext, mode, &v)) {
    std::fprintf(stderr, "%s\n", g_err.c_str());
    return NAN;
  }
  return v;
}

void one_read_batch(const char* read, const unsigned char* qual, int n, const char* const* haps, const int* lengths,
                    int count, double gap, double ext, double* out, int mode) {
  if (count <= 0) return;
  EagerLock lk;
  if (one_read_fast_locked(read, qual, n, haps, lengths, count, gap, ext, mode, out)) {
    std::fprintf(stderr, "%s\n", g_err.c_str());
    for (int k = 0; k < count; ++k) out[k] = NAN;
  }
}

}  // namespace

// ---- C ABI ------------------------------------------------------------------

extern "C" {

void vd_params_default(vd_params* p, double gap, double ext) {
  p->gap = gap;
  p->ext = ext;
  // batch_hmm.cpp:14: 10^(-min(q,60)/10) (bins 60..63 all equal the q>=60 value)
  for (int q = 0; q < 64; ++q) p->err[q] = std::pow(10., -std::min(q, 60) / 10.);
}

int vd_forward(const vd_batch* batch, const vd_params* params, int mode, int precision, double* out,
               vd_stats* stats) {
  if (!batch || !params || !out) return fail("vdhmm: null argument");
  EagerLock lk;
  return forward_host_locked(batch, params, mode, precision, out, stats);
}

int vd_forward_device(const vd_batch* dev_batch, const vd_params* params, int mode, int precision, int64_t npairs,
                      int max_hap_len, double* dev_out, void* stream, vd_stats* stats) {
  if (!dev_batch || !params || !dev_out) return fail("vdhmm: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  CaptureScope capture((hipStream_t)stream);
  if (int rc = engine_init()) return rc;
  auto t0 = std::chrono::steady_clock::now();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int rc = forward_device_locked(dev_batch, params, mode, precision, npairs, max_hap_len, dev_out,
                                 (hipStream_t)stream, stats);
  if (stats && !rc)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int vd_forward_device_noalloc(const vd_batch* dev_batch, const vd_params* params, int mode, int precision,
                              int64_t npairs, int max_hap_len, double* dev_out, void* stream) {
  if (!dev_batch || !params || !dev_out) return fail("vdhmm: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  return forward_noalloc_locked(dev_batch, params, mode, precision, npairs, max_hap_len, dev_out,
                                (hipStream_t)stream);
}

int vd_grad(const vd_batch* batch, const vd_params* params, int mode, const double* weights, double* out,
            double* grad, vd_stats* stats) {
  if (!batch || !params || !grad) return fail("vdhmm: null argument");
  EagerLock lk;
  return grad_host_locked(batch, params, mode, weights, out, grad, stats);
}

int vd_grad_device(const vd_batch* dev_batch, const vd_params* params, int mode, int64_t npairs, int max_hap_len,
                   int max_read_len, const double* dev_weights, double* dev_out, double* dev_grad, void* stream,
                   vd_stats* stats) {
  if (!dev_batch || !params || !dev_out || !dev_grad) return fail("vdhmm: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  CaptureScope capture((hipStream_t)stream);
  if (int rc = engine_init()) return rc;
  auto t0 = std::chrono::steady_clock::now();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int rc = grad_device_locked(dev_batch, params, mode, npairs, max_hap_len, max_read_len, dev_weights, dev_out,
                              dev_grad, (hipStream_t)stream, stats);
  if (stats && !rc)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int vd_forward_plain(const vd_batch* batch, const vd_params* params, int mode, int64_t max_pairs, double* out,
                     double* gpu_ms) {
  if (!batch || !params || !out) return fail("vdhmm: null argument");
  EagerLock lk;
  return plain_host_locked(batch, params, mode, max_pairs, out, gpu_ms);
}

int vd_forward_refdesign(const vd_batch* batch, const vd_params* params, int mode, int64_t budget_mb, double* out,
                         double* gpu_ms, int64_t* launches) {
  if (!batch || !params || !out) return fail("vdhmm: null argument");
  EagerLock lk;
  return refdesign_host_locked(batch, params, mode, budget_mb, out, gpu_ms, launches);
}

int vd_mixture_general(int rows, int n, const double* lh, double prior, int triangular, double* out) {
  EagerLock lk;
  return mixture_host_locked(rows, n, lh, prior, triangular ? 1 : 0, out);
}

int vd_mixture_two(int rows, const double* lh, double prior, double* out) {
  if (int rc = check_prior(prior, false)) return rc;  // mixture_cpu.py:14: 0 <= prior < 1
  EagerLock lk;
  return mixture_host_locked(rows, 2, lh, prior, 0, out);
}

int vd_normalized_rows(int rows, int n, const double* ll, const unsigned char* mq, double* values,
                       double* certainty) {
  EagerLock lk;
  return normalized_host_locked(rows, n, ll, mq, values, certainty);
}

int vd_mixture_regions_device(const vd_geno_batch* batch, const double* lh, const signed char* keep, double prior,
                              int triangular, double* gl, void* stream, vd_geno_stats* stats) {
  std::lock_guard<std::mutex> lk(g_mu);
  CaptureScope capture((hipStream_t)stream);
  auto t0 = std::chrono::steady_clock::now();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int rc = mixture_device_locked(batch, lh, (const int8_t*)keep, prior, triangular ? 1 : 0, gl, (hipStream_t)stream,
                                 stats);
  if (stats && !rc)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int vd_genotype_device(const vd_geno_batch* batch, const double* ll, const unsigned char* mq, double min_certainty,
                       double prior, double* gl, double* values, double* certainty, void* stream,
                       vd_geno_stats* stats) {
  std::lock_guard<std::mutex> lk(g_mu);
  CaptureScope capture((hipStream_t)stream);
  auto t0 = std::chrono::steady_clock::now();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int rc = genotype_device_locked(batch, ll, mq, min_certainty, prior, gl, values, certainty, (hipStream_t)stream,
                                  stats);
  if (stats && !rc)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int vd_genotype_guard(const vd_geno_batch* batch, const double* gl, const double* certainty,
                      const double* allele_freq, double min_certainty, double margin_scale, signed char* flags,
                      double* gaps, void* stream, int64_t* nflagged) {
  std::lock_guard<std::mutex> lk(g_mu);
  CaptureScope capture((hipStream_t)stream);
  return guard_device_locked(batch, gl, certainty, allele_freq, min_certainty, margin_scale, (int8_t*)flags, gaps,
                             (hipStream_t)stream, nflagged);
}

int vd_max_hap_len(int gradients) {
  std::lock_guard<std::mutex> lk(g_mu);
  return max_hap_len_for(gradients ? 1 : 0);
}

int vd_tuning_get(vd_tuning* t) {
  if (!t) return fail("vdhmm: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  tune_defaults();
  *t = g.tune;
  t->class_overlap = g.overlap ? 1 : 0;
  return 0;
}

int vd_tuning_set(const vd_tuning* t) {
  if (!t) return fail("vdhmm: null argument");
  for (int r = 1; r < 17; ++r) {
    if (t->fwd_unroll[r] != 2 && t->fwd_unroll[r] != 4) return fail("vdhmm: fwd_unroll must be 2 or 4");
    if (t->grad_variant[r] != 0 && t->grad_variant[r] != 1) return fail("vdhmm: grad_variant must be 0 or 1");
    if (t->fwd_blocks_per_cu[r] < 0 || t->grad_blocks_per_cu[r] < 0) return fail("vdhmm: negative grid size");
    if (t->fwd_share[r] < 0 || t->fwd_share[r] > 3) return fail("vdhmm: fwd_share must be 0, 1, 2 or 3");
    if (t->fwd_route64[r] != 0 && t->fwd_route64[r] != 1) return fail("vdhmm: fwd_route64 must be 0 or 1");
  }
  if (t->share_chunk < 1 || t->share_chunk > 64) return fail("vdhmm: share_chunk must be in 1..64");
  if (t->class_overlap != 0 && t->class_overlap != 1) return fail("vdhmm: class_overlap must be 0 or 1");
  if (t->fwd_pack != 0 && t->fwd_pack != 1) return fail("vdhmm: fwd_pack must be 0 or 1");
  std::lock_guard<std::mutex> lk(g_mu);
  g.tune = *t;
  g.tune_init = true;
  g.overlap = t->class_overlap != 0;
  return 0;
}

int vd_prefix_sharing(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  const int was = g.share ? 1 : 0;
  if (on >= 0) g.share = on != 0;
  return was;
}

int vd_suffix_sharing(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  const int was = g.suffix ? 1 : 0;
  if (on >= 0) g.suffix = on != 0;
  return was;
}

int vd_class_overlap(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  const int was = g.overlap ? 1 : 0;
  if (on >= 0) g.overlap = on != 0;
  return was;
}

void vd_class_timing(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g.class_timing = on != 0;
}

int vd_class_ms(double* fwd_ms, double* grad_ms) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int r = 0; r < 17; ++r) {
    if (fwd_ms) fwd_ms[r] = g.fwd_cls_ms[r];
    if (grad_ms) grad_ms[r] = g.grad_cls_ms[r];
  }
  return 0;
}

const char* vd_last_error(void) { return g_err.c_str(); }

int vd_device(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return engine_init() ? -1 : g.dev;
}

int pairhmm_pairs(const char* const* reads, const unsigned char* const* quals, const int* rn,
                  const char* const* haps, const int* hn, int count, int mode, int lanes, double* output) {
  (void)lanes;
  if (count < 0 || mode < 0 || mode > 2) return 1;  // batch_hmm.cpp:194
  if (count == 0) return 0;
  HostBatch hb;
  for (int p = 0; p < count; ++p) {
    hb.add_read(reads[p], quals[p], rn[p]);
    hb.add_hap(haps[p], hn[p]);
    hb.close_region();
  }
  EagerLock lk;
  return run_host(hb, 0.001, 0.1, mode, VD_PREC_FP64, output) ? 2 : 0;
}

int trie_many(const char* const* reads, const unsigned char* const* quals, const int* rn, int nr,
              const char* const* haps, const int* hn, int nh, int mode, int lanes, double* out, uint64_t* cells) {
  if (nr < 0 || nh < 0 || mode < 0 || mode > 2) return 1;  // trie_vector.cpp:286
  *cells = trie_cells(rn, nr, haps, hn, nh, lanes);
  if (nr == 0 || nh == 0) return 0;
  HostBatch hb;
  for (int a = 0; a < nr; ++a) hb.add_read(reads[a], quals[a], rn[a]);
  for (int h = 0; h < nh; ++h) hb.add_hap(haps[h], hn[h]);
  hb.close_region();
  {
    EagerLock lk;
    if (run_host(hb, 0.001, 0.1, mode, env_precision(), out)) return 2;
  }
  // trie_vector.cpp:277: -inf for empty inputs, NaN when the likelihood is <= 1e-200
  const double lim = std::log(1e-200);
  for (int64_t i = 0; i < (int64_t)nr * nh; ++i)
    if (!std::isinf(out[i]) && !(out[i] > lim)) out[i] = NAN;
  return 0;
}

double read_haplotype(const char* read, const unsigned char* qual, int n, const char* hap, int m, double gap,
                      double ext) {
  return single(read, qual, n, hap, m, gap, ext, VD_MODE_SEMIGLOBAL_RESCALE);
}

void read_haplotype_batch(const char* read, const unsigned char* qual, int n, const char* const* hap,
                          const int* lengths, int count, double gap, double ext, double* out) {
  one_read_batch(read, qual, n, hap, lengths, count, gap, ext, out, VD_MODE_SEMIGLOBAL_RESCALE);
}

void read_haplotype_batch_mode(const char* read, const unsigned char* qual, int n, const char* const* haps,
                               const int* lengths, int count, double gap, double ext, double* out, int mode) {
  (void)mode;
  one_read_batch(read, qual, n, haps, lengths, count, gap, ext, out, VD_MODE_SEMIGLOBAL_RESCALE);
}

int wide_capability(void) { return 8; }

double anchored_pairhmm(const char* read, const unsigned char* qual, int n, const char* hap, int m, double gap,
                        double extension) {
  return single(read, qual, n, hap, m, gap, extension, VD_MODE_ANCHORED);
}

double vd_core_read_haplotype(const char* read, const unsigned char* qual, int n, const char* hap, int m, double gap,
                              double ext) {
  return single(read, qual, n, hap, m, gap, ext, VD_MODE_SEMIGLOBAL);
}

void vd_core_read_haplotype_batch(const char* read, const unsigned char* qual, int n, const char* const* hap,
                                  const int* lengths, int count, double gap, double ext, double* out) {
  one_read_batch(read, qual, n, hap, lengths, count, gap, ext, out, VD_MODE_SEMIGLOBAL);
}

}  // extern "C"

// Internal types shared by the HIP kernels (vdhmm_kernels.hip) and the host
// engine (vdhmm_engine.cpp).  Not part of the public C-ABI (include/vdhmm.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vd {

// Read classes.  Class c in [1,16]: the FP32 anti-diagonal kernel with c rows
// per lane (a 16-lane group covers 16*c rows).  CLS_EMPTY: read of length 0
// (every pair is -inf).  CLS_F64: rows beyond 256, or FP64 requested -> the
// FP64 row-tiled kernel.
constexpr int CLS_EMPTY = 0;
constexpr int CLS_F64 = 17;
constexpr int NCLS = 18;
constexpr int MAX_R32 = 16;

constexpr int GROUP = 16;        // lanes per pair (one DPP row)
constexpr int PAIRS_PER_UNIT = 4;
// unit_cnt: the nb + 1 bucket counts (scanned in place), then the unit scan's
// per-tile sums, one per SCAN_TILE buckets
constexpr int SCAN_TILE = 4096;
__host__ __device__ constexpr int64_t unit_cnt_len(int64_t nb) { return nb + 1 + (nb + SCAN_TILE - 1) / SCAN_TILE; }
// bytes of a unit's hap-code slots in LDS (uint16 per slot, 16-byte multiple)
__host__ __device__ constexpr size_t lds_codes_bytes(int smax) {
  return ((size_t)PAIRS_PER_UNIT * smax * 2 + 15) & ~size_t(15);
}
constexpr int R64 = 4;           // FP64 rows per lane: 16 lanes x 4 = 64-row tiles

// hap code stream sentinels (hap bytes are 0..255)
constexpr uint32_t SENT_POST = 0x100u;  // column beyond the haplotype end
constexpr uint32_t SENT_PRE = 0x101u;   // column <= 0
constexpr uint32_t NOMATCH = 0xFFFFu;   // read code of dummy rows

// FP32 start scale 2^108: every scaled value stays below 2^108 * 1.2 * (m+1),
// so haplotypes up to 2^18 columns cannot overflow.  A pair whose scaled
// terminal is below 2^-100 leaves the FP32 window (its dominant paths would
// pass within 2^26 of the denormal range) and goes to the FP64 rescue.
constexpr int CEXP32 = 108;
constexpr int CEXP64 = 500;   // FP64 start / tile-renormalised scale 2^500
constexpr float FLAG32_MIN = 0x1p-100f;
constexpr int CEXPG64 = 900;  // FP64 gradient start scale 2^900 (single tile of <= 256 rows)
constexpr int NGRAD = 66;     // gradient layout: [0] d/dgap, [1] d/dext, [2 + q] d/derr[q]

// Per-quality-bin row constants of the rescaled recurrence, for one
// precision.  With a = 1-2g, b = 1-x, c = g/4, d = x/4 and a per-bin
// normaliser rho_q:
//   good = (1-e) a / rho, bad = (e/3) a / rho     (match emission, scaled)
//   mu   = d / rho                                 (I recurrence of the next row)
//   beta = (b c / a) / rho                         (I term of the next row's M)
//   cfin = c / rho                                 (terminal weight of I)
//   gamma = b g / a                                (D term of the next row's M)
// FP64: rho_q = 2^E_q >= max(1-e, e/3) a, a power of two, so that the FP64
// results are exact rescalings (the per-read sum of E_q is an integer).
// FP32: rho_q = max((1-e) a, (e/3) a, d, c, b c / a) exactly, so that the
// largest per-row factor is 1 and a matching row costs no FP32 range (a
// power of two would cost up to one bit per row); the per-read sum of ln rho
// is taken in FP64.  Either way every row factor is at most 1, and the
// constants are rounded once, when the table is built.
template <typename T>
struct QTab {
  T good[64], bad[64], mu[64], beta[64], cfin[64];
  T gamma, gap, ext;
  double lnrho[64];  // ln rho_q: read by the planner only, so kernels stage the fields before it
};
// Bytes of a table that kernels stage in LDS (every field before lnrho).
template <typename T>
constexpr size_t qtab_lds_bytes() {
  return offsetof(QTab<T>, lnrho);
}

// Partial derivatives of the scaled row constants with respect to g, x and
// e_q with rho_q held fixed (the likelihood is the same for any fixed choice
// of normalisers, which the log correction cancels, so the derivative at
// fixed rho is the derivative), used to turn the adjoint sums into gradient
// terms:
//   d good/dg = -2(1-e)/rho  d bad/dg = -2(e/3)/rho  d good/de = -a/rho  d bad/de = a/(3 rho)
//   d mu/dx = 1/(4 rho)      d beta/dx = -g/(4 a rho) d beta/dg = (1-x)/(4 rho a^2)
//   d cfin/dg = 1/(4 rho)    d gamma/dg = (1-x)/a^2   d gamma/dx = -g/a
struct GradTab {
  double dgood_dg[64], dbad_dg[64], dgood_de[64], dbad_de[64];
  double dgood_de32[64];  // d good/de with the FP32 normaliser (the others hold the FP64 one)
  double dmu_dx[64], dbeta_dx[64], dbeta_dg[64], dcfin_dg[64];
  double dgam_dg, dgam_dx;
};

// One read x haplotype pair addressed by global read / hap index.
struct PairRef {
  int64_t read, hap, out;
};

// Device view of a region batch (CSR, see vardepth_amd/workloads.py).
struct BatchDev {
  const uint8_t* read_bytes;
  const uint8_t* qual_bytes;
  const int64_t* read_off;
  const uint8_t* hap_bytes;
  const int64_t* hap_off;
  const int64_t* region_reads;
  const int64_t* region_haps;
  const int64_t* out_off;
  int64_t nregions, nreads, nhaps;
};

// Device scratch of one planned batch.
struct PlanDev {
  int8_t* read_cls;         // [nreads]
  int32_t* read_exp;        // [nreads]  sum of E_q over the read (FP64 normalisers)
  double* read_ln32;        // [nreads]  sum of ln rho_q over the read (FP32 normalisers)
  int32_t* class_off;       // [nregions * (NCLS+1)] per-region class offsets
  int64_t* read_perm;       // [nreads]  reads grouped by class inside each region
  int64_t* unit_cnt;        // [unit_cnt_len(nregions * NCLS)] FP32 units per bucket, scanned in place
  int32_t* unit_bucket;     // [total units] bucket of every unit
  PairRef* f64_list;        // [capacity] FP64 work list
  unsigned long long* ctr;  // counters, see CTR_*
  int64_t f64_cap;
  int class_major;          // bucket order: 0 region-major (forward), 1 class-major (gradients)
  // haplotype prefix sharing of the FP32 forward (trie = 1): a unit is 4 reads
  // of one (region, class), each group sweeping the region's haplotypes in
  // lexicographic order and resuming each from the fork column it shares with
  // its predecessor
  int trie;                 // bit c set: class c runs read units with prefix sharing
  int trie_chunk;           // haplotypes per read unit (1..64)
  int32_t* trie_order;      // [nhaps] region slot h0 + rank -> global haplotype index
  int32_t* trie_start;      // [nhaps] per rank: resume column s (0 = full sweep)
  int32_t* trie_fork;       // [nhaps] per region: sorted distinct fork columns at h0 + i, i < trie_nfork[k]
  int32_t* trie_nfork;      // [nregions]
  float* trie_ws;           // fork wavefronts: [block][group][TRIE_SLOTS][16 lanes][3R+3 padded to 4]
  // suffix continuation (sfx = 1): a region whose haplotypes all end in the
  // same S >= 16 bases gets, per read, one adjoint sweep over those S columns;
  // each haplotype is then swept to column m - S only and its terminal
  // completed by a dot product with the adjoint
  int sfx;                  // bit c set: class c runs with the suffix continuation (planner finds S if any)
  int split;                // bit c set (with sfx): the suffix kernel takes only the read units of regions
                            // with a shared suffix, the prefix-only kernel the others
  int32_t* trie_sfx;        // [nregions] common suffix S (0: none)
  float* sfx_ws;            // [block][64 lanes][84]: adjoint weights, boundary mass, the column m - S state
  // pair units packed across regions (class-major plans of the gradients): the
  // per-bucket counts are pairs, not units, and unit_bucket maps every pair of
  // a class to its bucket, so a unit of 4 consecutive pairs of a class may
  // span regions and small regions leave no lanes idle
  int pack_pairs;
};

constexpr int TRIE_SLOTS = 32;  // distinct fork columns kept per region (later forks are swept in full)
constexpr int SFX_MIN = 16;     // shortest common suffix worth the adjoint sweep
constexpr int SFX_SEED_EXP = 100;  // adjoint seed 2^100 (FP32 headroom on both sides)
constexpr int SFX_WS = (5 * MAX_R32 + 4) * 64;  // floats of sfx_ws per block (64 lanes x 84)

enum {
  CTR_UNIT32 = 0,   // next FP32 unit
  CTR_F64 = 1,      // FP64 list length
  CTR_UNIT64 = 2,   // next FP64 unit
  CTR_FLAGGED = 3,  // FP32 pairs flagged for rescue
  CTR_OVERFLOW = 4, // FP64 list overflow (pairs dropped: reported as error)
  CTR_NUNITS = 5,   // total FP32 units (written by the scan)
  CTR_F64BAD = 6,   // FP64 results that were not finite and positive
  CTR_SWEPT = 7,    // FP32 cells swept (read rows x haplotype columns computed)
  CTR_CLS = 8,      // [8 + c]: next unit of class c (per-class gradient kernels)
  CTR_BIN64 = 32,   // [32 + b]: FP64 gradient sub-list lengths (8 row bins)
  CTR_CLS2 = 40,    // [40 + c]: next unit of class c for the second kernel of a split class
  CTR_BIN64T = 57,  // FP64 gradient list entries of reads beyond 16 rows per lane (the tiled adjoint)
  CTR_COUNT = 58
};

// Device view of a genotype batch: region k owns rows row_off[k]..row_off[k+1]
// (global row index: mapping quality and keep flag), haplotypes
// col_off[k]..col_off[k+1], the row-major rows x haplotypes block at
// mat_off[k] of the likelihood / molecule matrices, and its pair outputs at
// gl_off[k].
struct GenoDev {
  const int64_t* row_off;
  const int64_t* col_off;
  const int64_t* mat_off;
  const int64_t* gl_off;
  int64_t nregions, nrows;
  int64_t mat_size, gl_size;  // entries of the matrices / of the pair outputs (bounds of every block)
  int64_t unit_cap;           // capacity of each wave-unit map (units of 64 items)
};

enum {
  GCTR_UNITS = 0,    // mixture wave units (written by the scan)
  GCTR_INVALID = 1,  // regions whose matrix held a non-finite or negative entry
  GCTR_KEPT = 2,     // rows that passed the certainty filter
  GCTR_LAYOUT = 3,   // regions skipped: block outside the matrix / output bounds, or unit map full
  GCTR_GUARD = 4,    // regions flagged by the near-tie guard
  GCTR_UNITS_WIDE = 5,  // k_mix_wide wave units (written by the scan)
  GCTR_COUNT = 8
};

struct LaunchCfg {
  int blocks32;   // persistent FP32 grid (one wave per block)
  int blocks64;   // persistent FP64 grid
  int smax;       // LDS hap-code slots per group (>= max hap length + 40)
};

}  // namespace vd

namespace vd {
size_t lds_fwd32(int smax);
size_t lds_fwd64(int smax);
hipError_t occupancy(int anchored, int r, int u, int smax, int* blocks_per_cu);  // r = 0: FP64 kernel
hipError_t occupancy_trie(int anchored, int r, int sfx, int smax, int* blocks_per_cu);    // prefix-sharing FP32 kernel
void launch_plan(const BatchDev& b, const PlanDev& p, const int32_t* qE, const QTab<float>* tab32, int anchored,
                 int route64, double* out, hipStream_t st);  // p.class_major selects the bucket order; route64
                                                              // as k_prep_reads (bit 0: every read to FP64)
// A second stream for the FP32 class kernels: consecutive classes alternate
// between the caller's stream and `side`, so that the last units of one
// persistent kernel overlap the first of the next.  Kernels on `side` use the
// second half of the prefix-sharing workspaces (offsets in floats).
struct SideLane {
  hipStream_t side;
  hipEvent_t fork, join;
  size_t trie_half, sfx_half;
};
hipError_t launch_pipeline(const BatchDev& b, const PlanDev& p, const int32_t* qE, const QTab<float>* tab32,
                           const QTab<double>* tab64, int anchored, int route64, int smax, const int* blocks32,
                           const int* blocks32s,
                           const int* unroll32, int blocks64, double* ws64, PairRef* f64_sorted,
                           unsigned long long* f64_bins, double* out, hipStream_t st, hipEvent_t* ev,
                           hipEvent_t* ev_cls, const SideLane* lane2);
// FP64 list binning (f64_sorted / f64_bins non-null): capacity of the sorted
// copy = the list's, bins = 2 * F64_BINS_TOTAL counters
constexpr int F64_BINS_TOTAL = 16 * 32;
// One read against `count` haplotypes in FP64, one unit of 4 pairs per block
// (the scalar drop-ins): pairs, ws (4 * 4 * smax doubles per block) and out
// on the device.
hipError_t launch_scalar64(const BatchDev& b, const PlanDev& p, const PairRef* pairs, int count,
                           const QTab<double>* tab64, int anchored, int smax, double* ws, double* out,
                           hipStream_t st);
hipError_t launch_ref_pairs(const BatchDev& b, const PairRef* pairs, int count, int cols, const float* err,
                            float gap, float ext, int anchored, float* ws, double* out, hipStream_t st);
hipError_t launch_plain_pair(const uint8_t* read, const uint8_t* qual, int n, const uint8_t* hap, int m,
                             const QTab<float>* tab, int anchored, double lnk, double* out, hipStream_t st);
size_t lds_grad32(int smax, int r = 0);  // r rows per lane; 0: the largest class
size_t lds_grad64(int smax);
size_t grad_ck_stride(int smax, int fp64);
hipError_t grad_occupancy(int anchored, int r, int fp64, int variant, int smax, int* blocks_per_cu);
// The tiled FP64 adjoint of reads beyond 256 rows: tiles of 64 rows, at most
// ntmax per read (0: no such read, the kernel is not launched); a workspace
// of ws_stride doubles per block (grad_tile_ws), blocks <= the partial
// gradient rows.
struct GradTiled {
  int ntmax, blocks;
  double* ws;
  size_t ws_stride;
};
size_t grad_tile_ws(int smax, int ntmax);
hipError_t grad_tiled_occupancy(int anchored, int smax, int* blocks_per_cu);
hipError_t launch_grad(const BatchDev& b, const PlanDev& p, const int32_t* qE, const QTab<float>* tab32,
                       const QTab<double>* tab64, const GradTab* gt, int anchored, int smax,
                       const int* blocks32, const int* variant32, const int* blocks64, int max_blocks, float* ck32,
                       double* ck64, PairRef* bins, int64_t bin_cap, const double* weights, double* out,
                       double* gpart, double* grad, hipStream_t st, hipEvent_t* ev, hipEvent_t* ev_cls,
                       const GradTiled* tiled);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
// scratch of the no-allocation forward (vdhmm_noalloc.hip): device address of
// the library's static array on device `dev`, and its size
hipError_t noalloc_scratch(int dev, void** ptr, size_t* bytes);
hipError_t launch_put_tables(const QTab<float>& t32, const QTab<double>& t64, const int32_t* E, QTab<float>* d32,
                             QTab<double>* d64, int32_t* dE, hipStream_t st);
hipError_t launch_guard(const GenoDev& g, const double* gl, const double* cert, const double* af, double min_cert,
                        double margin_scale, int8_t* flag, double* gaps, unsigned long long* ctr, hipStream_t st);
hipError_t launch_molecule(const GenoDev& g, const double* ll, const uint8_t* mq, const double* mq_err,
                           double min_cert, double* values, double* cert, double* mol, int8_t* keep,
                           unsigned long long* ctr, int cus, hipStream_t st);
// Workspaces of the mixture: cnt 2 x (nregions + 1) int64; unit_first
// unit_cap int32; kidx nrows int32; nkept nregions int32; status nregions int8.
struct MixScratch {
  int64_t* cnt;
  int32_t* unit_first;
  int32_t* kidx;
  int32_t* nkept;
  int8_t* status;
};
hipError_t launch_mixture(const GenoDev& g, const double* mat, const int8_t* keep, double prior, int triangular,
                          const MixScratch& w, double* gl, unsigned long long* ctr, int cus, hipStream_t st);
}  // namespace vd

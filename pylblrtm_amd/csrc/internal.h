// Internal interfaces shared by the HIP translation units of libpylblrtm_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include "../../include/pylblrtm_amd.h"

namespace lbl {

// Physical constants of the reference, used verbatim
// (pylblrtm/forward.py:32-36).
constexpr double kTRef = 296.0;
constexpr double kC2 = 1.4387752;
constexpr double kDoppler = 3.5810e-7;
constexpr double kSqrtLn2 = 0x1.aa4499161cd47p-1;     // math.sqrt(math.log(2.0))
constexpr double kVoigtK = 0x1.e0fdec495104dp-2;      // kSqrtLn2 / math.sqrt(math.pi)

constexpr int kMaxGroups = 1024;   // groups whose q-ratios fit the LDS table
constexpr int kMaxMu = 8;          // viewing cosines per call
constexpr int kBlock = 256;        // threads per workgroup of the line sum

// Error word, written with atomicMin so the host reads the first failure:
//   kErrBadGroup / kErrBadGas / kErrBadTips: an index or table of the inputs
//     that the kernels cannot use (reported before any numerical failure);
//   ((2*layer + kind + 1) << 11) | group, kind 0 = strength, 1 = od: the
//     reference's first numerical failure in its loop order (layer, then
//     strength before od, then group).
constexpr uint32_t kErrNone = 0xFFFFFFFFu;
constexpr int kErrGroupBits = 11;
constexpr uint32_t kErrBadGroup = 1;   // a line's group outside [0, n_groups)
constexpr uint32_t kErrBadGas = 2;     // a group's gas row outside [0, n_gases)
constexpr uint32_t kErrBadTips = 3;    // a group's TIPS table empty (offsets not increasing)
__host__ __device__ __forceinline__ uint32_t err_numeric(int layer, int kind, int group) {
  return ((uint32_t)(2 * layer + kind + 1) << kErrGroupBits) | (uint32_t)group;
}

// Hardware assigns workgroup b to XCD b % 8.  Give every XCD a contiguous
// run of logical workgroups so neighbouring tiles (which share most of their
// line records) share an L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n) {
  const int64_t per = n >> 3, rem = n & 7;
  const int64_t xcd = b & 7, idx = b >> 3;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Grid point i exactly as numpy builds it: nu_min + (double(i) * step),
// rounded after the multiply and after the add (contraction off: the
// __d*_rn helpers are plain operators, which hipcc would otherwise fuse).
__device__ __forceinline__ double grid_nu(double nu_min, double step, int64_t i) {
#pragma clang fp contract(off)
  return __dadd_rn(nu_min, __dmul_rn((double)i, step));
}

// Sets the calling thread's lbl_last_error() text (printf format) and
// returns code.
int set_error(int code, const char* fmt, ...);

// Grid points per thread of the mixed-precision line sums.
constexpr int kPointsMixed = 8;

// Far-field series of the line sums.  A wavefront's 64 lanes own
// kWavePoints consecutive grid points, whose middle is nu_w (wave_middle) and
// half-span H = kWavePoints/2 * step.  A record whose centre lies at least
// kSeriesRatio * H from nu_w contributes a smooth function of
// tau = (nu - nu_w) / H, |tau| < 1, to all of them: its one-node form's Taylor
// series in tau, kSeriesTerms terms, is summed per lane into kSeriesTerms
// coefficients, reduced across the wavefront and evaluated at each point.
constexpr int kWavePoints = 64 * kPointsMixed;
// A second level for the farther records: the series about a slice's middle
// (kSlicePoints points, half-span Hs = kSlicePoints/2 * step = 4 H), for
// records at least kSeriesRatio * Hs away, summed once per slice.
constexpr int kSlicePoints = 4 * kWavePoints;
constexpr double kSeriesRatio = 3.0;
constexpr int kSeriesTerms = 18;     // mixed: one-node form, FP32
constexpr int kSeriesTerms64 = 28;   // FP64: second-order shifted form

// A third level for the records inside a wavefront's near range (mixed modes):
// the series about the middle of each group of kGroupPoints points (8 lanes),
// half-span Hg = kGroupPoints/2 * step = H/8, for records at least
// kSeriesRatio * Hg from it; the near field evaluates point by point only the
// records within that distance of some group of the wavefront.
constexpr int kGroupPoints = 8 * kPointsMixed;
constexpr int kGroupsPerWave = kWavePoints / kGroupPoints;

// Middle of the group whose first grid point is i: grid(i) + (kGroupPoints-1)/2 step.
__device__ __forceinline__ double group_middle(double nu_min, double step, int64_t i) {
#pragma clang fp contract(off)
  return __dadd_rn(__dadd_rn(nu_min, __dmul_rn((double)i, step)),
                   0.5 * (double)(kGroupPoints - 1) * step);
}

// The group level's test, shared bit for bit by k_far_series_mixed (which
// sums a pair where it holds) and the near field's staging (which evaluates
// the pair point by point where it fails), so each (group, record) pair is
// taken exactly once: the one-node form holds at every point of group g of a
// wavefront iff |nu_g - c| >= ds1 - (H - Hg).  nu_g - c is formed as
// d0 + g * dg, with d0 = (float)(middle of group 0 - c) and dg the groups'
// spacing in FP32; contraction is off in these helpers, so that no step is
// fused into an fma in one kernel and not in the other.
struct GroupTest {
  double nu0;   // middle of the wavefront's group 0
  float dg;     // kGroupPoints * step
  float dH;     // H - Hg = (kWavePoints - kGroupPoints)/2 * step
  float hg;     // Hg = kGroupPoints/2 * step
};
__device__ __forceinline__ GroupTest group_test(double nu_min, double step, int64_t w_first) {
#pragma clang fp contract(off)
  GroupTest t;
  t.nu0 = (nu_min + (double)w_first * step) + 0.5 * (double)(kGroupPoints - 1) * step;
  t.dg = (float)((double)kGroupPoints * step);
  t.dH = (float)(0.5 * (double)(kWavePoints - kGroupPoints) * step);
  t.hg = (float)(0.5 * (double)kGroupPoints * step);
  return t;
}
// offset of group g's middle from group 0's, in FP32
__device__ __forceinline__ float group_offset(const GroupTest& t, int g) {
#pragma clang fp contract(off)
  return (float)g * t.dg;
}
// per record: d0 = (float)(nu0 - c) and thr = ds1 - (H - Hg)
__device__ __forceinline__ float group_d0(const GroupTest& t, double c) {
#pragma clang fp contract(off)
  return (float)(t.nu0 - c);
}
__device__ __forceinline__ float group_thr(const GroupTest& t, float ds1) {
#pragma clang fp contract(off)
  return ds1 - t.dH;
}
__device__ __forceinline__ bool group_one_node_holds(float d0, float thr, float goff) {
#pragma clang fp contract(off)
  return fabsf(d0 + goff) >= thr;
}
// Where the one-node form fails at a group, the group level sums the record's
// middle-tier form instead (the second-order shifted form for y^2 < kF32S2Y2,
// else the two-node form) if that holds at every point of the group: at
// distance |nu - c| >= dm (MixedRec), i.e. iff |nu_g - c| - (Hg - step/2) >=
// dm, which |nu_g - c| >= dm + Hg implies.  The test is group_one_node_holds's
// with this threshold.
__device__ __forceinline__ float group_mid_thr(const GroupTest& t, float dm) {
#pragma clang fp contract(off)
  return dm + t.hg;
}

// H = kWavePoints/2 * step in FP32, as the mixed records' ds1 and the series'
// unit sh = s H are formed.
__device__ __forceinline__ float wave_half_span32(double step) {
  return 0.5f * (float)kWavePoints * (float)step;
}

// Middle of the wavefront whose first grid point is i: grid(i) + (kWavePoints-1)/2 step.
__device__ __forceinline__ double wave_middle(double nu_min, double step, int64_t i) {
#pragma clang fp contract(off)
  return __dadd_rn(__dadd_rn(nu_min, __dmul_rn((double)i, step)),
                   0.5 * (double)(kWavePoints - 1) * step);
}
// Middle of the slice whose first grid point is i: grid(i) + (kSlicePoints-1)/2 step.
__device__ __forceinline__ double slice_middle(double nu_min, double step, int64_t i) {
#pragma clang fp contract(off)
  return __dadd_rn(__dadd_rn(nu_min, __dmul_rn((double)i, step)),
                   0.5 * (double)(kSlicePoints - 1) * step);
}

// Per-(layer, line) record of the mixed-precision paths (48 bytes, one
// 16-byte-aligned struct so the line sum stages it with three 16-byte loads).
// The first 16 bytes are what the line sum's per-record test reads, the next
// 16 what its far-field branches read:
//   c  shifted centre (FP64), s = sqrt(ln2)/alpha_D,
//   hw1 = sqrt(max(s1_r2_min(y^2) - y^2, 0)) + (kPointsMixed - 1)/2 * hs: a
//        thread's kPointsMixed points all lie where the shifted one-node form
//        holds (faddeeva.h) iff |x| at the middle of the run is >= hw1,
//   hs = step * s, c1 = y^2 - 3/2, ay = A y / sqrt(pi) (one-node form; the
//        two-node form's y^2 - 1/2 = c1 + 1 and 2 y^2 = 2 c1 + 3 are exact),
//   hw2 = the same for the middle tier: the second-order shifted form
//        (faddeeva.h s2_r2_min32) when y^2 < kF32S2Y2, else the two-node form
//        (|z| >= 25); the tier is chosen per record from c1,
//   y = sqrt(ln2) gamma_L / alpha_D, a = A = S sqrt(ln2/pi) / alpha_D,
//   ds1 = H + sqrt(max(s1_r2_min(y^2) - y^2, 0)) / s: the one-node form holds
//        at all kWavePoints points of a wavefront iff |nu_w - c| >= ds1,
//        at the group level the form holds at all points of a group iff
//        |nu_g - c| >= ds1 - (H - Hg),
//   dm = sqrt(max(r2_mid - y^2, 0)) / s, r2_mid the middle tier's (hw2):
//        the middle-tier form holds where |nu - c| >= dm (group_mid_thr).
// The far-field series' unit is sh = s H (H = kWavePoints/2 * step,
// wave_half_span32), sh/8 at the group level.  (The near field's LDS copy
// carries other values in ds1 and dm: the core form's rho and the mask of
// the wavefront's groups that evaluate the record point by point.)
// A skipped (zero-column) record has c = NaN, ay = a = 0.
struct alignas(16) MixedRec {
  double c;
  float s, hw1;
  float hs, c1, ay, hw2;
  float y, a, ds1, dm;
};
static_assert(sizeof(MixedRec) == 48, "record layout");

struct Records {
  // FP64 mode: SoA arrays [n_layers * n_lines]
  double* c;     // shifted line centre
  void* s;       // sqrt(ln2) / alpha_D
  void* y;       // sqrt(ln2) * gamma_L / alpha_D
  void* a;       // strength * sqrt(ln2/pi) / alpha_D (0 when the layer is skipped)
  // mixed / surrogate modes: AoS [n_layers * n_lines]
  MixedRec* mix;
};

struct LineParamsArgs {
  lbl_lines_t lines;
  lbl_groups_t groups;
  lbl_layers_t layers;
  Records rec;
  uint32_t* err;
  double step;   // grid step (for MixedRec::hs)
};

struct TileRangeArgs {
  const double* nu0;   // sorted ascending
  int64_t n_lines;
  double nu_min, step, cutoff, margin;
  int64_t n_nu, ts, n_tiles;
  int32_t* lo;
  int32_t* hi;
};

// Near ranges of the far-field series: [wlo, whi) are the records whose nu0
// lies within half_width of a wavefront's middle point, one entry per
// (workgroup slice, wavefront) in launch order; [slo, shi) those within
// slice_half_width of a slice's middle point (kSlicePoints points), one per
// slice; [glo, ghi) those within group_half_width of a group's middle point
// (kGroupPoints points), kGroupsPerWave per wavefront entry.
struct WaveRangeArgs {
  const double* nu0;   // sorted ascending
  int64_t n_lines;
  double nu_min, step, half_width, slice_half_width, group_half_width;
  int64_t ts, n_wg;
  int32_t slices_per_tile, waves_per_slice;
  int32_t* wlo;
  int32_t* whi;
  int32_t* slo;
  int32_t* shi;
  int32_t* glo;
  int32_t* ghi;
};

struct SumArgs {
  Records rec;
  const int32_t* lo;
  const int32_t* hi;
  const int32_t* wlo;   // near range of every wavefront (far-field series)
  const int32_t* whi;
  const int32_t* slo;   // near range of every slice
  const int32_t* shi;
  const int32_t* glo;   // near range of every group (kGroupsPerWave per wavefront)
  const int32_t* ghi;
  const double* t_k;
  int64_t n_lines;
  int32_t n_layers;
  double nu_min, step, cutoff;
  int64_t n_nu, ts, n_tiles;
  int32_t slices_per_tile;
  int64_t n_wg;
  int32_t n_mu;
  double inv_mu[kMaxMu];
  void* trans;
  void* tau;
  void* tau_layers;
  uint32_t* err;
  // surrogate mode only
  const float* mlp_params;
  const unsigned short* mlp_wbf;
  float mlp_gate;   // a point uses the network iff A * c * K_ref(x, y) <= mlp_gate
  unsigned long long* mlp_count;     // network evaluations (nullable)
  unsigned long long* exact_count;   // exact near-centre evaluations (nullable)
  // Layer groups: when the grid has too few tile slices to fill the GPU, the
  // layers are split into layer_groups ranges, one workgroup per (slice,
  // range); each writes its partial tau to tau_part[group][point] (mixed:
  // float sum and Kahan compensation; FP64: double) and launch_tau_epilogue
  // adds them and writes T (and tau).  layer_groups == 1: T is written by the
  // line sum itself.
  int32_t layer_groups;
  void* tau_part;
  // far-field sum per (layer, point), FP32 (mixed modes) or FP64, written
  // by k_far_series_* and added by the line sum
  void* far_sum;
  // when set, the counting variants of the line-sum kernels add their work
  // units to these kWorkSlots counters (same results otherwise)
  unsigned long long* work;
  // mixed modes: per (layer, wavefront) [n_layers][n_wg * 4], the number of
  // (record, group) pairs the group level left to the near field (written by
  // k_far_series_mixed for every live wavefront, read by the near field)
  uint32_t* deferred;
};

// Work units the line-sum kernels count in their counting variant:
//   near-field point evaluations (records inside the tile window x points of
//   the wavefront), (record, slice) and (record, wavefront) series pairs,
//   records that fail the slice level's validity test (passed on to the
//   wavefront level) and the wavefront level's (evaluated point by point,
//   and their point evaluations), queued (record, lane) pairs, queued core
//   points (mixed), slice-list overflows (rescans), and the wavefront-level
//   candidates taken from the slice level's list (each of its records is one
//   for every live wavefront of the slice).
enum WorkCount {
  kWorkNearEvals = 0,
  kWorkSliceSeries,
  kWorkWaveSeries,
  kWorkSliceExceptions,
  kWorkWaveExceptions,
  kWorkExceptionEvals,
  kWorkQueued,
  kWorkCorePoints,
  kWorkRescans,
  kWorkWaveFromSliceList,
  kWorkGroupSeries,         // (record, group) pairs summed as group series
  kWorkGroupDeferred,       // ... left to the near field (the form fails at the group)
  kWorkGroupDeferredEvals,  // the near field's point evaluations of those
  kWorkGroupMiddle,         // group series pairs in the middle-tier form (part of kWorkGroupSeries)
  kWorkSeriesTerms,         // terms summed: mixed, over the one-node pairs of all levels;
                            // FP64, over the slice level's pairs
  kWorkSeriesTerms2,        // FP64: terms of each wavefront- and group-level pair (both pole
                            // pairs); mixed: terms of the middle-tier group pairs
  kWorkSlots
};

// Workgroups a line-sum launch aims for.  Many short workgroups (a layer
// range of a tile slice each) let the hardware scheduler balance the uneven
// tiles; fewer layer groups cost fewer workgroup prologues and less traffic
// of partial optical depths.  On the five-gas workload (1001 slices, 50
// layers) the mixed near field (with the epilogue) measured 5.39 / 5.27 ms
// at 50 / 25 layer groups (one / two layers each; uneven ranges such as 20,
// 30 or 40 groups of one to three layers measured 5.6-6.1 ms), the FP64 one
// 24.8 / 24.4 ms at 17 / 25 groups.
constexpr int64_t kTargetWorkgroupsMixed = 25000;
constexpr int64_t kTargetWorkgroupsFp64 = 25000;

// Number of layer ranges for a grid of n_wg tile slices.
inline int32_t layer_groups_for(int64_t n_wg, int32_t n_layers, int precision) {
  const int64_t target = precision == LBL_FP64 ? kTargetWorkgroupsFp64 : kTargetWorkgroupsMixed;
  if (n_wg <= 0 || n_layers <= 1 || n_wg >= target) return 1;
  const int64_t g = (target + n_wg - 1) / n_wg;
  return (int32_t)std::min<int64_t>(g, n_layers);
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
hipError_t launch_line_params(const LineParamsArgs& a, int precision, hipStream_t s);
hipError_t launch_tile_ranges(const TileRangeArgs& a, hipStream_t s);
hipError_t launch_wave_ranges(const WaveRangeArgs& a, hipStream_t s);
// mixed modes: k_far_series_mixed, before launch_voigt_sum (no-op for FP64)
hipError_t launch_far_series(const SumArgs& a, int precision, hipStream_t s);
hipError_t launch_voigt_sum(const SumArgs& a, int precision, hipStream_t s);
// Sum of the layer-group partials, T = exp(-tau/mu) (only when layer_groups > 1).
hipError_t launch_tau_epilogue(const SumArgs& a, int precision, hipStream_t s);
hipError_t launch_voigt_eval(int64_t n, const void* x, const void* y, void* k, int precision,
                             hipStream_t s);

hipError_t launch_line_grad(const lbl_lines_t& lines, const lbl_groups_t& groups,
                            const lbl_layers_t& layers, float* at, float* dyt, float* av,
                            float* bxv, float* dyv, hipStream_t s);
hipError_t launch_vjp(const Records& rec, const float* at, const float* dyt, const float* av,
                      const float* bxv, const float* dyv, const int32_t* group,
                      const int32_t* gas_of_group, const int32_t* lo, const int32_t* hi,
                      const int32_t* wlo, const int32_t* whi, const int32_t* slo,
                      const int32_t* shi, const double* t_k, int64_t n_lines, int n_layers,
                      int n_gases,
                      const lbl_grid_t& grid, int64_t n_tiles, int n_mu, const double* mu,
                      const float* trans, const float* grad_trans, const float* od_layers,
                      double* grad_t, double* grad_vmr, hipStream_t s);
int vjp_points_per_thread();

int mlp_num_params();
hipError_t launch_mlp_make_data(int64_t n, uint64_t seed, float* x, float* ly, float* t,
                                hipStream_t s);
hipError_t launch_mlp_pack(const float* p, unsigned short* wbf, hipStream_t s);
hipError_t launch_mlp_step(float* p, float* m, float* v, unsigned short* wbf, const float* x,
                           const float* ly, const float* t, int64_t n_data, uint64_t perm_a,
                           uint64_t perm_b, int64_t base, int batch, float lr, int adam_t,
                           float* partial, float* loss_partial, float* grad, float* loss,
                           hipStream_t s);
hipError_t launch_mlp_eval(int64_t n, const float* x, const float* ly, const float* p,
                           const unsigned short* wbf, float* t, hipStream_t s);

// Points per thread and threads per workgroup of the line-sum kernels (a
// workgroup covers block * P consecutive grid points of one reference tile).
int sum_points_per_thread(int precision);
int sum_block(int precision);

}  // namespace lbl

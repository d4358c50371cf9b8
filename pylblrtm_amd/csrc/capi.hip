// extern "C" entry points of libpylblrtm_amd (declared in include/pylblrtm_amd.h).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "internal.h"

namespace {

thread_local char g_last_error[512] = "";

// Optional per-kernel timing of lbl_optical_depth with HIP events recorded on
// the caller's stream (one event set per device, created on first use).
constexpr int kMaxDevices = 64;
bool g_timing = false;
constexpr int kEvents = 6;
hipEvent_t g_events[kMaxDevices][kEvents];
bool g_events_made[kMaxDevices] = {};
thread_local double g_ms_line_params = 0.0, g_ms_voigt_sum = 0.0;
thread_local double g_ms_far = 0.0, g_ms_near = 0.0, g_ms_epilogue = 0.0;
// Work counting (lbl_enable_work_count): the line sums run their counting
// variants and the counts of the calling thread's last call are kept here.
bool g_count = false;
thread_local uint64_t g_work[lbl::kWorkSlots] = {};

hipError_t timing_events(hipEvent_t** ev) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_events_made[dev]) {
    for (int i = 0; i < kEvents; ++i) {
      e = hipEventCreate(&g_events[dev][i]);
      if (e != hipSuccess) return e;
    }
    g_events_made[dev] = true;
  }
  *ev = g_events[dev];
  return hipSuccess;
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

// Scratch of lbl_optical_depth_noalloc: a fixed device array in this code
// object, one copy per GPU the library runs on, held until the process ends.
// Host code never passes the array's name to a kernel (on the host it is only
// a stand-in symbol): its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = LBL_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (looked up on the
// first call on that device, with that device current).
int scratch_of_current_device(void** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(LBL_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(LBL_EHIP, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(LBL_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return LBL_OK;
}

}  // namespace

namespace lbl {

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace lbl

namespace {

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) return fail(LBL_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Layout {
  size_t c, s, y, a, lo, hi, wlo, whi, slo, shi, glo, ghi, err, work, deferred, tau_part, far, total;
  size_t g[5];   // adjoint coefficient arrays (only with_grad)
  int64_t n_tiles, n_wg;
  int32_t slices_per_tile, layer_groups;
};

// Largest grid the size arithmetic below handles without overflow (with at
// most 2^19 layers and 2^31 lines every product stays below 2^63 bytes).
constexpr int64_t kMaxPoints = (int64_t)1 << 40;

bool make_layout(const lbl_lines_t* lines, const lbl_layers_t* layers, const lbl_grid_t* grid,
                 int precision, Layout* L, bool with_grad = false) {
  if (!lines || !layers || !grid) return false;
  if (grid->tile_samples <= 0 || grid->n_nu <= 0 || grid->n_nu > kMaxPoints) return false;
  if (lines->n_lines < 0 || lines->n_lines > INT32_MAX) return false;
  if (layers->n_layers < 0 || layers->n_layers > (1 << 19)) return false;
  if (precision != LBL_FP64 && precision != LBL_MIXED && precision != LBL_SURROGATE) return false;
  const size_t nrec = (size_t)std::max<int64_t>(lines->n_lines, 0) *
                      (size_t)std::max<int32_t>(layers->n_layers, 0);
  const bool fp64 = precision == LBL_FP64;
  L->n_tiles = (grid->n_nu + grid->tile_samples - 1) / grid->tile_samples;
  const int64_t pts = (int64_t)lbl::sum_block(precision) * lbl::sum_points_per_thread(precision);
  L->slices_per_tile = (int32_t)((grid->tile_samples + pts - 1) / pts);
  L->n_wg = L->n_tiles * L->slices_per_tile;
  L->layer_groups = lbl::layer_groups_for(L->n_wg, std::max<int32_t>(layers->n_layers, 0), precision);
  size_t off = 0;
  // FP64: SoA c, s, y, a (8 B each); mixed: one 48-byte MixedRec per record at L->c
  L->c = off; off = align_up(off + nrec * (fp64 ? 8 : sizeof(lbl::MixedRec)));
  L->s = off; off = align_up(off + (fp64 ? nrec * 8 : 0));
  L->y = off; off = align_up(off + (fp64 ? nrec * 8 : 0));
  L->a = off; off = align_up(off + (fp64 ? nrec * 8 : 0));
  L->lo = off; off = align_up(off + (size_t)L->n_tiles * 4);
  L->hi = off; off = align_up(off + (size_t)L->n_tiles * 4);
  const size_t n_waves = (size_t)L->n_wg * (size_t)(lbl::sum_block(precision) / 64);
  L->wlo = off; off = align_up(off + n_waves * 4);
  L->whi = off; off = align_up(off + n_waves * 4);
  L->slo = off; off = align_up(off + (size_t)L->n_wg * 4);
  L->shi = off; off = align_up(off + (size_t)L->n_wg * 4);
  L->glo = off; off = align_up(off + n_waves * lbl::kGroupsPerWave * 4);
  L->ghi = off; off = align_up(off + n_waves * lbl::kGroupsPerWave * 4);
  L->err = off; off = align_up(off + 4);
  L->work = off; off = align_up(off + lbl::kWorkSlots * sizeof(unsigned long long));
  // mixed modes: the group level's left-over pair count per (layer, wavefront)
  L->deferred = off;
  if (!with_grad)
    off = align_up(off + (size_t)std::max<int32_t>(layers->n_layers, 0) * n_waves * 4);
  // partial tau per layer group (float sum + compensation, or double): 8 B
  L->tau_part = off;
  // (the line sums keep their column totals here even with one group)
  if (L->layer_groups > 1 || !with_grad)
    off = align_up(off + (size_t)std::max<int32_t>(L->layer_groups, 1) * (size_t)grid->n_nu * 8);
  // far-field sum per (layer, point): FP32 (mixed modes) or FP64
  L->far = off;
  if (!with_grad)
    off = align_up(off + (size_t)std::max<int32_t>(layers->n_layers, 0) * (size_t)grid->n_nu *
                             (fp64 ? 8 : 4));
  for (int i = 0; i < 5; ++i) {
    L->g[i] = off;
    if (with_grad) off = align_up(off + nrec * 4);
  }
  L->total = off;
  return true;
}

int check_inputs(const lbl_lines_t* lines, const lbl_groups_t* groups, const lbl_layers_t* layers,
                 const lbl_grid_t* grid, int precision) {
  if (!lines || !groups || !layers || !grid) return fail(LBL_EINVAL, "null argument");
  if (precision != LBL_FP64 && precision != LBL_MIXED && precision != LBL_SURROGATE)
    return fail(LBL_EINVAL, "unknown precision %d", precision);
  if (lines->n_lines < 0 || lines->n_lines > INT32_MAX)
    return fail(LBL_EINVAL, "n_lines %lld out of range", (long long)lines->n_lines);
  if (layers->n_layers < 0 || layers->n_layers > (1 << 19))
    return fail(LBL_EINVAL, "n_layers %d out of range", layers->n_layers);
  if (groups->n_groups < 0 || groups->n_groups > lbl::kMaxGroups)
    return fail(LBL_EINVAL, "n_groups %d exceeds %d", groups->n_groups, lbl::kMaxGroups);
  if (grid->n_nu < 1 || grid->tile_samples < 1)
    return fail(LBL_EINVAL, "empty grid or tile");
  if (grid->n_nu > kMaxPoints)
    return fail(LBL_EINVAL, "n_nu %lld exceeds 2^40 points", (long long)grid->n_nu);
  if (!(grid->step > 0.0) || !isfinite(grid->nu_min) || !isfinite(grid->cutoff))
    return fail(LBL_EINVAL, "invalid grid parameters");
  if (lines->n_lines > 0) {
    if (!lines->nu0 || !lines->sref || !lines->g_air || !lines->g_self || !lines->elow ||
        !lines->n_air || !lines->shift || !lines->group)
      return fail(LBL_EINVAL, "null line array");
    if (groups->n_groups < 1 || !groups->gas || !groups->molar_mass || !groups->tips_offset ||
        !groups->tips_t || !groups->tips_q)
      return fail(LBL_EINVAL, "null group array");
  }
  if (layers->n_layers > 0 && (!layers->p_atm || !layers->t_k))
    return fail(LBL_EINVAL, "null layer array");
  if (lines->n_lines > 0 && layers->n_layers > 0 && (!layers->vmr || !layers->col))
    return fail(LBL_EINVAL, "null vmr/col array");
  if (lines->n_lines > 0 && layers->n_layers > 0 && layers->n_gases < 1)
    return fail(LBL_EINVAL, "n_gases must be >= 1 when there are lines and layers");
  return LBL_OK;
}

// A tile at least as long as the grid is the whole grid: the kernels get
// tile_samples = min(tile_samples, n_nu), the same tiling without the empty
// workgroups of the part of the tile past the grid's end.
lbl_grid_t effective_grid(const lbl_grid_t* grid) {
  lbl_grid_t g = *grid;
  if (g.tile_samples > g.n_nu) g.tile_samples = g.n_nu;
  return g;
}

}  // namespace

extern "C" {

const char* lbl_last_error(void) { return g_last_error; }

void lbl_enable_timing(int on) { g_timing = on != 0; }

int lbl_last_timings(double* ms_line_params, double* ms_voigt_sum) {
  if (ms_line_params) *ms_line_params = g_ms_line_params;
  if (ms_voigt_sum) *ms_voigt_sum = g_ms_voigt_sum;
  return LBL_OK;
}

int lbl_last_timing_epilogue(double* ms_epilogue) {
  if (ms_epilogue) *ms_epilogue = g_ms_epilogue;
  return LBL_OK;
}

int lbl_last_timings_split(double* ms_far_series, double* ms_near_sum) {
  if (ms_far_series) *ms_far_series = g_ms_far;
  if (ms_near_sum) *ms_near_sum = g_ms_near;
  return LBL_OK;
}

void lbl_enable_work_count(int on) { g_count = on != 0; }

int lbl_last_work_count(uint64_t* counts, int32_t n) {
  if (!counts || n < 0) return fail(LBL_EINVAL, "lbl_last_work_count: null counts");
  for (int i = 0; i < n; ++i) counts[i] = i < lbl::kWorkSlots ? g_work[i] : 0;
  return LBL_OK;
}

int lbl_workspace_bytes(const lbl_lines_t* lines, const lbl_layers_t* layers,
                        const lbl_grid_t* grid, int precision, size_t* bytes) {
  lbl_grid_t eg{};
  if (grid) { eg = effective_grid(grid); grid = &eg; }
  Layout L;
  if (!bytes || !make_layout(lines, layers, grid, precision, &L))
    return fail(LBL_EINVAL, "invalid arguments to lbl_workspace_bytes");
  *bytes = L.total;
  return LBL_OK;
}

// Near ranges of the far-field series: per (slice, wavefront), per slice and
// per group, the lines whose nu0 lies within kSeriesRatio half-spans (+ the
// shift bound) of the unit's middle point.
static hipError_t launch_near_ranges(const lbl_lines_t* lines, const lbl_grid_t* grid,
                                     double margin, const Layout& L, int precision, char* ws,
                                     hipStream_t s) {
  lbl::WaveRangeArgs wr{};
  wr.nu0 = lines->nu0;
  wr.n_lines = lines->n_lines;
  wr.nu_min = grid->nu_min;
  wr.step = grid->step;
  wr.half_width = lbl::kSeriesRatio * 0.5 * lbl::kWavePoints * grid->step + margin;
  wr.slice_half_width = lbl::kSeriesRatio * 0.5 * lbl::kSlicePoints * grid->step + margin;
  wr.group_half_width = lbl::kSeriesRatio * 0.5 * lbl::kGroupPoints * grid->step + margin;
  wr.ts = grid->tile_samples;
  wr.n_wg = L.n_wg;
  wr.slices_per_tile = L.slices_per_tile;
  wr.waves_per_slice = lbl::sum_block(precision) / 64;
  wr.wlo = reinterpret_cast<int32_t*>(ws + L.wlo);
  wr.whi = reinterpret_cast<int32_t*>(ws + L.whi);
  wr.slo = reinterpret_cast<int32_t*>(ws + L.slo);
  wr.shi = reinterpret_cast<int32_t*>(ws + L.shi);
  wr.glo = reinterpret_cast<int32_t*>(ws + L.glo);
  wr.ghi = reinterpret_cast<int32_t*>(ws + L.ghi);
  return lbl::launch_wave_ranges(wr, s);
}

// An index or table the kernels flagged as unusable (the error word's
// lowest codes): LBL_EINVAL, nothing was computed from it.
static int invalid_index(uint32_t code) {
  switch (code) {
    case lbl::kErrBadGroup:
      return fail(LBL_EINVAL, "a line's group index is outside [0, n_groups)");
    case lbl::kErrBadGas:
      return fail(LBL_EINVAL, "a group's gas index is outside [0, n_gases)");
    default:
      return fail(LBL_EINVAL, "a group's TIPS table is empty (tips_offset not increasing)");
  }
}

// Argument checks shared by the optical-depth entry points (host only).
static int check_call(const lbl_lines_t* lines, const lbl_groups_t* groups,
                      const lbl_layers_t* layers, const lbl_grid_t* grid, int precision,
                      const lbl_outputs_t* out, const lbl_surrogate_t* sur) {
  int rc = check_inputs(lines, groups, layers, grid, precision);
  if (rc != LBL_OK) return rc;
  if (precision == LBL_SURROGATE && (!sur || !sur->params || !sur->wbf16 || !(sur->gate >= 0.0)))
    return fail(LBL_EINVAL, "precision LBL_SURROGATE needs network parameters and a gate >= 0");
  if (!out || out->n_mu < 1 || out->n_mu > lbl::kMaxMu || !out->mu || !out->transmittance)
    return fail(LBL_EINVAL, "outputs: need 1..%d viewing cosines and a transmittance buffer",
                lbl::kMaxMu);
  for (int m = 0; m < out->n_mu; ++m)
    if (!(out->mu[m] > 0.0 && out->mu[m] <= 1.0))
      return fail(LBL_EINVAL, "mu_view must be in (0, 1]");
  return LBL_OK;
}

// The whole forward path on a workspace of at least make_layout's total
// bytes (arguments already checked by check_call).  allow_timing: record
// the per-kernel HIP events when lbl_enable_timing is on (their one-time
// creation is left out of the entry point that must not allocate).
static int optical_depth_impl(const lbl_lines_t* lines, const lbl_groups_t* groups,
                              const lbl_layers_t* layers, const lbl_grid_t* grid, int precision,
                              const lbl_outputs_t* out, const lbl_surrogate_t* sur,
                              void* workspace, size_t workspace_bytes, int32_t* err_layer,
                              int32_t* err_group, void* stream, bool allow_timing) {
  Layout L;
  make_layout(lines, layers, grid, precision, &L);
  if (!workspace || workspace_bytes < L.total)
    return fail(LBL_EINVAL, "workspace of %zu bytes needed, %zu given", L.total, workspace_bytes);
  if (L.n_tiles > INT32_MAX) return fail(LBL_EINVAL, "too many tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* ws = static_cast<char*>(workspace);

  lbl::Records rec{reinterpret_cast<double*>(ws + L.c), ws + L.s, ws + L.y, ws + L.a,
                   reinterpret_cast<lbl::MixedRec*>(ws + L.c)};
  uint32_t* err = reinterpret_cast<uint32_t*>(ws + L.err);
  int32_t* lo = reinterpret_cast<int32_t*>(ws + L.lo);
  int32_t* hi = reinterpret_cast<int32_t*>(ws + L.hi);

  hipEvent_t* ev = nullptr;
  if (g_timing && allow_timing) HIP_TRY(timing_events(&ev));

  HIP_TRY(lbl::launch_fill_u32(err, 1, lbl::kErrNone, s));

  lbl::LineParamsArgs lp{*lines, *groups, *layers, rec, err, grid->step};
  if (ev) HIP_TRY(hipEventRecord(ev[0], s));
  HIP_TRY(lbl::launch_line_params(lp, precision, s));
  if (ev) HIP_TRY(hipEventRecord(ev[1], s));

  // |centre - nu0| <= |shift| * P_air <= shift_bound: the candidate range per
  // tile is widened by it (and a rounding slack); the per-record test is exact.
  lbl::TileRangeArgs tr{};
  tr.nu0 = lines->nu0;
  tr.n_lines = lines->n_lines;
  tr.nu_min = grid->nu_min;
  tr.step = grid->step;
  tr.cutoff = grid->cutoff;
  tr.n_nu = grid->n_nu;
  tr.ts = grid->tile_samples;
  tr.n_tiles = L.n_tiles;
  tr.lo = lo;
  tr.hi = hi;
  tr.margin = fabs(grid->shift_bound) * (1.0 + 1e-9) + 1e-9;
  HIP_TRY(lbl::launch_tile_ranges(tr, s));

  // near range of every wavefront for the far-field series: records whose
  // nu0 is within kSeriesRatio * H (+ the shift bound) of its middle point
  int32_t* wlo = reinterpret_cast<int32_t*>(ws + L.wlo);
  int32_t* whi = reinterpret_cast<int32_t*>(ws + L.whi);
  int32_t* slo = reinterpret_cast<int32_t*>(ws + L.slo);
  int32_t* shi = reinterpret_cast<int32_t*>(ws + L.shi);
  HIP_TRY(launch_near_ranges(lines, grid, tr.margin, L, precision, ws, s));

  lbl::SumArgs sa{};
  sa.rec = rec;
  sa.lo = lo;
  sa.hi = hi;
  sa.wlo = wlo;
  sa.whi = whi;
  sa.glo = reinterpret_cast<const int32_t*>(ws + L.glo);
  sa.ghi = reinterpret_cast<const int32_t*>(ws + L.ghi);
  sa.slo = slo;
  sa.shi = shi;
  sa.t_k = layers->t_k;
  sa.n_lines = lines->n_lines;
  sa.n_layers = layers->n_layers;
  sa.nu_min = grid->nu_min;
  sa.step = grid->step;
  sa.cutoff = grid->cutoff;
  sa.n_nu = grid->n_nu;
  sa.ts = grid->tile_samples;
  sa.n_tiles = L.n_tiles;
  sa.slices_per_tile = L.slices_per_tile;
  sa.n_wg = L.n_wg;
  if (sa.n_wg > INT32_MAX) return fail(LBL_EINVAL, "grid too large");
  sa.layer_groups = L.layer_groups;
  sa.tau_part = ws + L.tau_part;
  sa.far_sum = ws + L.far;
  sa.deferred = reinterpret_cast<uint32_t*>(ws + L.deferred);
  sa.n_mu = out->n_mu;
  for (int m = 0; m < out->n_mu; ++m) sa.inv_mu[m] = 1.0 / out->mu[m];
  sa.trans = out->transmittance;
  sa.tau = out->optical_depth;
  sa.tau_layers = out->optical_depth_layers;
  sa.err = err;
  if (precision == LBL_SURROGATE) {
    sa.mlp_params = sur->params;
    sa.mlp_wbf = sur->wbf16;
    sa.mlp_gate = (float)sur->gate;
    sa.mlp_count = reinterpret_cast<unsigned long long*>(sur->n_network);
    sa.exact_count = reinterpret_cast<unsigned long long*>(sur->n_exact_near);
  }
  unsigned long long* work = nullptr;
  if (g_count && allow_timing && precision != LBL_SURROGATE) {
    work = reinterpret_cast<unsigned long long*>(ws + L.work);
    HIP_TRY(lbl::launch_fill_u32(reinterpret_cast<uint32_t*>(work), 2 * lbl::kWorkSlots, 0u, s));
    sa.work = work;
  }
  if (ev) HIP_TRY(hipEventRecord(ev[2], s));
  HIP_TRY(lbl::launch_far_series(sa, precision, s));
  if (ev) HIP_TRY(hipEventRecord(ev[4], s));
  HIP_TRY(lbl::launch_voigt_sum(sa, precision, s));
  if (ev) HIP_TRY(hipEventRecord(ev[5], s));
  HIP_TRY(lbl::launch_tau_epilogue(sa, precision, s));
  if (ev) HIP_TRY(hipEventRecord(ev[3], s));

  uint32_t code = lbl::kErrNone;
  HIP_TRY(hipMemcpyAsync(&code, err, sizeof(code), hipMemcpyDeviceToHost, s));
  if (work) HIP_TRY(hipMemcpyAsync(g_work, work, sizeof(g_work), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (!work) memset(g_work, 0, sizeof(g_work));
  if (ev) {
    float a = 0.f, b = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, ev[2], ev[3]));
    g_ms_line_params = a;
    g_ms_voigt_sum = b;
    float f = 0.f, n = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, ev[2], ev[4]));
    HIP_TRY(hipEventElapsedTime(&n, ev[4], ev[3]));
    g_ms_far = f;
    g_ms_near = n;
    float ep = 0.f;
    HIP_TRY(hipEventElapsedTime(&ep, ev[5], ev[3]));
    g_ms_epilogue = ep;
  }
  if (code != lbl::kErrNone && code < (1u << lbl::kErrGroupBits)) return invalid_index(code);
  if (code != lbl::kErrNone) {
    const int kind2 = (int)(code >> lbl::kErrGroupBits) - 1;
    const int layer = kind2 >> 1;
    const int group = (int)(code & ((1u << lbl::kErrGroupBits) - 1));
    if (err_layer) *err_layer = layer;
    if (err_group) *err_group = (kind2 & 1) ? -1 : group;
    if (kind2 & 1) return fail(LBL_ENONFINITE, "Layer %d: non-finite optical depth", layer);
    return fail(LBL_ESTRENGTH, "Layer %d, group %d: non-finite or negative line strength",
                layer, group);
  }
  return LBL_OK;
}

int lbl_optical_depth(const lbl_lines_t* lines, const lbl_groups_t* groups,
                      const lbl_layers_t* layers, const lbl_grid_t* grid, int precision,
                      const lbl_outputs_t* out, void* workspace, size_t workspace_bytes,
                      int32_t* err_layer, int32_t* err_group, void* stream) {
  g_last_error[0] = '\0';
  if (precision == LBL_SURROGATE)
    return fail(LBL_EINVAL, "precision LBL_SURROGATE needs lbl_optical_depth_surrogate");
  int rc = check_call(lines, groups, layers, grid, precision, out, nullptr);
  if (rc != LBL_OK) return rc;
  const lbl_grid_t eg = effective_grid(grid);
  grid = &eg;
  return optical_depth_impl(lines, groups, layers, grid, precision, out, nullptr, workspace,
                            workspace_bytes, err_layer, err_group, stream, true);
}

int lbl_optical_depth_noalloc(const lbl_lines_t* lines, const lbl_groups_t* groups,
                              const lbl_layers_t* layers, const lbl_grid_t* grid, int precision,
                              const lbl_outputs_t* out, int32_t* err_layer, int32_t* err_group,
                              void* stream) {
  g_last_error[0] = '\0';
  if (precision != LBL_FP64 && precision != LBL_MIXED)
    return fail(LBL_EINVAL, "lbl_optical_depth_noalloc: precision must be LBL_FP64 or LBL_MIXED");
  int rc = check_call(lines, groups, layers, grid, precision, out, nullptr);
  if (rc != LBL_OK) return rc;
  const lbl_grid_t eg = effective_grid(grid);
  grid = &eg;
  Layout L;
  make_layout(lines, layers, grid, precision, &L);
  if (L.total > kScratchBytes)
    return fail(LBL_EINVAL,
                "lbl_optical_depth_noalloc: this problem needs %zu bytes of scratch, the fixed "
                "device array holds %zu; call lbl_optical_depth with a workspace instead",
                L.total, kScratchBytes);
  void* ptr = nullptr;
  rc = scratch_of_current_device(&ptr);
  if (rc != LBL_OK) return rc;
  return optical_depth_impl(lines, groups, layers, grid, precision, out, nullptr, ptr,
                            kScratchBytes, err_layer, err_group, stream, false);
}

int lbl_optical_depth_surrogate(const lbl_lines_t* lines, const lbl_groups_t* groups,
                                const lbl_layers_t* layers, const lbl_grid_t* grid,
                                const lbl_outputs_t* out, const lbl_surrogate_t* sur,
                                void* workspace, size_t workspace_bytes, int32_t* err_layer,
                                int32_t* err_group, void* stream) {
  g_last_error[0] = '\0';
  int rc = check_call(lines, groups, layers, grid, LBL_SURROGATE, out, sur);
  if (rc != LBL_OK) return rc;
  const lbl_grid_t eg = effective_grid(grid);
  grid = &eg;
  return optical_depth_impl(lines, groups, layers, grid, LBL_SURROGATE, out, sur, workspace,
                            workspace_bytes, err_layer, err_group, stream, true);
}

int lbl_vjp_workspace_bytes(const lbl_lines_t* lines, const lbl_layers_t* layers,
                            const lbl_grid_t* grid, size_t* bytes) {
  lbl_grid_t eg{};
  if (grid) { eg = effective_grid(grid); grid = &eg; }
  Layout L;
  if (!bytes || !make_layout(lines, layers, grid, LBL_MIXED, &L, true))
    return fail(LBL_EINVAL, "invalid arguments to lbl_vjp_workspace_bytes");
  *bytes = L.total;
  return LBL_OK;
}

int lbl_transmittance_vjp(const lbl_lines_t* lines, const lbl_groups_t* groups,
                          const lbl_layers_t* layers, const lbl_grid_t* grid, int32_t n_mu,
                          const double* mu, const float* transmittance,
                          const float* optical_depth_layers, const float* grad_transmittance,
                          double* grad_t, double* grad_vmr, void* workspace,
                          size_t workspace_bytes, void* stream) {
  g_last_error[0] = '\0';
  int rc = check_inputs(lines, groups, layers, grid, LBL_MIXED);
  if (rc != LBL_OK) return rc;
  const lbl_grid_t eg = effective_grid(grid);
  grid = &eg;
  if (n_mu < 1 || n_mu > lbl::kMaxMu || !mu || !transmittance || !optical_depth_layers ||
      !grad_transmittance || !grad_t || !grad_vmr)
    return fail(LBL_EINVAL, "vjp: need 1..%d viewing cosines and all gradient buffers",
                lbl::kMaxMu);
  for (int m = 0; m < n_mu; ++m)
    if (!(mu[m] > 0.0 && mu[m] <= 1.0)) return fail(LBL_EINVAL, "mu_view must be in (0, 1]");
  if (layers->n_gases > 8) return fail(LBL_EINVAL, "vjp supports at most 8 gases");
  Layout L;
  make_layout(lines, layers, grid, LBL_MIXED, &L, true);
  if (!workspace || workspace_bytes < L.total)
    return fail(LBL_EINVAL, "workspace of %zu bytes needed, %zu given", L.total, workspace_bytes);
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* ws = static_cast<char*>(workspace);
  lbl::Records rec{nullptr, nullptr, nullptr, nullptr, reinterpret_cast<lbl::MixedRec*>(ws + L.c)};
  uint32_t* err = reinterpret_cast<uint32_t*>(ws + L.err);
  int32_t* lo = reinterpret_cast<int32_t*>(ws + L.lo);
  int32_t* hi = reinterpret_cast<int32_t*>(ws + L.hi);
  float* gc[5];
  for (int i = 0; i < 5; ++i) gc[i] = reinterpret_cast<float*>(ws + L.g[i]);

  HIP_TRY(lbl::launch_fill_u32(err, 1, lbl::kErrNone, s));
  lbl::LineParamsArgs lp{*lines, *groups, *layers, rec, err, grid->step};
  HIP_TRY(lbl::launch_line_params(lp, LBL_MIXED, s));
  // the gradient kernels index groups and gases directly: stop here if the
  // per-(layer, line) pass found an index it could not use
  uint32_t code = lbl::kErrNone;
  HIP_TRY(hipMemcpyAsync(&code, err, sizeof(code), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (code != lbl::kErrNone && code < (1u << lbl::kErrGroupBits)) return invalid_index(code);
  lbl::TileRangeArgs tr{};
  tr.nu0 = lines->nu0;
  tr.n_lines = lines->n_lines;
  tr.nu_min = grid->nu_min;
  tr.step = grid->step;
  tr.cutoff = grid->cutoff;
  tr.n_nu = grid->n_nu;
  tr.ts = grid->tile_samples;
  tr.n_tiles = L.n_tiles;
  tr.lo = lo;
  tr.hi = hi;
  tr.margin = fabs(grid->shift_bound) * (1.0 + 1e-9) + 1e-9;
  HIP_TRY(lbl::launch_tile_ranges(tr, s));
  // near ranges of the far-field series, as for the forward's mixed line sum
  int32_t* wlo = reinterpret_cast<int32_t*>(ws + L.wlo);
  int32_t* whi = reinterpret_cast<int32_t*>(ws + L.whi);
  int32_t* slo = reinterpret_cast<int32_t*>(ws + L.slo);
  int32_t* shi = reinterpret_cast<int32_t*>(ws + L.shi);
  HIP_TRY(launch_near_ranges(lines, grid, tr.margin, L, LBL_MIXED, ws, s));
  HIP_TRY(lbl::launch_line_grad(*lines, *groups, *layers, gc[0], gc[1], gc[2], gc[3], gc[4], s));
  HIP_TRY(lbl::launch_vjp(rec, gc[0], gc[1], gc[2], gc[3], gc[4], lines->group, groups->gas, lo,
                          hi, wlo, whi, slo, shi, layers->t_k, lines->n_lines, layers->n_layers,
                          layers->n_gases,
                          *grid, L.n_tiles, n_mu, mu, transmittance, grad_transmittance,
                          optical_depth_layers, grad_t, grad_vmr, s));
  HIP_TRY(hipMemcpyAsync(&code, err, sizeof(code), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (code != lbl::kErrNone)
    return fail(LBL_ESTRENGTH, "vjp: non-finite or negative line strength (layer %d)",
                ((int)(code >> lbl::kErrGroupBits) - 1) >> 1);
  return LBL_OK;
}

int lbl_mlp_num_params(void) { return lbl::mlp_num_params(); }

int lbl_mlp_make_data(int64_t n, uint64_t seed, float* x, float* log10y, float* target,
                      void* stream) {
  g_last_error[0] = '\0';
  if (n < 0 || (n > 0 && (!x || !log10y || !target)))
    return fail(LBL_EINVAL, "invalid arguments to lbl_mlp_make_data");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(lbl::launch_mlp_make_data(n, seed, x, log10y, target, s));
  HIP_TRY(hipStreamSynchronize(s));
  return LBL_OK;
}

int lbl_mlp_pack(const float* params, unsigned short* wbf16, void* stream) {
  g_last_error[0] = '\0';
  if (!params || !wbf16) return fail(LBL_EINVAL, "invalid arguments to lbl_mlp_pack");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(lbl::launch_mlp_pack(params, wbf16, s));
  HIP_TRY(hipStreamSynchronize(s));
  return LBL_OK;
}

int lbl_mlp_train_workspace_bytes(int batch, size_t* bytes) {
  if (!bytes || batch <= 0 || batch % 64) return fail(LBL_EINVAL, "batch must be a positive multiple of 64");
  const size_t rows = (size_t)batch / 64;
  *bytes = align_up(rows * lbl::mlp_num_params() * sizeof(float)) + align_up(rows * sizeof(float));
  return LBL_OK;
}

int lbl_mlp_train_step(float* params, float* adam_m, float* adam_v, unsigned short* wbf16,
                       const float* x, const float* log10y, const float* target, int64_t n_data,
                       uint64_t perm_a, uint64_t perm_b, int64_t base, int batch, float lr,
                       int adam_t, float* grad, float* loss, void* workspace,
                       size_t workspace_bytes, void* stream) {
  g_last_error[0] = '\0';
  size_t need = 0;
  int rc = lbl_mlp_train_workspace_bytes(batch, &need);
  if (rc != LBL_OK) return rc;
  if (!params || !wbf16 || !x || !log10y || !target || !grad || !loss || !workspace ||
      (lr > 0.0f && (!adam_m || !adam_v || adam_t < 1)))
    return fail(LBL_EINVAL, "invalid arguments to lbl_mlp_train_step");
  if (n_data <= 0 || (n_data & (n_data - 1)) || (perm_a & 1u) == 0)
    return fail(LBL_EINVAL, "n_data must be a power of two and perm_a odd");
  if (workspace_bytes < need)
    return fail(LBL_EINVAL, "workspace of %zu bytes needed, %zu given", need, workspace_bytes);
  char* ws = static_cast<char*>(workspace);
  const size_t rows = (size_t)batch / 64;
  float* partial = reinterpret_cast<float*>(ws);
  float* loss_partial =
      reinterpret_cast<float*>(ws + align_up(rows * lbl::mlp_num_params() * sizeof(float)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(lbl::launch_mlp_step(params, adam_m, adam_v, wbf16, x, log10y, target, n_data, perm_a,
                               perm_b, base, batch, lr, adam_t, partial, loss_partial, grad, loss,
                               s));
  return LBL_OK;
}

int lbl_mlp_eval(int64_t n, const float* x, const float* log10y, const float* params,
                 const unsigned short* wbf16, float* t_out, void* stream) {
  g_last_error[0] = '\0';
  if (n < 0 || (n > 0 && (!x || !log10y || !params || !wbf16 || !t_out)))
    return fail(LBL_EINVAL, "invalid arguments to lbl_mlp_eval");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(lbl::launch_mlp_eval(n, x, log10y, params, wbf16, t_out, s));
  HIP_TRY(hipStreamSynchronize(s));
  return LBL_OK;
}

// ---------------------------------------------------------------- grid --
// nu_min + i * step rounded after the multiply and after the add, as numpy
// forms nu_min + np.arange(n) * step (forward.py:141-142).
static double affine_point(double nu_min, double step, int64_t i) {
#pragma clang fp contract(off)
  const double m = (double)i * step;
  return nu_min + m;
}

// Does nu_min + i * step reproduce every nu[i] bit for bit?  The last and
// middle points first (a wrong step fails there), then all of them.
static bool reproduces(const double* nu, int64_t n, double step) {
  if (!(step > 0.0) || !isfinite(step)) return false;
  const double a = nu[0];
  if (affine_point(a, step, n - 1) != nu[n - 1] || affine_point(a, step, n / 2) != nu[n / 2])
    return false;
  for (int64_t i = 1; i < n; ++i)
    if (affine_point(a, step, i) != nu[i]) return false;
  return true;
}

// Steps that may have built nu: the decimal roundings (1 to 17 significant
// digits) of the end-to-end step and of nu[1] - nu[0] (a caller's step is
// usually a short decimal such as 0.001 or 5e-4), then the end-to-end step
// and its neighbours up to 64 ulps away.
static bool find_step(const double* nu, int64_t n, double* step) {
  const double s_end = (nu[n - 1] - nu[0]) / (double)(n - 1);
  const double s_one = nu[1] - nu[0];
  const double bases[2] = {s_end, s_one};
  char buf[64];
  for (double b : bases) {
    if (!(b > 0.0) || !isfinite(b)) continue;
    for (int digits = 1; digits <= 17; ++digits) {
      snprintf(buf, sizeof(buf), "%.*g", digits, b);
      const double c = strtod(buf, nullptr);
      if (reproduces(nu, n, c)) { *step = c; return true; }
    }
  }
  if (!(s_end > 0.0) || !isfinite(s_end)) return false;
  double up = s_end, down = s_end;
  for (int k = 0; k <= 64; ++k) {
    if (reproduces(nu, n, up)) { *step = up; return true; }
    if (k && reproduces(nu, n, down)) { *step = down; return true; }
    up = nextafter(up, INFINITY);
    down = nextafter(down, 0.0);
  }
  return false;
}

int lbl_grid_from_nu(const double* nu, int64_t n_nu, double tile_width_cm1,
                     double line_cutoff_cm1, lbl_grid_t* grid) {
  g_last_error[0] = '\0';
  if (!nu || !grid || n_nu < 1) return fail(LBL_EINVAL, "empty grid");
  if (n_nu > kMaxPoints) return fail(LBL_EINVAL, "n_nu %lld exceeds 2^40 points", (long long)n_nu);
  if (!isfinite(tile_width_cm1) || !(tile_width_cm1 > 0.0) || !isfinite(line_cutoff_cm1))
    return fail(LBL_EINVAL, "invalid tile width or cutoff");
  for (int64_t i = 0; i < n_nu; ++i)
    if (!isfinite(nu[i])) return fail(LBL_EINVAL, "nu[%lld] is not finite", (long long)i);
  // forward.py:175-176: the tile size follows from the first two points
  const double ref_step = n_nu > 1 ? nu[1] - nu[0] : 0.001;
  const double ts = floor(tile_width_cm1 / ref_step + 1e-8);
  double step = 0.001;
  if (n_nu > 1 && !find_step(nu, n_nu, &step))
    return fail(LBL_EINVAL,
                "nu is not nu[0] + i*step for any step (the line sum needs "
                "the evenly spaced grid compute_transmittance and run_forward build)");
  grid->nu_min = nu[0];
  grid->step = step;
  grid->n_nu = n_nu;
  // (a step so small that the tile exceeds 2^62 points keeps 2^62: either way
  // the tile holds the whole grid, and the conversion stays in range)
  grid->tile_samples = !(ts >= 1.0) ? 1 : ts < 0x1p62 ? (int64_t)ts : ((int64_t)1 << 62);
  grid->cutoff = line_cutoff_cm1;
  grid->shift_bound = 0.0;
  return LBL_OK;
}

int lbl_voigt_eval(int64_t n, const void* x, const void* y, void* k_out, int precision,
                   void* stream) {
  g_last_error[0] = '\0';
  if (n < 0 || (n > 0 && (!x || !y || !k_out)))
    return fail(LBL_EINVAL, "invalid arguments to lbl_voigt_eval");
  if (precision != LBL_FP64 && precision != LBL_MIXED)
    return fail(LBL_EINVAL, "unknown precision %d", precision);
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(lbl::launch_voigt_eval(n, x, y, k_out, precision, s));
  HIP_TRY(hipStreamSynchronize(s));
  return LBL_OK;
}

}  // extern "C"

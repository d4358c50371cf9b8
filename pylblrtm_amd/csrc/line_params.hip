// Per-(layer, line) line-parameter scaling and the per-tile line ranges.
//
// k_line_params restates pylblrtm/forward.py:236-240 (partition-sum ratio)
// and :257-292 (pressure shift, temperature-scaled strength with the JRAD=0
// radiation term divided out at the shifted centre, Lorentz and Doppler
// half widths) for every (layer, line) pair in one pass, and folds the Voigt
// scale factors of forward.py:61-62 into the record it writes.
//
// Arithmetic is written operation by operation in the reference's evaluation
// order with contraction off, so the FP64 mode reproduces numpy's rounding
// up to the last-bit differences of exp/expm1/tanh/pow.
#include <math.h>

#include "faddeeva.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace lbl {
namespace {

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// numpy.interp for one query point: clamped ends, exact node hits return the
// node value, otherwise slope * (x - xp[j]) + fp[j] with the slope computed
// first (the order numpy's compiled interp uses).
__device__ double interp_like_numpy(double x, const double* xp, const double* fp, int n) {
  if (isnan(x)) return x;
  if (n == 1) return fp[0];
  if (x < xp[0]) return fp[0];
  if (x > xp[n - 1]) return fp[n - 1];
  if (x == xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;  // invariant xp[lo] <= x < xp[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  if (xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  double res = slope * (x - xp[lo]) + fp[lo];
  if (isnan(res)) {
    res = slope * (x - xp[lo + 1]) + fp[lo + 1];
    if (isnan(res) && fp[lo] == fp[lo + 1]) res = fp[lo];
  }
  return res;
}

__device__ __forceinline__ void store64(const Records& r, int64_t i, double c, double s, double y,
                                        double a) {
  r.c[i] = c;
  static_cast<double*>(r.s)[i] = s;
  static_cast<double*>(r.y)[i] = y;
  static_cast<double*>(r.a)[i] = a;
}

__device__ __forceinline__ void store_mixed(const Records& r, int64_t i, double c, float s, float y,
                                            float a, float step) {
  MixedRec m;
  const float y2 = y * y;
  m.c = c;
  m.s = s;
  m.hs = step * s;
  m.c1 = y2 - 1.5f;
  m.ay = a * y * (float)kInvSqrtPi;
  const float half_run = 0.5f * (float)(kPointsMixed - 1) * m.hs;
  const float x_s1 = sqrtf(fmaxf(s1_r2_min(y2) - y2, 0.0f));
  m.hw1 = half_run + x_s1;
  // the line sum picks the middle tier by the same test on c1
  const float r2_mid = (m.c1 < kF32S2Y2 - 1.5f) ? s2_r2_min32(y2) : kF32Gh2R2;
  const float x_s2 = sqrtf(fmaxf(r2_mid - y2, 0.0f));
  m.hw2 = half_run + x_s2;
  m.y = y;
  m.a = a;
  const float H = wave_half_span32(step);
  m.ds1 = H + x_s1 / s;
  m.dm = x_s2 / s;
  r.mix[i] = m;
}

// R(nu, T) = nu * tanh(c2 nu / 2T)  (forward.py:41-48)
__device__ __forceinline__ double radiation64(double nu, double T) {
  return nu * tanh(0.5 * kC2 * nu / T);
}
__device__ __forceinline__ float radiation32(float nu, float T) {
  return nu * tanhf(0.5f * (float)kC2 * nu / T);
}

// One thread per (layer, line).  Grid: one workgroup per (256-line chunk,
// layer), chunk-major and XCD-aware, so the layers of a chunk run on one XCD
// at about the same time and read the chunk's line parameters from its L2
// instead of streaming the whole line list once per layer; a 2-D grid (x over
// chunks, y over layers) where the 1-D one would exceed its size limit.
// R = double: everything in FP64 (LBL_FP64).
// R = float : the centre and the Boltzmann exponent in FP64, the rest FP32.
template <typename R>
__global__ __launch_bounds__(256) void k_line_params(const LineParamsArgs a) {
  __shared__ double s_q[kMaxGroups];
  int64_t chunk = blockIdx.x;
  int k = blockIdx.y;
  if (gridDim.y == 1 && a.layers.n_layers > 1) {
    const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
    chunk = w / a.layers.n_layers;
    k = (int)(w - chunk * a.layers.n_layers);
  }
  const double T = a.layers.t_k[k];
  const double P = a.layers.p_atm[k];
  // Q(T_ref) / Q(T_k) per group (forward.py:236-240), once per workgroup.
  // (an empty table, or a group whose gas row does not exist, is flagged and
  // never read: the host then returns LBL_EINVAL)
  for (int g = threadIdx.x; g < a.groups.n_groups; g += blockDim.x) {
    const int o = a.groups.tips_offset[g];
    const int n = a.groups.tips_offset[g + 1] - o;
    const int gas = a.groups.gas[g];
    if (gas < 0 || gas >= a.layers.n_gases) atomicMin(a.err, kErrBadGas);
    if (o < 0 || n < 1) {
      atomicMin(a.err, kErrBadTips);
      s_q[g] = __builtin_nan("");
      continue;
    }
    const double* tt = a.groups.tips_t + o;
    const double* qq = a.groups.tips_q + o;
    s_q[g] = interp_like_numpy(kTRef, tt, qq, n) / interp_like_numpy(T, tt, qq, n);
  }
  __syncthreads();

  const int64_t l = chunk * blockDim.x + threadIdx.x;
  if (l >= a.lines.n_lines) return;
  const int64_t out = (int64_t)k * a.lines.n_lines + l;
  const int g = a.lines.group[l];
  const int gas = (g >= 0 && g < a.groups.n_groups) ? a.groups.gas[g] : -1;
  const bool valid = gas >= 0 && gas < a.layers.n_gases;
  if (!valid) atomicMin(a.err, (g >= 0 && g < a.groups.n_groups) ? kErrBadGas : kErrBadGroup);
  const double vmr = valid ? a.layers.vmr[(int64_t)gas * a.layers.n_layers + k] : 0.0;
  const double col = valid ? a.layers.col[(int64_t)gas * a.layers.n_layers + k] : 0.0;
  if (col == 0.0) {
    // forward.py:253-255: the group is skipped for this layer.  A NaN centre
    // fails every tile's inclusion test, so the record is never summed.
    if constexpr (sizeof(R) == 8)
      store64(a.rec, out, __builtin_nan(""), 0.0, 1.0, 0.0);
    else
      store_mixed(a.rec, out, __builtin_nan(""), 0.0f, 1.0f, 0.0f, 0.0f);
    return;
  }
  const double nu0 = a.lines.nu0[l];
  const double p_self = vmr * P;
  const double p_air = P - p_self;
  const double c = nu0 + a.lines.shift[l] * p_air;  // shifted centre (forward.py:263)
  // Boltzmann factor exp(-c2 E'' (1/T - 1/T_ref)), FP64 in every mode.
  const double boltz = exp(-kC2 * a.lines.elow[l] * (1.0 / T - 1.0 / kTRef));
  const double molar_mass = a.groups.molar_mass[g];

  if constexpr (sizeof(R) == 8) {
    const double strength = a.lines.sref[l] * s_q[g] * boltz * (-expm1(-kC2 * nu0 / T)) /
                            (-expm1(-kC2 * nu0 / kTRef)) * col / radiation64(c, T);
    if (!(isfinite(strength) && strength >= 0.0)) {
      atomicMin(a.err, err_numeric(k, 0, g));
    }
    const double lorentz =
        pow(kTRef / T, a.lines.n_air[l]) * (a.lines.g_air[l] * p_air + a.lines.g_self[l] * p_self);
    const double doppler = nu0 * kDoppler * sqrt(T / molar_mass);
    store64(a.rec, out, c, kSqrtLn2 / doppler, kSqrtLn2 * lorentz / doppler,
            strength * kVoigtK / doppler);
  } else {
    const float Tf = (float)T;
    const float nu0f = (float)nu0;
    const float c2 = (float)kC2;
    const float strength = (float)a.lines.sref[l] * (float)s_q[g] * (float)boltz *
                           (-expm1f(-c2 * nu0f / Tf)) / (-expm1f(-c2 * nu0f / (float)kTRef)) *
                           (float)col / radiation32((float)c, Tf);
    if (!(isfinite(strength) && strength >= 0.0f)) {
      atomicMin(a.err, err_numeric(k, 0, g));
    }
    const float lorentz = powf((float)kTRef / Tf, (float)a.lines.n_air[l]) *
                          ((float)a.lines.g_air[l] * (float)p_air +
                           (float)a.lines.g_self[l] * (float)p_self);
    const float doppler = nu0f * (float)kDoppler * sqrtf(Tf / (float)molar_mass);
    store_mixed(a.rec, out, c, (float)kSqrtLn2 / doppler, (float)kSqrtLn2 * lorentz / doppler,
                strength * (float)kVoigtK / doppler, (float)a.step);
  }
}

// Line index range [lo, hi) of the nu0-sorted list that can fall inside tile
// t's window [nu[t0] - cutoff, nu[t1-1] + cutoff] (forward.py:296-301) once
// shifted by at most `margin`.  The exact per-record test happens later.
__global__ void k_tile_ranges(const TileRangeArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_tiles) return;
  const int64_t t0 = t * a.ts;
  const int64_t t1 = min(t0 + a.ts, a.n_nu);
  const double lo_v = (grid_nu(a.nu_min, a.step, t0) - a.cutoff) - a.margin;
  const double hi_v = (grid_nu(a.nu_min, a.step, t1 - 1) + a.cutoff) + a.margin;
  // lower_bound(lo_v)
  int64_t lo = 0, hi = a.n_lines;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a.nu0[mid] < lo_v) lo = mid + 1; else hi = mid;
  }
  const int64_t first = lo;
  // upper_bound(hi_v)
  hi = a.n_lines;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a.nu0[mid] <= hi_v) lo = mid + 1; else hi = mid;
  }
  a.lo[t] = (int32_t)first;
  a.hi[t] = (int32_t)lo;
}

}  // namespace

// Near ranges of every wavefront, slice and 64-point group: the records
// whose nu0 lies in [nu_m - half width, nu_m + half width] (lower and upper
// bound in the sorted nu0), nu_m the unit's middle point as the line sums
// form it.  One thread per (wavefront, range): ranges 0..kGroupsPerWave-1
// are the wavefront's groups', the last its own (and, for its slice's first
// wavefront, the slice's); each thread's binary searches are independent,
// so their dependent loads overlap across threads.
__global__ void k_wave_ranges(const WaveRangeArgs a) {
  constexpr int kParts = kGroupsPerWave + 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_wg * a.waves_per_slice * kParts) return;
  const int64_t g = t / kParts;
  const int part = (int)(t - g * kParts);
  const int64_t wg = g / a.waves_per_slice;
  const int w = (int)(g - wg * a.waves_per_slice);
  const int64_t tile = wg / a.slices_per_tile;
  const int64_t slice = wg - tile * a.slices_per_tile;
  const int64_t first =
      tile * a.ts + (slice * a.waves_per_slice + w) * (int64_t)kWavePoints;
  // [lower_bound(lo_v), upper_bound(hi_v)) of the nu0-sorted list
  auto range = [&](double lo_v, double hi_v, int32_t* out_lo, int32_t* out_hi) {
    int64_t lo = 0, hi = a.n_lines;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (a.nu0[mid] < lo_v) lo = mid + 1; else hi = mid;
    }
    *out_lo = (int32_t)lo;
    hi = a.n_lines;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (a.nu0[mid] <= hi_v) lo = mid + 1; else hi = mid;
    }
    *out_hi = (int32_t)lo;
  };
  if (part < kGroupsPerWave) {   // a group of the wavefront
    const double nu_g = group_middle(a.nu_min, a.step, first + (int64_t)part * kGroupPoints);
    range(nu_g - a.group_half_width, nu_g + a.group_half_width,
          a.glo + g * kGroupsPerWave + part, a.ghi + g * kGroupsPerWave + part);
    return;
  }
  const double nu_w = wave_middle(a.nu_min, a.step, first);
  range(nu_w - a.half_width, nu_w + a.half_width, a.wlo + g, a.whi + g);
  if (w == 0) {   // the slice's near range
    const double nu_s = slice_middle(a.nu_min, a.step, tile * a.ts + slice * (int64_t)kSlicePoints);
    range(nu_s - a.slice_half_width, nu_s + a.slice_half_width, a.slo + wg, a.shi + wg);
  }
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_line_params(const LineParamsArgs& a, int precision, hipStream_t s) {
  if (a.lines.n_lines <= 0 || a.layers.n_layers <= 0) return hipSuccess;
  const int64_t chunks = (a.lines.n_lines + 255) / 256;
  const int64_t blocks = chunks * a.layers.n_layers;
  // 1-D while the grid's thread count fits in 32 bits
  const dim3 grid = blocks < (int64_t(1) << 24) ? dim3((unsigned)blocks, 1)
                                                 : dim3((unsigned)chunks, (unsigned)a.layers.n_layers);
  if (precision == LBL_FP64)
    k_line_params<double><<<grid, 256, 0, s>>>(a);
  else
    k_line_params<float><<<grid, 256, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_tile_ranges(const TileRangeArgs& a, hipStream_t s) {
  if (a.n_tiles <= 0) return hipSuccess;
  k_tile_ranges<<<(unsigned)((a.n_tiles + 255) / 256), 256, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_wave_ranges(const WaveRangeArgs& a, hipStream_t s) {
  const int64_t n = a.n_wg * a.waves_per_slice * (kGroupsPerWave + 1);
  if (n <= 0) return hipSuccess;
  k_wave_ranges<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace lbl

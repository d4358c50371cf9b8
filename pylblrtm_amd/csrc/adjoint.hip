// Adjoint (vector-Jacobian product) of the transmittance with respect to the
// layer temperatures T_k and the layer mixing ratios v_{k,g}, with the layer
// column of gas g taken as v_{k,g} * air_k (so d col / d v = air_k).
//
// Forward model (pylblrtm/forward.py:257-312, :151):
//   T_m(nu) = exp(-tau(nu)/mu_m),  tau = sum_k R(nu, T_k) sum_l A_kl K(x_kl, y_kl),
//   A = S sqrt(ln2/pi)/alpha_D,  x = s (nu - c),  y = sqrt(ln2) gamma_L / alpha_D.
// Given G_m(nu) = dL/dT_m(nu), with W(nu) = -sum_m G_m T_m / mu_m,
//   dL/dtheta = sum_nu W(nu) d tau(nu)/d theta.
// Per record and parameter theta the chain rule needs d ln A/d theta,
// dx/d theta = bx + ax * x and dy/d theta; k_line_grad computes them in FP64:
//   temperature: d ln A/dT = -Q'(T)/Q(T) + c2 E''/T^2 - (a/T^2)/expm1(a/T)
//                           + (c2 c/T^2)/sinh(c2 c/T) - 1/(2T),   a = c2 nu0
//                ax = -1/(2T), bx = 0, dy/dT = -y (n_air + 1/2)/T
//   mixing ratio (c = nu0 + delta P (1 - v)):
//                d ln A/dv = 1/v + delta P (1/c + (c2/T)/sinh(c2 c/T))
//                ax = 0, bx = s delta P, dy/dv = sqrt(ln2) (296/T)^n P (g_self - g_air)/alpha_D
// and, at every grid point, dR/dT = -R (c2 nu/T^2)/sinh(c2 nu/T).
// The backward sum recomputes K and its derivatives (no lines x grid storage),
// with the forward's region split:
//   far, shifted one-node form (faddeeva.h, D1 = x^2 + y^2 - 3/2):
//     K = y/(sqrt(pi) D1),  dK/dx = -2xy/(sqrt(pi) D1^2),  dK/dy = (D1 - 2y^2)/(sqrt(pi) D1^2),
//   far, two-node form (|z| >= 25, w = x^2 + y^2 - 1/2, D = w^2 + 2y^2):
//     K = y (w+1)/(sqrt(pi) D),  dK/dx = 2xy (2y^2 - w^2 - 2w)/(sqrt(pi) D^2),
//     dK/dy = ((w + 1 + 2y^2) D - 4 y^2 (w+1)^2)/(sqrt(pi) D^2),
//   both accumulated per record as point sums weighted once per record (VjpFar);
//   near: K, dK/dx, dK/dy from kgrad_mixed (each region's approximant differentiated).
#include <math.h>

#include "faddeeva.h"
#include "internal.h"

namespace lbl {
namespace {

constexpr int kVjpPoints = 8;
constexpr int kVjpChunk = kBlock;
constexpr int kVjpFold = 32;
// Far-field series of the adjoint: terms of the one-node form's r = 1/(x^2 + c1)
// and r^2 (whose double pole converges more slowly): outside the near range
// 22 terms leave <= 1.1e-9 of r and 1.8e-7 of r^2 (and of x r^2, formed as x
// times the truncated r^2 series).  A record farther out needs fewer for the
// same 1.8e-7 (adj_terms).
constexpr int kAdjTerms = 22;
constexpr int kAdjSliceExCap = 1024;
constexpr int kAdjQueueCap = 256;   // near-centre (record, lane) pairs per wavefront per gas run
constexpr int kAdjCoreCap = 512;    // core points of queued pairs listed per wavefront (>= 64 x 8)
// The drain's first pass: in FP32 the six-node form's K and gradient are
// within 1e-6 of K and of |grad K| from |z| = 5.5 (from 7 for y < 1e-4,
// where the Gaussian exp(-x^2) it leaves out is not yet negligible against
// y/(sqrt(pi) x^2)), better than the four-node form from |z| = 8, where
// kgrad_mixed switches to it; inside, the core.
__device__ __forceinline__ float adj_core_r2(float y) { return y < 1e-4f ? 49.0f : 30.25f; }

struct GradArgs {
  lbl_lines_t lines;
  lbl_groups_t groups;
  lbl_layers_t layers;
  float *at, *dyt, *av, *bxv, *dyv;   // [n_layers * n_lines]
};

// Slope of the numpy.interp segment used at x (0 outside the table).
__device__ double interp_slope(double x, const double* xp, const double* fp, int n) {
  if (n < 2 || !(x >= xp[0]) || !(x <= xp[n - 1])) return 0.0;
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  return (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
}

__device__ double interp_value(double x, const double* xp, const double* fp, int n) {
  if (n == 1 || x <= xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  return fp[lo] + (x - xp[lo]) * (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
}

__global__ __launch_bounds__(256) void k_line_grad(const GradArgs a) {
  __shared__ double s_dlnq[kMaxGroups];
  const int k = blockIdx.y;
  const double T = a.layers.t_k[k];
  const double P = a.layers.p_atm[k];
  for (int g = threadIdx.x; g < a.groups.n_groups; g += blockDim.x) {
    const int o = a.groups.tips_offset[g];
    const int n = a.groups.tips_offset[g + 1] - o;
    const double* tt = a.groups.tips_t + o;
    const double* qq = a.groups.tips_q + o;
    s_dlnq[g] = -interp_slope(T, tt, qq, n) / interp_value(T, tt, qq, n);
  }
  __syncthreads();
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= a.lines.n_lines) return;
  const int64_t out = (int64_t)k * a.lines.n_lines + l;
  const int g = a.lines.group[l];
  const int gas = a.groups.gas[g];
  const double v = a.layers.vmr[(int64_t)gas * a.layers.n_layers + k];
  const double col = a.layers.col[(int64_t)gas * a.layers.n_layers + k];
  if (col == 0.0 || v == 0.0) {
    a.at[out] = a.dyt[out] = a.av[out] = a.bxv[out] = a.dyv[out] = 0.0f;
    return;
  }
  const double nu0 = a.lines.nu0[l];
  const double delta = a.lines.shift[l];
  const double n_air = a.lines.n_air[l];
  const double p_self = v * P, p_air = P - p_self;
  const double c = nu0 + delta * p_air;
  const double doppler = nu0 * kDoppler * sqrt(T / a.groups.molar_mass[g]);
  const double gscale = pow(kTRef / T, n_air);
  const double lorentz = gscale * (a.lines.g_air[l] * p_air + a.lines.g_self[l] * p_self);
  const double y = kSqrtLn2 * lorentz / doppler;
  const double s = kSqrtLn2 / doppler;
  const double aa = kC2 * nu0 / T;
  const double u = kC2 * c / T;
  const double inv_t = 1.0 / T;
  a.at[out] = (float)(s_dlnq[g] + kC2 * a.lines.elow[l] * inv_t * inv_t -
                      (aa * inv_t) / expm1(aa) + (u * inv_t) / sinh(u) - 0.5 * inv_t);
  a.dyt[out] = (float)(-y * (n_air + 0.5) * inv_t);
  a.av[out] = (float)(1.0 / v + delta * P * (1.0 / c + (kC2 * inv_t) / sinh(u)));
  a.bxv[out] = (float)(s * delta * P);
  a.dyv[out] = (float)(kSqrtLn2 * gscale * P * (a.lines.g_self[l] - a.lines.g_air[l]) / doppler);
}

struct VjpArgs {
  Records rec;                 // mixed-precision records of the forward
  const float *at, *dyt, *av, *bxv, *dyv;
  const int32_t* group;        // per line
  const int32_t* gas_of_group;
  const int32_t* lo;
  const int32_t* hi;
  const int32_t* wlo;          // near range of each (slice, wavefront), k_wave_ranges
  const int32_t* whi;
  const int32_t* slo;          // near range of each slice
  const int32_t* shi;
  const double* t_k;
  int64_t n_lines;
  int32_t n_layers, n_gases;
  double nu_min, step, cutoff;
  int64_t n_nu, ts;
  int32_t slices_per_tile;
  int64_t n_wg;
  int32_t n_mu;
  double inv_mu[kMaxMu];
  const float* trans;          // [n_mu][n_nu] forward transmittance
  const float* grad_trans;     // [n_mu][n_nu] upstream gradient
  const float* od_layers;      // [n_layers][n_nu] forward per-layer optical depth
  double* grad_t;              // [n_layers]
  double* grad_vmr;            // [n_gases][n_layers]
};

// Far-field record of one (layer, line) pair as staged in LDS.  Two-node form:
// K = ay F with F = (w+1)/D, and per grid point
//   dK/dx = 2 ay x F',  dK/dy = aq (F + d2 F' - 2 d2 F/D),  F' = (d2 - w^2 - 2w)/D^2,
// so a record's contribution to both gradients is a fixed combination of five
// running sums over its points: S1 = sum W F, S2 = sum W F', S3 = sum W x F',
// S4 = sum W x^2 F', S5 = sum W F/D (W = the weighted upstream gradient).
// The coefficients below are those combinations' weights; dx/dT = -x/(2T)
// multiplies S4, dx/dv = bx multiplies S3.  Shifted one-node form (r = 1/D1):
// K = ay r, dK/dx = -2 ay x r^2, dK/dy = aq (r - d2 r^2), so with S1 = sum W r,
// S2 = sum W r^2, S3 = sum W x r^2, S4 = sum W x^2 r^2 = S1 - c1 S2 the same
// coefficients apply as tF S1 - tX2 S4 - tD S2 and vF S1 - vX S3 - vD S2.
struct alignas(16) VjpFar {
  double c;
  float s, hs;
  float c1, hw1, hw2, tA;    // y^2 - 3/2, mid-run |x| thresholds (one-node, two-node), tF - tX2
  float tF, tX2, tD, tQ;     // temperature: weights of S1, S4, S2, S5
  float vF, vX, vD, vQ;      // mixing ratio: weights of S1, S3, S2, S5
};

// What the near-centre branch (complex Faddeeva, per point) needs.
struct alignas(16) VjpNear {
  float y, a, at, dyt;
  float av, bxv, dyv;
  int32_t gas;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// One workgroup per (tile slice, layer): P consecutive points per thread.
//
// Far field.  As in the forward's series (k_far_series_mixed), a record far
// from a slice or wavefront middle point contributes smooth functions of the
// unit's normalised offset tau; here they are the one-node form's
// r = 1/(x^2 + c1), r^2 and x r^2, x = xm + sh tau, whose sums against WR over
// the unit's points are the point sums S1, S2, S3 of the one-node branch
// below.  With the unit's moments M_n = sum_j WR_j tau_j^n (computed once per
// workgroup),
//   S1 = sum_n a_n M_n,  S2 = sum_n b_n M_n,  S3 = xm S2 + sh sum_n b_n M_(n+1),
// where a_n are r's Taylor coefficients (a_0 = 1/D0, a_1 = al a_0,
// a_n = al a_(n-1) + be a_(n-2); D0 = xm^2 + c1, al = -2 xm sh / D0,
// be = -sh^2 / D0) and b_n = -d a_n / d c1 those of r^2:
// b_n = a_0 a_n + al b_(n-1) + be b_(n-2).  The split of records is the
// forward's: slice level for records outside the slice's near range where
// the one-node form holds at every point of the slice, wavefront level for
// the rest outside the wavefront's near range where it holds at every point
// of the wavefront, point by point for everything else.
//
// Near field.  The records of the slice's near range (and the slice level's
// exceptions) are staged through LDS in chunks, each chunk counting-sorted
// by gas (ballot ranks, so the order is deterministic) so the per-gas
// gradient is summed in FP32 over runs of one gas and folded into FP64
// every kVjpFold records; each wavefront skips the records its series took.
template <int P>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_vjp_mixed(const VjpArgs a) {
  static_assert(P == kPointsMixed && kBlock * P == kSlicePoints && 64 * P == kWavePoints,
                "slice and wavefront geometry of the forward's series");
  constexpr int kWaves = kBlock / 64;
  constexpr int kNT = kAdjTerms;
  __shared__ VjpFar s_far[kVjpChunk];
  __shared__ VjpNear s_near[kVjpChunk];
  __shared__ int32_t s_wcnt[kWaves][8];
  __shared__ int32_t s_idx[kVjpChunk];       // line index of each staged record
  __shared__ float s_ds1[kVjpChunk];         // its one-node distance ds1
  __shared__ float s_ms[kWaves][kNT + 1];    // slice moments, per wavefront's share
  __shared__ float s_mw[kWaves][kNT + 1];    // wavefront moments
  __shared__ float s_msum[kNT + 1];          // slice moments
  __shared__ int32_t s_sx[kAdjSliceExCap];   // slice-level exceptions
  __shared__ int s_sxn;
  __shared__ uint16_t s_aq[kWaves][kAdjQueueCap];   // near-centre queue: (staged record, lane)
  __shared__ uint16_t s_acl[kWaves][kAdjCoreCap];   // its core points: (queue entry, point)
  __shared__ float s_at[kBlock], s_av[kBlock];       // queue results per owning thread
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t lw = xcd_remap(blockIdx.x, a.n_wg * a.n_layers);
  const int k = (int)(lw / a.n_wg);
  const int64_t wg = lw - (int64_t)k * a.n_wg;
  const int64_t tile = wg / a.slices_per_tile;
  const int slice = (int)(wg - tile * a.slices_per_tile);
  const int64_t t0 = tile * a.ts;
  const int64_t t1 = min(t0 + a.ts, a.n_nu);
  // a slice wholly past the end of a short (last) tile has no points: the
  // whole workgroup leaves before its first barrier
  if (t0 + (int64_t)slice * (kBlock * P) >= t1) return;
  const int64_t i0 = t0 + (int64_t)slice * (kBlock * P) + (int64_t)tid * P;
  const double lo_b = __dsub_rn(grid_nu(a.nu_min, a.step, t0), a.cutoff);
  const double hi_b = __dadd_rn(grid_nu(a.nu_min, a.step, t1 - 1), a.cutoff);
  const int32_t rlo = a.lo[tile], rhi = a.hi[tile];
  const double nu_m = __dadd_rn(grid_nu(a.nu_min, a.step, i0), 0.5 * (P - 1) * a.step);
  const float stepf = (float)a.step;
  const MixedRec* __restrict__ recs = a.rec.mix;
  const int n_gas = a.n_gases;
  const float T = (float)a.t_k[k];
  const float ax_t = -0.5f / T;

  // W(nu) = -sum_m G_m T_m / mu_m and WR = W R(nu, T_k) at this thread's
  // points; the radiation term's own derivative W (d ln R/dT) od_k(nu) needs
  // only the forward's layer optical depth.
  float WR[P];
  double g_t;
  {
    float rad = 0.0f;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int64_t i = i0 + j;
      float w = 0.0f, od = 0.0f;
      if (i < t1) {
        for (int m = 0; m < a.n_mu; ++m)
          w -= a.grad_trans[(int64_t)m * a.n_nu + i] * a.trans[(int64_t)m * a.n_nu + i] *
               (float)a.inv_mu[m];
        od = a.od_layers[(int64_t)k * a.n_nu + i];
      }
      const float nu = (float)grid_nu(a.nu_min, a.step, i);
      const float u = (float)kC2 * nu / T;
      WR[j] = w * nu * tanhf(0.5f * u);
      rad = fmaf(w * od, -(u / T) / sinhf(u), rad);
    }
    g_t = (double)rad;
  }
  // (j - (P-1)/2) WR_j: sum_j WR_j x_j r_j^2 = xm S2 + hs sum_j oW_j r_j^2
  float oW[P];
#pragma unroll
  for (int j = 0; j < P; ++j) oW[j] = ((float)j - 0.5f * (float)(P - 1)) * WR[j];
  double g_v[8];
#pragma unroll
  for (int g = 0; g < 8; ++g) g_v[g] = 0.0;

  const int64_t kb = (int64_t)k * a.n_lines;
  // ---- far field: moments, then the slice and wavefront series
  const int64_t s0 = t0 + (int64_t)slice * kSlicePoints;
  const int64_t w_first = s0 + (int64_t)wv * kWavePoints;
  const double nu_s = slice_middle(a.nu_min, a.step, s0);
  const double nu_w = wave_middle(a.nu_min, a.step, w_first);
  const float dS = 0.5f * (float)(kSlicePoints - kWavePoints) * stepf;   // Hs - H
  const int32_t slo = min(max(a.slo[wg], rlo), rhi);
  const int32_t shi = max(min(a.shi[wg], rhi), slo);
  const int32_t nlo = min(max(a.wlo[wg * kWaves + wv], slo), shi);
  const int32_t nhi = max(min(a.whi[wg * kWaves + wv], shi), nlo);
  {
    // moments of WR about the slice's and this wavefront's middle (WR = 0
    // past the tile end), one set at a time to keep registers free
#pragma unroll
    for (int lvl = 0; lvl < 2; ++lvl) {
      float mo[kNT + 1];
#pragma unroll
      for (int n = 0; n <= kNT; ++n) mo[n] = 0.0f;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const float t = lvl == 0
            ? ((float)(wv * kWavePoints + lane * P + j) - 0.5f * (float)(kSlicePoints - 1)) *
                  (2.0f / (float)kSlicePoints)
            : ((float)(lane * P + j) - 0.5f * (float)(kWavePoints - 1)) * (2.0f / (float)kWavePoints);
        float pw = WR[j];
#pragma unroll
        for (int n = 0; n <= kNT; ++n) {
          mo[n] += pw;
          pw *= t;
        }
      }
#pragma unroll
      for (int n = 0; n <= kNT; ++n) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) mo[n] += __shfl_xor(mo[n], off);
      }
      if (lane == 0) {
#pragma unroll
        for (int n = 0; n <= kNT; ++n) (lvl == 0 ? s_ms : s_mw)[wv][n] = mo[n];
      }
    }
    if (tid == 0) s_sxn = 0;
  }
  __syncthreads();
  if (tid <= kNT) {
    float t = 0.0f;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += s_ms[w][tid];
    s_msum[tid] = t;
  }
  __syncthreads();
  // S1 = sum W r, S2 = sum W r^2, S3 = sum W x r^2 of one record from the
  // moments M (read from LDS as they are needed: registers are scarce here)
  auto series = [&](const float* M, float xm, float sh, float c1, float& S1, float& S2,
                    float& S3, const int nt) {
    const float a0 = __builtin_amdgcn_rcpf(fmaf(xm, xm, c1));
    const float u = sh * a0;
    const float al = -2.0f * xm * u, be = -sh * u;
    float ap = a0, an = al * a0;
    float bp = a0 * a0, bn = 2.0f * al * bp;
    float s1 = fmaf(an, M[1], a0 * M[0]);
    float s2 = fmaf(bn, M[1], bp * M[0]);
    float s3 = fmaf(bn, M[2], bp * M[1]);
#pragma unroll
    for (int n = 2; n < kNT; ++n) {
      if ((n == 6 || n == 8 || n == 11 || n == 15 || n == 19) && n >= nt) break;
      const float a_n = fmaf(al, an, be * ap);
      const float b_n = fmaf(a0, a_n, fmaf(al, bn, be * bp));
      s1 = fmaf(a_n, M[n], s1);
      s2 = fmaf(b_n, M[n], s2);
      s3 = fmaf(b_n, M[n + 1], s3);
      ap = an; an = a_n;
      bp = bn; bn = b_n;
    }
    S1 = s1;
    S2 = s2;
    S3 = fmaf(xm, s2, sh * s3);
  };
  // adds record li's series (middle point at offset d, half-span sh/s) to the gradients
  auto far_record = [&](const float* M, const MixedRec& m, int64_t li, float d, float sh,
                        const int nt) {
    const int64_t r = kb + li;
    float S1, S2, S3;
    series(M, d * m.s, sh, m.c1, S1, S2, S3, nt);
    const float y = m.y, am = m.a;
    const float aq = am * (float)kInvSqrtPi, ay = m.ay;
    const float cty = a.dyt[r] * aq, cvy = a.dyv[r] * aq;
    const float d2 = 2.0f * y * y;
    const float tX2 = 2.0f * ax_t * ay;
    const float tA = fmaf(a.at[r], ay, cty) - tX2;
    const float ft = fmaf(tA, S1, fmaf(tX2, m.c1, -cty * d2) * S2);
    const float vF = fmaf(a.av[r], ay, cvy);
    const float fv = fmaf(vF, S1, fmaf(-2.0f * a.bxv[r] * ay, S3, -cvy * d2 * S2));
    g_t += (double)ft;
    int gas = a.gas_of_group[a.group[li]];
#pragma unroll
    for (int gg = 0; gg < 8; ++gg)
      if (gg == gas) g_v[gg] += (double)fv;
  };
  // the term count of a round (wavefront-uniform): the fewest terms that keep
  // r, r^2 and x r^2 within 1.8e-7 for every record the round sums, each at
  // R = |nu_m - c| / Hu unit half-spans (>= 3): 22 terms from R = 3, 19 from
  // 3.5, 15 from 5, 11 from 8, 8 from 14, 6 from 30
  // (tests/test_far_series.py)
  auto adj_terms = [&](bool use, float R) -> int {
    if (__ballot(use && R < 3.5f)) return kNT;
    if (__ballot(use && R < 5.0f)) return 19;
    if (__ballot(use && R < 8.0f)) return 15;
    if (__ballot(use && R < 14.0f)) return 11;
    if (__ballot(use && R < 30.0f)) return 8;
    return 6;
  };
  static_assert(kNT == 22, "adj_terms's table ends at 22 terms");
  // the one-node form holds at every point of the slice / of this wavefront
  const float Hw = wave_half_span32(a.step);   // the series' unit sh = s Hw
  const float inv_hs = 0.25f / Hw;             // 1 / Hs
  auto slice_ok = [&](const MixedRec& m) { return fabsf((float)(nu_s - m.c)) >= m.ds1 + dS; };
  auto wave_ok = [&](const MixedRec& m) { return fabsf((float)(nu_w - m.c)) >= m.ds1; };
  const MixedRec* __restrict__ lrec = recs + kb;
  {
    // slice level: records of the tile outside [slo, shi), one per thread
    const float* const M = s_msum;
    const int32_t n_below = slo - rlo, n_sl = n_below + (rhi - shi);
    auto sl_index = [&](int32_t v) -> int64_t {
      return v < n_below ? (int64_t)rlo + v : (int64_t)shi + (v - n_below);
    };
    // the next record is loaded one iteration ahead
    MixedRec nx;
    if (tid < n_sl) nx = lrec[sl_index(tid)];
    for (int32_t v = tid; v < n_sl; v += kBlock) {
      const int64_t li = sl_index(v);
      const MixedRec m = nx;
      if (v + kBlock < n_sl) nx = lrec[sl_index(v + kBlock)];
      // (outside the tile window, or skipped: nothing)
      const bool inw = m.c >= lo_b && m.c <= hi_b;
      const float d = (float)(nu_s - m.c);
      const bool ser = inw && slice_ok(m);
      const int nt = adj_terms(ser, fabsf(d) * inv_hs);
      if (ser) {
        far_record(M, m, li, d, 4.0f * (m.s * Hw), nt);
      } else if (inw) {
        const int pos = atomicAdd(&s_sxn, 1);
        if (pos < kAdjSliceExCap) s_sx[pos] = (int32_t)li;
      }
    }
  }
  __syncthreads();   // s_sx, s_sxn complete
  const int n_sx = s_sxn;
  const bool rescan = n_sx > kAdjSliceExCap;
  {
    // wavefront level: [slo, nlo), [nhi, shi) and the slice level's
    // exceptions (all of the tile outside [slo, shi) that failed there, if
    // the list overflowed), one per lane
    const float* const M = s_mw[wv];
    const int32_t nA = nlo - slo, nB = shi - nhi;
    const int32_t nC = rescan ? (slo - rlo) + (rhi - shi) : n_sx;
    for (int32_t v = lane; v < nA + nB + nC; v += 64) {
      int64_t li;
      if (v < nA) li = (int64_t)slo + v;
      else if (v < nA + nB) li = (int64_t)nhi + (v - nA);
      else if (!rescan) li = s_sx[v - nA - nB];
      else {
        const int32_t u = v - nA - nB;
        li = u < slo - rlo ? (int64_t)rlo + u : (int64_t)shi + (u - (slo - rlo));
      }
      const MixedRec m = lrec[li];
      if (!(m.c >= lo_b && m.c <= hi_b)) continue;
      if (rescan && v >= nA + nB && slice_ok(m)) continue;   // the slice level summed it
      if (wave_ok(m)) far_record(M, m, li, (float)(nu_w - m.c), m.s * Hw, kNT);
    }
  }

  s_at[tid] = 0.0f;
  s_av[tid] = 0.0f;
  const bool live = i0 < t1;   // threads wholly past the tile end have WR = 0
  // ---- near field: the slice's near range and the slice level's exceptions
  // (the whole tile if their list overflowed), point by point, each record
  // skipped by the wavefronts whose series took it
  const int32_t n_stage = rescan ? rhi - rlo : (shi - slo) + n_sx;
  auto stage_index = [&](int32_t v) -> int64_t {
    if (rescan) return (int64_t)rlo + v;
    return v < shi - slo ? (int64_t)slo + v : (int64_t)s_sx[v - (shi - slo)];
  };
  // the near-centre sums of staged record qi at the points of lane o of this
  // wavefront (x exactly as that lane forms it), given that lane's WR
  auto near_pair = [&](int qi, int o, const float (&wr_o)[P], float& acc_t, float& acc_v) {
    const VjpFar& R = s_far[qi];
    const VjpNear& N = s_near[qi];
    const int64_t i0o = t0 + (int64_t)slice * (kBlock * P) + (int64_t)(wv * 64 + o) * P;
    const double nu_mo = __dadd_rn(grid_nu(a.nu_min, a.step, i0o), 0.5 * (P - 1) * a.step);
    const float xm = (float)(nu_mo - R.c) * R.s;
    acc_t = 0.0f;
    acc_v = 0.0f;
#pragma unroll 1
    for (int j = 0; j < P; ++j) {
      const float x = fmaf((float)j - 0.5f * (float)(P - 1), R.hs, xm);
      float K, Kx, Ky;
      kgrad_mixed(x, N.y, K, Kx, Ky);
      const float AK = N.a * K;
      Kx *= N.a;
      Ky *= N.a;
      float wr = 0.0f;
#pragma unroll
      for (int jj = 0; jj < P; ++jj) wr = (jj == j) ? wr_o[jj] : wr;
      acc_t = fmaf(wr, N.at * AK + Kx * (ax_t * x) + Ky * N.dyt, acc_t);
      acc_v = fmaf(wr, N.av * AK + Kx * N.bxv + Ky * N.dyv, acc_v);
    }
  };
  // The drain's first pass: the same sums at the points of lane o outside
  // the core (|z|^2 >= adj_core_r2(y)) in the six-node form, its gradient in the
  // node-pair form of kgrad_mixed; the core points are flagged in cmask for
  // the second pass (kgrad_cr32, one point per lane).  (A select: at a core
  // point the node sum may be infinite.)
  auto near_pair_a = [&](int qi, int o, const float (&wr_o)[P], float& acc_t, float& acc_v,
                         uint32_t& cmask) {
    const VjpFar& R = s_far[qi];
    const VjpNear& N = s_near[qi];
    const int64_t i0o = t0 + (int64_t)slice * (kBlock * P) + (int64_t)(wv * 64 + o) * P;
    const double nu_mo = __dadd_rn(grid_nu(a.nu_min, a.step, i0o), 0.5 * (P - 1) * a.step);
    const float xm = (float)(nu_mo - R.c) * R.s;
    const float y = N.y, rc2 = adj_core_r2(y);
    const float cs[3] = {kGh6W[0] * (float)(2.0 * kInvPi), kGh6W[1] * (float)(2.0 * kInvPi),
                         kGh6W[2] * (float)(2.0 * kInvPi)};
    acc_t = 0.0f;
    acc_v = 0.0f;
    cmask = 0u;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const float x = fmaf((float)j - 0.5f * (float)(P - 1), R.hs, xm);
      const bool core = x * x + y * y < rc2;
      float K, Kx, Ky;
      kgrad_pairs32<3>(x, y, kGh6T2, cs, K, Kx, Ky);
      const float AK = N.a * K;
      Kx *= N.a;
      Ky *= N.a;
      const float ct = wr_o[j] * (N.at * AK + Kx * (ax_t * x) + Ky * N.dyt);
      const float cv = wr_o[j] * (N.av * AK + Kx * N.bxv + Ky * N.dyv);
      acc_t += core ? 0.0f : ct;
      acc_v += core ? 0.0f : cv;
      cmask |= core ? (1u << j) : 0u;
    }
  };
  for (int32_t base = 0; base < n_stage; base += kVjpChunk) {
    const int cnt = min(kVjpChunk, n_stage - base);
    __syncthreads();   // previous chunk consumed (s_far, s_near, s_wcnt, s_idx, s_ds1)
    VjpFar f;
    VjpNear n;
    int gas = -1;
    int32_t li = -1;
    float ds1 = 0.0f;
    if (tid < cnt) {
      li = (int32_t)stage_index(base + tid);
      const int64_t r = kb + li;
      const MixedRec& mr = recs[r];
      double c = mr.c;
      float s = mr.s, y = mr.y, am = mr.a;
      float at = a.at[r], dyt = a.dyt[r], av = a.av[r], bxv = a.bxv[r], dyv = a.dyv[r];
      if (!(c >= lo_b && c <= hi_b)) {
        // outside this tile's window: a record that adds exactly zero
        c = 0.0; s = 0.0f; y = 1.0e3f; am = 0.0f;
        at = dyt = av = bxv = dyv = 0.0f;
      }
      const float y2 = y * y;
      const float d2 = 2.0f * y2;
      const float ay = am * y * (float)kInvSqrtPi;
      const float aq = am * (float)kInvSqrtPi;
      const float cty = dyt * aq, cvy = dyv * aq;
      f.c = c; f.s = s; f.hs = stepf * s;
      f.c1 = y2 - 1.5f;
      // thresholds on |x| at the middle of a thread's run of P points
      const float half_run = 0.5f * (float)(P - 1) * f.hs;
      f.hw1 = half_run + sqrtf(fmaxf(s1_r2_min(y2) - y2, 0.0f));
      f.hw2 = half_run + sqrtf(fmaxf(kF32Gh2R2 - y2, 0.0f));
      f.tF = at * ay + cty; f.tX2 = 2.0f * ax_t * ay; f.tD = cty * d2; f.tQ = -2.0f * d2 * cty;
      f.tA = f.tF - f.tX2;
      f.vF = av * ay + cvy; f.vX = 2.0f * bxv * ay; f.vD = cvy * d2; f.vQ = -2.0f * d2 * cvy;
      n.y = y; n.a = am; n.at = at; n.dyt = dyt; n.av = av; n.bxv = bxv; n.dyv = dyv;
      ds1 = mr.ds1;
      gas = a.gas_of_group[a.group[li]];
      if (gas >= n_gas) gas = -1;   // cannot happen for validated inputs
      n.gas = gas;
    }
    // Counting sort by gas: rank inside the wave, then per-wave counts.
    int rank = 0;
    for (int g = 0; g < n_gas; ++g) {
      const unsigned long long m = __ballot(gas == g);
      if (gas == g) rank = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) s_wcnt[wv][g] = __popcll(m);
    }
    __syncthreads();
    if (gas >= 0) {
      int off = rank;
      for (int g = 0; g < gas; ++g)
        for (int w = 0; w < kWaves; ++w) off += s_wcnt[w][g];
      for (int w = 0; w < wv; ++w) off += s_wcnt[w][gas];
      s_far[off] = f;
      s_near[off] = n;
      s_idx[off] = li;
      s_ds1[off] = ds1;
    }
    __syncthreads();
    int q0 = 0;
    for (int g = 0; g < n_gas; ++g) {
      int len = 0;
      for (int w = 0; w < kWaves; ++w) len += s_wcnt[w][g];
      // uniform across the workgroup: keep the record loop on the scalar unit
      const int q1 = q0 + __builtin_amdgcn_readfirstlane(len);
      float ft = 0.0f, fv = 0.0f;
      double dv = 0.0;
      int naq = 0;   // queue entries of this run (uniform across the wavefront)
      for (int qi = q0; qi < q1; ++qi) {
        const VjpFar& R = s_far[qi];
        {
          // taken by a series (uniform across the wavefront): the same tests
          // as the far field's, on the same values
          const int32_t li = s_idx[qi];
          const float ds1 = s_ds1[qi];
          const bool by_slice = (li < slo || li >= shi) &&
                                fabsf((float)(nu_s - R.c)) >= ds1 + dS;
          const bool by_wave = (li < nlo || li >= nhi) && fabsf((float)(nu_w - R.c)) >= ds1;
          if (by_slice || by_wave) continue;
        }
        // x at the middle of the run; point j at xm + (j - (P-1)/2) hs
        const float xm = (float)(nu_m - R.c) * R.s;
        const float hs = R.hs;
        if (fabsf(xm) >= R.hw1) {
          const float c1 = R.c1;
          float S1 = 0.0f, S2 = 0.0f, S3o = 0.0f;
#pragma unroll
          for (int j = 0; j < P; ++j) {
            const float x = fmaf((float)j - 0.5f * (float)(P - 1), hs, xm);
            const float r = __builtin_amdgcn_rcpf(fmaf(x, x, c1));
            const float r2 = r * r;
            S1 = fmaf(WR[j], r, S1);
            S2 = fmaf(WR[j], r2, S2);
            S3o = fmaf(oW[j], r2, S3o);
          }
          const float S3 = fmaf(xm, S2, hs * S3o);   // sum W x r^2
          // x^2 r^2 = r - c1 r^2, so sum W x^2 r^2 = S1 - c1 S2 and the temperature
          // share tF S1 - tX2 (S1 - c1 S2) - tD S2 is (tF - tX2) S1 + (tX2 c1 - tD) S2;
          // both shares go straight into the running sums
          ft = fmaf(R.tA, S1, fmaf(fmaf(R.tX2, c1, -R.tD), S2, ft));
          fv = fmaf(R.vF, S1, fmaf(-R.vX, S3, fmaf(-R.vD, S2, fv)));
        } else if (fabsf(xm) >= R.hw2) {
          const float cw = R.c1 + 1.0f, d2 = fmaf(2.0f, R.c1, 3.0f);
          float S1 = 0.0f, S2 = 0.0f, S3 = 0.0f, S4 = 0.0f, S5 = 0.0f;
#pragma unroll
          for (int j = 0; j < P; ++j) {
            const float x = fmaf((float)j - 0.5f * (float)(P - 1), hs, xm);
            const float w = fmaf(x, x, cw);
            const float D = fmaf(w, w, d2);
            const float rD = __builtin_amdgcn_rcpf(D);
            const float w1 = w + 1.0f;
            S1 = fmaf(WR[j], w1 * rD, S1);
            const float q = (WR[j] * rD) * rD;
            const float qe = q * fmaf(-w, w + 2.0f, d2);
            const float xqe = x * qe;
            S2 += qe;
            S3 += xqe;
            S4 = fmaf(x, xqe, S4);
            S5 = fmaf(q, w1, S5);
          }
          ft += fmaf(R.tF, S1, fmaf(R.tX2, S4, fmaf(R.tD, S2, R.tQ * S5)));
          fv += fmaf(R.vF, S1, fmaf(R.vX, S3, fmaf(R.vD, S2, R.vQ * S5)));
        }
        // near-centre pairs: appended to this wavefront's queue (one ballot)
        // and drained densely after the run, instead of one lane's kgrad
        // loop holding the whole wavefront
        const bool nearc = live && !(fabsf(xm) >= R.hw1) && !(fabsf(xm) >= R.hw2);
        const uint64_t msk = __ballot(nearc);
        if (msk) {   // uniform
          const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
              (uint32_t)(msk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)msk, 0u));
          const int pos = naq + (int)rank;
          if (nearc) {
            if (pos < kAdjQueueCap) {
              s_aq[wv][pos] = (uint16_t)((qi << 6) | lane);
            } else {   // queue full: evaluate here
              float at_, av_;
              near_pair(qi, lane, WR, at_, av_);
              ft += at_;
              fv += av_;
            }
          }
          naq += (int)__popcll(msk);
        }
        if (((qi - q0) & (kVjpFold - 1)) == kVjpFold - 1) {
          g_t += (double)ft;
          dv += (double)fv;
          ft = fv = 0.0f;
        }
      }
      naq = min(naq, kAdjQueueCap);
      if (naq > 0) {   // uniform
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // second pass: the listed core points, one per lane (every lane takes
        // part in the owners' WR shuffles)
        auto core_pass = [&](const int m) {
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the list writes
          __builtin_amdgcn_wave_barrier();
          for (int b0 = 0; b0 < m; b0 += 64) {   // uniform
            const int b = b0 + lane;
            const bool ok = b < m;
            const uint32_t item = ok ? (uint32_t)s_acl[wv][b] : 0u;
            const int j = (int)(item & 7u);
            const uint32_t ent = (uint32_t)s_aq[wv][item >> 3];
            const int o = (int)(ent & 63u), qi = (int)(ent >> 6);
            float wr = 0.0f;
#pragma unroll
            for (int jj = 0; jj < P; ++jj) {
              const float w_ = __shfl(WR[jj], o);
              wr = jj == j ? w_ : wr;
            }
            if (ok) {
              const VjpFar& R = s_far[qi];
              const VjpNear& N = s_near[qi];
              const int64_t i0o = t0 + (int64_t)slice * (kBlock * P) + (int64_t)(wv * 64 + o) * P;
              const double nu_mo = __dadd_rn(grid_nu(a.nu_min, a.step, i0o), 0.5 * (P - 1) * a.step);
              const float xm = (float)(nu_mo - R.c) * R.s;
              const float x = fmaf((float)j - 0.5f * (float)(P - 1), R.hs, xm);
              float K, Kx, Ky;
              kgrad_cr32(x, N.y, K, Kx, Ky);
              const float AK = N.a * K;
              Kx *= N.a;
              Ky *= N.a;
              atomicAdd(&s_at[wv * 64 + o], wr * (N.at * AK + Kx * (ax_t * x) + Ky * N.dyt));
              atomicAdd(&s_av[wv * 64 + o], wr * (N.av * AK + Kx * N.bxv + Ky * N.dyv));
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // before the list is reused
          __builtin_amdgcn_wave_barrier();
        };
        int ncore = 0;   // uniform
        for (int e0 = 0; e0 < naq; e0 += 64) {   // every lane takes part (the owners' WR by shuffle)
          const int e = e0 + lane;
          const uint32_t ent = e < naq ? (uint32_t)s_aq[wv][e] : 0u;
          const int o = (int)(ent & 63u);
          float wro[P];
#pragma unroll
          for (int j = 0; j < P; ++j) wro[j] = __shfl(WR[j], o);
          float at_ = 0.0f, av_ = 0.0f;
          uint32_t cmask = 0u;
          if (e < naq) near_pair_a((int)(ent >> 6), o, wro, at_, av_, cmask);
          // this lane's core points go to the list at an exclusive prefix sum
          // of the lanes' counts (bit planes of the 0..8 count, one ballot each)
          const uint32_t cn = __builtin_popcount(cmask);
          uint32_t pre = 0, total = 0;
#pragma unroll
          for (int bit = 0; bit < 4; ++bit) {
            const uint64_t m = __ballot((cn >> bit) & 1u);
            pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << bit;
            total += (uint32_t)__popcll(m) << bit;
          }
          if (total) {   // uniform (<= 64 P = kAdjCoreCap)
            if (ncore + (int)total > kAdjCoreCap) {
              core_pass(ncore);
              ncore = 0;
            }
            int off = ncore + (int)pre;
#pragma unroll
            for (int j = 0; j < P; ++j) {
              if (cmask & (1u << j)) {
                s_acl[wv][off] = (uint16_t)((uint32_t)e << 3 | (uint32_t)j);
                ++off;
              }
            }
            ncore += (int)total;
          }
          if (e < naq) {
            atomicAdd(&s_at[wv * 64 + o], at_);
            atomicAdd(&s_av[wv * 64 + o], av_);
          }
        }
        if (ncore > 0) core_pass(ncore);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        ft += s_at[tid];
        fv += s_av[tid];
        s_at[tid] = 0.0f;
        s_av[tid] = 0.0f;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      g_t += (double)ft;
      dv += (double)fv;
#pragma unroll
      for (int gg = 0; gg < 8; ++gg)
        if (gg == g) g_v[gg] += dv;
      q0 = q1;
    }
  }
  g_t = wave_sum(g_t);
  if (lane == 0) atomicAdd(&a.grad_t[k], g_t);
  for (int g = 0; g < n_gas && g < 8; ++g) {
    const double sv = wave_sum(g_v[g]);
    if (lane == 0) atomicAdd(&a.grad_vmr[(int64_t)g * a.n_layers + k], sv);
  }
}

__global__ void k_fill_f64(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace

int vjp_points_per_thread() { return kVjpPoints; }

hipError_t launch_line_grad(const lbl_lines_t& lines, const lbl_groups_t& groups,
                            const lbl_layers_t& layers, float* at, float* dyt, float* av,
                            float* bxv, float* dyv, hipStream_t s) {
  if (lines.n_lines <= 0 || layers.n_layers <= 0) return hipSuccess;
  GradArgs g{lines, groups, layers, at, dyt, av, bxv, dyv};
  dim3 grid((unsigned)((lines.n_lines + 255) / 256), (unsigned)layers.n_layers);
  k_line_grad<<<grid, 256, 0, s>>>(g);
  return hipGetLastError();
}

hipError_t launch_vjp(const Records& rec, const float* at, const float* dyt, const float* av,
                      const float* bxv, const float* dyv, const int32_t* group,
                      const int32_t* gas_of_group, const int32_t* lo, const int32_t* hi,
                      const int32_t* wlo, const int32_t* whi, const int32_t* slo,
                      const int32_t* shi, const double* t_k, int64_t n_lines, int n_layers,
                      int n_gases,
                      const lbl_grid_t& grid, int64_t n_tiles, int n_mu, const double* mu,
                      const float* trans, const float* grad_trans, const float* od_layers,
                      double* grad_t, double* grad_vmr, hipStream_t s) {
  const int64_t nt = (int64_t)n_layers * (1 + n_gases);
  if (n_layers > 0) {
    k_fill_f64<<<(unsigned)((n_layers + 255) / 256), 256, 0, s>>>(grad_t, n_layers);
    if (n_gases > 0)
      k_fill_f64<<<(unsigned)(((int64_t)n_gases * n_layers + 255) / 256), 256, 0, s>>>(
          grad_vmr, (int64_t)n_gases * n_layers);
  }
  (void)nt;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_tiles <= 0 || n_layers <= 0) return e;
  VjpArgs a{};
  a.rec = rec;
  a.at = at; a.dyt = dyt; a.av = av; a.bxv = bxv; a.dyv = dyv;
  a.group = group;
  a.gas_of_group = gas_of_group;
  a.lo = lo;
  a.hi = hi;
  a.wlo = wlo;
  a.whi = whi;
  a.slo = slo;
  a.shi = shi;
  a.t_k = t_k;
  a.n_lines = n_lines;
  a.n_layers = n_layers;
  a.n_gases = n_gases;
  a.nu_min = grid.nu_min;
  a.step = grid.step;
  a.cutoff = grid.cutoff;
  a.n_nu = grid.n_nu;
  a.ts = grid.tile_samples;
  const int64_t pts = (int64_t)kBlock * kVjpPoints;
  a.slices_per_tile = (int32_t)((grid.tile_samples + pts - 1) / pts);
  a.n_wg = n_tiles * a.slices_per_tile;
  a.n_mu = n_mu;
  for (int m = 0; m < n_mu; ++m) a.inv_mu[m] = 1.0 / mu[m];
  a.trans = trans;
  a.grad_trans = grad_trans;
  a.od_layers = od_layers;
  a.grad_t = grad_t;
  a.grad_vmr = grad_vmr;
  // one workgroup per (slice, layer); a 1-D grid holds at most 2^32 work-items
  const int64_t n_blocks = a.n_wg * (int64_t)n_layers;
  if (n_blocks * kBlock > (int64_t)UINT32_MAX) return hipErrorInvalidConfiguration;
  k_vjp_mixed<kVjpPoints><<<(unsigned)n_blocks, kBlock, 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace lbl

// Device Voigt function K(x, y) = Re w(x + iy), y >= 0, in FP64 and FP32.
//
// w(z) = exp(-z^2) erfc(-iz) is the Faddeeva function the reference evaluates
// with scipy.special.wofz (pylblrtm/forward.py:22,62).  Everything here is
// written from published mathematics:
//
// * Far region: n-point Gauss-Hermite quadrature of
//     w(z) = (i/pi) * integral exp(-t^2) / (z - t) dt,
//   i.e. the even truncations of the Laplace continued fraction.  Grouping the
//   symmetric nodes +-t with weight W gives, with r = x^2 + y^2,
//     Re[(i/pi) W (1/(z-t) + 1/(z+t))] = (2 y W / pi) (r + t^2) / D,
//     D = (x^2 - y^2 - t^2)^2 + 4 x^2 y^2,
//   a strictly positive multiple of y, so K keeps its relative accuracy on
//   the real axis.  Measured relative error in K (against mpmath):
//     2 nodes (Humlicek's region-1 form) 2.5/|z|^4, 4 nodes 13.5/|z|^8,
//     8 nodes < 1e-15 for |z| >= 15.
// * Core region: the trapezoid rule with step h on the same integral plus the
//   residue of its pole at t = z (Chiarella & Reichel, Math. Comp. 22 (1968)
//   137; Matta & Reichel, Math. Comp. 25 (1971) 339), truncation error
//   ~exp(-pi^2/h^2).  The node lattice (t = n h or t = (n + 1/2) h) is picked
//   per point so that x stays at least h/4 from every node, which keeps the
//   pole terms and the residue from cancelling.  With p = exp(2 pi i z / h),
//   rho = |p| = exp(-2 pi y / h) and sg = +1 for the half-shifted lattice,
//   -1 for the integer one, the residue adds (for y < pi/h)
//     Re C = 2 M (sg cos(theta - 2xy) + rho cos 2xy) / (1 + 2 sg rho cos theta + rho^2),
//     M = exp(y^2 - x^2 - 2 pi y / h), theta = 2 pi x / h,
//   which reduces to exp(-x^2) on the real axis.
//   FP64: h = 0.5, |t| <= 6.3 -> ~1e-15 relative.  FP32: h = 0.8,
//   |t| <= 4.2 -> ~3e-6 relative including FP32 rounding.
#pragma once

#include <hip/hip_runtime.h>

namespace lbl {

constexpr double kPi = 3.14159265358979323846;
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kInvSqrtPi = 0.56418958354775628695;

// Region thresholds on r = x^2 + y^2.  FP64 keeps every region below 1.1e-11
// relative error in K (the FP64 mode's target is 1e-9 in tau).
constexpr double kFp64Gh2R2 = 490000.0; // |z| >= 700: 2 nodes, rel err <= 1.04e-11
constexpr double kFp64Gh4R2 = 1225.0;   // |z| >= 35 : 4 nodes, rel err <= 5.1e-12
constexpr double kFp64Gh8R2 = 225.0;    // |z| >= 15 : 8 nodes, rel err <= 1e-15
// |z| >= 7: 10 nodes, rel err <= 1e-11 for y >= 1e-8 (measured against
// scipy's wofz: the 1e-11 boundary lies at |z| = 6.71 for y in [1e-3, 1],
// 6.72 for [1e-5, 1e-3), 6.98 for [1e-8, 1e-5) and below it for larger y;
// nearer the real axis exp(-x^2), which no node sum holds, keeps it out)
constexpr double kFp64Gh10R2 = 49.0;
constexpr double kFp64Gh10Ymin = 1e-8;
constexpr float kF32Gh2R2 = 625.0f;     // |z| >= 25: 2 nodes, rel err <= 6.41e-6
constexpr float kF32Gh4R2 = 64.0f;      // |z| >= 8 : 4 nodes, rel err <= 1e-6

// Shifted one-node far form (mixed line sum): K ~= y / (sqrt(pi) (x^2 + y^2 - 3/2)).
// The asymptotic expansion K = (1/sqrt(pi)) sum_n (2n-1)!!/2^n sin((2n+1) theta)/|z|^(2n+1)
// (theta = arg z) is y/(sqrt(pi)|z|^2) (1 + 3/(2|z|^2) + (15/4 - 2y^2)/|z|^4 + ...);
// the shifted term reproduces the first two orders, so its relative error is
// (2 y^2 - 3/2)/|z|^4 + O(|z|^-6): measured |err| <= 1.013 (3/2 + 2 y^2)/|z|^4 for
// |z| >= 20 against the series summed to 40 terms.  It is used where
// 1.05 (3/2 + 2 y^2)/|z|^4 <= 6.4e-6 (the two-node form's bound at |z| = 25)
// and |z| >= 25, i.e. r >= max(625, sqrt(kF32S1Coef (3/2 + 2 y^2))).
constexpr float kF32S1Coef = 164062.5f;  // 1.05 / 6.4e-6
__device__ __forceinline__ float s1_r2_min(float y2) {
  return fmaxf(kF32Gh2R2, sqrtf(kF32S1Coef * fmaf(2.0f, y2, 1.5f)));
}
// FP64 mode: the same form where 1.05 (3/2 + 2 y^2)/|z|^4 <= 1e-11 (|z| >= 630
// for small y, the two-node form's 1.04e-11 at |z| = 700).
constexpr double kFp64S1Coef = 1.05e11;   // 1.05 / 1e-11
__device__ __forceinline__ double s1_r2_min64(double y2) {
  return sqrt(kFp64S1Coef * fma(2.0, y2, 1.5));
}

// Second-order shifted form (FP64 line sum), r = |z|^2:
//   K ~= y r / (sqrt(pi) (r^2 - 3/2 r + 2 y^2 - 3/2)),
// i.e. y / (sqrt(pi) (r - 3/2 + (2y^2 - 3/2)/r)): the shift's own correction
// makes the first three orders of the expansion above exact, leaving
// (9 y^2 - 21/4)/|z|^6 + O(|z|^-8); measured |err| <= 1.1 (21/4 + 9 y^2)/|z|^6
// for |z| >= 20 against the series.  Used where that is <= 1e-11 (|z| >= 91
// for small y, 368 at y = 50).
constexpr double kFp64S2Coef = 1.1e11;    // 1.1 / 1e-11
__device__ __forceinline__ double s2_r2_min64(double y2) {
  return cbrt(kFp64S2Coef * fma(9.0, y2, 5.25));
}
// Mixed line sum: the same form where 1.1 (21/4 + 9 y^2)/|z|^6 <= 6.4e-6
// (|z| >= 9.8 for small y); it reaches below the two-node form's |z| = 25 while
// y^2 < kF32S2Y2, and there it replaces the two-node band.
constexpr float kF32S2Coef = 171875.0f;   // 1.1 / 6.4e-6
constexpr float kF32S2Y2 = 155.0f;        // s2 radius^2 = 625 at y^2 = 157.4
__device__ __forceinline__ float s2_r2_min32(float y2) {
  return cbrtf(kF32S2Coef * fmaf(9.0f, y2, 5.25f));
}

// Gauss-Hermite positive nodes / weights (numpy.polynomial.hermite.hermgauss).
constexpr double kGh2T = 0.7071067811865475, kGh2W = 0.8862269254527579;
constexpr double kGh4T[2] = {0.5246476232752904, 1.6506801238857847};
constexpr double kGh4W[2] = {0.8049140900055127, 0.08131283544724519};
constexpr double kGh8T[4] = {0.3811869902073221, 1.1571937124467802, 1.981656756695843,
                             2.930637420257244};
constexpr double kGh8W[4] = {0.6611470125582415, 0.20780232581489183, 0.017077983007413467,
                             0.00019960407221136783};

// 10-point Gauss-Hermite nodes and weights (t > 0), refined by Newton steps on
// H_10 in 50-digit arithmetic.
constexpr double kGh10T[5] = {0.34290132722370461, 1.0366108297895137, 1.7566836492998818,
                              2.5327316742327898, 3.4361591188377376};
constexpr double kGh10W[5] = {0.6108626337353258, 0.24013861108231469, 0.033874394455481063,
                              0.0013436457467812327, 7.6404328552326206e-6};

// exp(-t^2) at the trapezoid nodes: t = n h (n >= 1) and t = (n + 1/2) h (n >= 0).
constexpr double kCr64H = 0.5;
constexpr double kCr64E0[12] = {
    0.7788007830714049, 0.36787944117144233, 0.10539922456186433, 0.01831563888873418,
    0.0019304541362277093, 0.00012340980408667956, 4.785117392129009e-06,
    1.1253517471925912e-07, 1.6052280551856116e-09, 1.3887943864964021e-11,
    7.287724095819692e-14, 2.3195228302435696e-16};
constexpr double kCr64E5[13] = {
    0.9394130628134758, 0.569782824730923, 0.2096113871510978, 0.04677062238395898,
    0.006329715427485747, 0.0005195746821548384, 2.586810022265412e-05,
    7.811489408304491e-07, 1.4307241918567688e-08, 1.5893910094516368e-10,
    1.0709232382508077e-12, 4.37661850287085e-15, 1.0848552640429378e-17};
constexpr float kCr32H = 0.8f;
constexpr float kCr32E0[5] = {0.5272924240430485f, 0.07730474044329971f, 0.0031511115984444358f,
                              3.5712849641635144e-05f, 1.1253517471925912e-07f};
constexpr float kCr32E5[5] = {0.8521437889662113f, 0.23692775868212165f, 0.01831563888873418f,
                              0.0003936690406550776f, 2.352575200009771e-06f};

// ---------------------------------------------------------------- FP64 ----

// v_rcp_f64 (relative error about 2^-23) and one Newton step: error about
// 2^-46, far below the FP64 forms' own 1.1e-11.  Quotients below use it in
// place of the correctly rounded division sequence.
__device__ __forceinline__ double rcp64_1nr(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

__device__ __forceinline__ double gh_pair64(double r, double xmy, double fxy, double t, double w) {
  const double t2 = t * t;
  const double u = xmy - t2;
  return w * (r + t2) * rcp64_1nr(fma(u, u, fxy));
}

// sum_k W_k (r + t_k^2) / D_k over N node pairs as a single quotient: every
// term is positive, so the running numerator and denominator (the product of
// the D_k) keep their relative accuracy, and one reciprocal serves all pairs.
// The caller keeps the D_k away from 0 and the product inside the FP64 range
// (D_k = (r - t_k^2)^2 + 4 y^2 t_k^2: for |z| >= 7 and t_k <= 3.44 each lies
// in [1.3e3, ~r^2]).
template <int N>
__device__ __forceinline__ double gh_sum64(double r, double xmy, double fxy, const double (&T)[N],
                                           const double (&W)[N]) {
  double t2 = T[0] * T[0];
  double u = xmy - t2;
  double den = fma(u, u, fxy);
  double num = W[0] * (r + t2);
#pragma unroll
  for (int k = 1; k < N; ++k) {
    t2 = T[k] * T[k];
    u = xmy - t2;
    const double D = fma(u, u, fxy);
    num = fma(num, D, W[k] * (r + t2) * den);
    den *= D;
  }
  return num * rcp64_1nr(den);
}

// 2-node form, |z| >= 700: K = y (w + 1) / (sqrt(pi) (w^2 + 2 y^2)), w = x^2 + y^2 - 1/2.
__device__ __forceinline__ double K_gh2_64(double x2, double y) {
  const double y2 = y * y;
  const double w = x2 + (y2 - 0.5);
  return y * kInvSqrtPi * (w + 1.0) * rcp64_1nr(fma(w, w, 2.0 * y2));
}

// 4-node form, |z| >= 35.
__device__ __forceinline__ double K_gh4_64(double x2, double y) {
  const double y2 = y * y, r = x2 + y2, xmy = x2 - y2, fxy = 4.0 * x2 * y2;
  const double s = gh_pair64(r, xmy, fxy, kGh4T[0], kGh4W[0]) +
                   gh_pair64(r, xmy, fxy, kGh4T[1], kGh4W[1]);
  return 2.0 * y * kInvPi * s;
}

// 8-node form, |z| >= 15.
__device__ __forceinline__ double K_gh8_64(double x2, double y) {
  const double y2 = y * y, r = x2 + y2, xmy = x2 - y2, fxy = 4.0 * x2 * y2;
  double s = 0.0;
#pragma unroll
  for (int k = 3; k >= 0; --k) s += gh_pair64(r, xmy, fxy, kGh8T[k], kGh8W[k]);
  return 2.0 * y * kInvPi * s;
}

// 10-node form, |z| >= 7 (y >= 1e-8).
__device__ __forceinline__ double K_gh10_64(double x2, double y) {
  const double y2 = y * y, r = x2 + y2, xmy = x2 - y2, fxy = 4.0 * x2 * y2;
  return 2.0 * y * kInvPi * gh_sum64<5>(r, xmy, fxy, kGh10T, kGh10W);
}

// Trapezoid + residue, |z| < 7 (and |z| < 15 for y < 1e-8).
__device__ inline double K_cr64(double x, double y) {
  const double h = kCr64H;
  const double q = fabs(x) * (1.0 / h);
  const double fr = q - floor(q);
  const bool half = !(fr >= 0.25 && fr <= 0.75);
  const double x2 = x * x, y2 = y * y, r = x2 + y2, xmy = x2 - y2, fxy = 4.0 * x2 * y2;
  // One loop for both lattices (nodes (n + 1/2) h or n h, weights selected per
  // lane), so a wavefront whose lanes chose differently runs it once.  The
  // integer lattice's node t = 0 counts once: its slot has weight 1/2 of the
  // pair form, (r + 0)/D = 1/r.  The pairs are summed as one quotient (all
  // terms positive; on the chosen lattice x stays h/4 from every node, so
  // each D >= 3.9e-3, and for |z| < 15 each D = (r - t^2)^2 + 4 y^2 t^2 <
  // 1.3e5: 13 factors stay inside the FP64 range).
  const double off = half ? 0.5 : 0.0;
  double num = 0.0, den = 1.0;
#pragma unroll
  for (int n = 12; n >= 0; --n) {
    const double w = half ? kCr64E5[n] : (n >= 1 ? kCr64E0[n - 1] : 0.5);
    const double t = (n + off) * h, t2 = t * t;
    const double u = xmy - t2;
    const double D = fma(u, u, fxy);
    num = fma(num, D, w * (r + t2) * den);
    den *= D;
  }
  double K = 2.0 * y * h * kInvPi * num * rcp64_1nr(den);
  // The residue is at most 4M in size: its numerator is at most 2M (1 + rho)
  // and, on the lattice chosen above (cos theta > 0 on the half lattice,
  // <= 0 on the integer one), its denominator is at least 1.  Where 4M is
  // below 1e-15 of the pole-pair sum it is left out (never when y = 0).
  const double M = exp(y2 - x2 - 2.0 * kPi * y / h);
  if (y < kPi / h && 4.0 * M > 1e-15 * K) {
    const double sg = half ? 1.0 : -1.0;
    const double rho = exp(-2.0 * kPi * y / h);
    double st, ct, s2, c2;
    sincospi(2.0 * x / h, &st, &ct);
    // 2xy as a multiple of pi: the reduction is exact, and the two roundings
    // of the argument move the phase by < 3e-14 for 2xy < 90 (y < pi/h, |x| < 7)
    sincospi(2.0 * x * y * kInvPi, &s2, &c2);
    const double cphi = ct * c2 + st * s2;  // cos(theta - 2xy)
    K += 2.0 * M * (sg * cphi + rho * c2) * rcp64_1nr(1.0 + 2.0 * sg * rho * ct + rho * rho);
  }
  return K;
}

// Full FP64 region dispatch: relative error <= 1.1e-11 everywhere.
__device__ inline double K_fp64(double x, double y) {
  const double x2 = x * x;
  const double r = x2 + y * y;
  if (r >= kFp64Gh2R2) return K_gh2_64(x2, y);
  if (r >= kFp64Gh4R2) return K_gh4_64(x2, y);
  if (r >= kFp64Gh8R2) return K_gh8_64(x2, y);
  if (r >= kFp64Gh10R2 && y >= kFp64Gh10Ymin) return K_gh10_64(x2, y);
  return K_cr64(x, y);
}

// ---------------------------------------------------------------- FP32 ----
// Quotients use v_rcp_f32 (1 ulp) and a multiply instead of the correctly
// rounded division sequence: 2 ulp, far inside the FP32 forms' own error.

__device__ __forceinline__ float gh_pair32(float r, float xmy, float fxy, float t2, float w) {
  const float u = xmy - t2;
  return w * (r + t2) * __builtin_amdgcn_rcpf(fmaf(u, u, fxy));
}

// 2-node form: K = y (r + 1/2) / (sqrt(pi) D), D = x^4 + x^2 (2y^2 - 1) + (y^2 + 1/2)^2.
__device__ __forceinline__ float K_gh2_32(float x2, float y) {
  const float y2 = y * y;
  const float c = y2 + 0.5f;
  const float D = fmaf(x2, x2 + (2.0f * y2 - 1.0f), c * c);
  return y * (float)kInvSqrtPi * (x2 + c) * __builtin_amdgcn_rcpf(D);
}

// 4-node form, |z| >= 8.
__device__ __forceinline__ float K_gh4_32(float x2, float y) {
  const float y2 = y * y, r = x2 + y2, xmy = x2 - y2, fxy = 4.0f * x2 * y2;
  const float t0 = (float)(kGh4T[0] * kGh4T[0]), t1 = (float)(kGh4T[1] * kGh4T[1]);
  const float u0 = xmy - t0, u1 = xmy - t1;
  const float D0 = fmaf(u0, u0, fxy), D1 = fmaf(u1, u1, fxy);
  const float N = (float)kGh4W[0] * (r + t0) * D1 + (float)kGh4W[1] * (r + t1) * D0;
  return 2.0f * y * (float)kInvPi * N * __builtin_amdgcn_rcpf(D0 * D1);
}

// 6-point Gauss-Hermite nodes (t > 0), their squares and weights (50-digit
// Newton refinement on H_6, rounded to FP32).
constexpr float kGh6T2[3] = {0.19016350919348813f, 1.7844927485432516f, 5.5253437422632603f};
constexpr float kGh6W[3] = {0.72462959522439252f, 0.15706732032285664f, 0.0045300099055088456f};
// Squared radius below which the six-node form misses 5e-6 (FP32
// arithmetic), as a function of y^2: 7.0 for y < 1e-4 (near the real axis exp(-x^2) takes over), 5.5 for y < 1e-3,
// 5.0 for y < 2, 3.6 above (measured worst outside it: 3.2e-6).
__device__ __forceinline__ float gh6_core_r2(float y2) {
  return y2 < 1e-8f ? 49.0f : y2 < 1e-6f ? 30.25f : y2 < 4.0f ? 25.0f : 12.96f;
}
// sum_k W_k (r + t_k^2) / D_k over the six nodes as one quotient (all terms
// positive; for |z| >= 3.6 each D_k >= 52, and a product of three stays far
// inside the FP32 range for |z| < 1e6).
__device__ __forceinline__ float gh6_sum32(float r, float xmy, float fxy) {
  float u = xmy - kGh6T2[0];
  float den = fmaf(u, u, fxy);
  float num = kGh6W[0] * (r + kGh6T2[0]);
#pragma unroll
  for (int k = 1; k < 3; ++k) {
    u = xmy - kGh6T2[k];
    const float D = fmaf(u, u, fxy);
    num = fmaf(num, D, kGh6W[k] * (r + kGh6T2[k]) * den);
    den *= D;
  }
  return num * __builtin_amdgcn_rcpf(den);
}

// exp(-2 pi y / h), the residue's rho, per line (the near-field stage keeps it
// with the record).
__device__ __forceinline__ float cr32_rho(float y) {
  return expf(-(2.0f * (float)kPi / kCr32H) * y);
}

// Trapezoid + residue, |z| < 8, with rho = cr32_rho(y).
__device__ inline float K_cr32(float x, float y, float rho) {
  const float h = kCr32H;
  const float q = fabsf(x) * (1.0f / h);
  const float fr = q - floorf(q);
  const bool half = !(fr >= 0.25f && fr <= 0.75f);
  const float x2 = x * x, y2 = y * y, r = x2 + y2, xmy = x2 - y2, fxy = 4.0f * x2 * y2;
  // One loop for both lattices, as in K_cr64: half-lattice nodes (n + 1/2) h
  // in slots n = 0..4 (slot 5 has weight 0), integer-lattice nodes n h in
  // slots 0..5 (the node t = 0 counts once: weight 1/2 of the pair form,
  // (r + 0)/D = 1/r).  The pairs are summed as one quotient: all terms are
  // positive, on the chosen lattice x stays h/4 from every node so each
  // D >= 1.6e-3, and for |z| < 8 each D < 9.1e3 (six factors < 6e23).
  float num = 0.0f, den = 1.0f;
#pragma unroll
  for (int n = 5; n >= 0; --n) {
    const float w = half ? (n <= 4 ? kCr32E5[n] : 0.0f) : (n >= 1 ? kCr32E0[n - 1] : 0.5f);
    const float th = (n + 0.5f) * h, ti = n * h;
    const float t2 = half ? th * th : ti * ti;
    const float u = xmy - t2;
    const float D = fmaf(u, u, fxy);
    num = fmaf(num, D, w * (r + t2) * den);
    den *= D;
  }
  float K = 2.0f * h * (float)kInvPi * y * num * __builtin_amdgcn_rcpf(den);
  // |residue| <= 4M, as in K_cr64; left out where below 1e-8 of the pair sum
  const float a = 2.0f * (float)kPi / h;
  const float M = expf(fmaf(y, y - a, -x2));
  if (y < (float)kPi / h && 4.0f * M > 1e-8f * K) {
    const float sg = half ? 1.0f : -1.0f;
    // theta = 2 pi x / h and 2 x y by the hardware sine and cosine, whose
    // argument is in revolutions (x / h and x y / pi; |x| < 8 and x y < 32
    // here, far inside their +-256 range): absolute error <= 1.2e-7 over
    // +-64 revolutions, the size of the FP32 rounding of the arguments
    const float r1 = x * (1.0f / h), r2 = x * y * (float)kInvPi;
    const float st = __builtin_amdgcn_sinf(r1), ct = __builtin_amdgcn_cosf(r1);
    const float s2 = __builtin_amdgcn_sinf(r2), c2 = __builtin_amdgcn_cosf(r2);
    const float cphi = ct * c2 + st * s2;
    K += 2.0f * M * (sg * cphi + rho * c2) * __builtin_amdgcn_rcpf(1.0f + 2.0f * sg * rho * ct + rho * rho);
  }
  return K;
}

__device__ inline float K_cr32(float x, float y) { return K_cr32(x, y, cr32_rho(y)); }

// Full FP32 region dispatch: relative error <= ~6.4e-6 everywhere.
__device__ inline float K_mixed(float x, float y) {
  const float x2 = x * x;
  const float r = x2 + y * y;
  if (r >= kF32Gh2R2) return K_gh2_32(x2, y);
  if (r >= kF32Gh4R2) return K_gh4_32(x2, y);
  return K_cr32(x, y);
}

// ------------------------------------------- FP32, K and its gradient ----
// K, dK/dx and dK/dy with the same region split as K_mixed, each region's
// approximant differentiated analytically (used by the adjoint).  The
// identity w' = -2 z w + 2i/sqrt(pi) is not used: for |z| >= 8 it cancels
// to O(1/|z|^2) and would multiply the approximant's error by ~2|z|^2.
//
// Symmetric node pair (node t, weight W), w = x^2 + y^2 - t^2, d = 4 t^2 y^2,
// D = w^2 + d (= (x^2 - y^2 - t^2)^2 + 4 x^2 y^2):
//   K  += (2W/pi) y F,            F  = (w + 2t^2)/D,
//   Kx += (2W/pi) 2 x y F',       F' = dF/dw = (4 t^2 (t^2 - x^2) - w^2)/D^2,
//   Ky += (2W/pi) (F + 2 y^2 F' - 2 d F/D).
template <int N>
__device__ inline void kgrad_pairs32(float x, float y, const float (&t2s)[N],
                                     const float (&cs)[N], float& K, float& Kx, float& Ky) {
  const float x2 = x * x, y2 = y * y, r = x2 + y2;
  float sf = 0.0f, sp = 0.0f, sy = 0.0f;
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    const float t2 = t2s[k], c = cs[k];
    const float w = r - t2;
    const float d = 4.0f * t2 * y2;
    const float rD = __builtin_amdgcn_rcpf(fmaf(w, w, d));
    const float F = (w + 2.0f * t2) * rD;
    const float Fp = fmaf(4.0f * t2, t2 - x2, -w * w) * rD * rD;
    sf = fmaf(c, F, sf);
    sp = fmaf(c, Fp, sp);
    sy = fmaf(c, F - 2.0f * d * F * rD, sy);
  }
  K = y * sf;
  Kx = 2.0f * x * y * sp;
  Ky = fmaf(2.0f * y2, sp, sy);
}

// Trapezoid + residue, |z| < 8.  The pole terms are node pairs with weight
// h exp(-t^2) (plus (h/pi) y/r for the pole at 0 on the integer lattice);
// the residue C = 2 exp(-z^2) sg p/(1 + sg p), p = exp(2 pi i z/h), has
//   dC/dz = -2 z C + (2 pi i/h) C/(1 + sg p),
// and dK/dx = Re dC/dz, dK/dy = -Im dC/dz for its share.
__device__ inline void kgrad_cr32(float x, float y, float& K, float& Kx, float& Ky) {
  const float h = kCr32H;
  const float q = fabsf(x) * (1.0f / h);
  const float fr = q - floorf(q);
  const bool half = !(fr >= 0.25f && fr <= 0.75f);
  float t2s[5], cs[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const float t = half ? (n + 0.5f) * h : (n + 1) * h;
    t2s[n] = t * t;
    cs[n] = (float)(2.0 * kInvPi) * h * (half ? kCr32E5[n] : kCr32E0[n]);
  }
  kgrad_pairs32<5>(x, y, t2s, cs, K, Kx, Ky);
  const float x2 = x * x, y2 = y * y;
  if (!half) {
    const float c = h * (float)kInvPi;
    const float rr = __builtin_amdgcn_rcpf(x2 + y2);
    K = fmaf(c * y, rr, K);
    Kx = fmaf(-2.0f * c * x * y, rr * rr, Kx);
    Ky = fmaf(c * (x2 - y2), rr * rr, Ky);
  }
  if (y < (float)kPi / h) {
    const float sg = half ? 1.0f : -1.0f;
    const float a = 2.0f * (float)kPi / h;
    const float rho = expf(-a * y);
    const float M = expf(fmaf(y, y - a, -x2));
    float st, ct, s2, c2;
    sincospif(x * (2.0f / h), &st, &ct);   // 2x/h (1 ulp; no run-time division)
    sincosf(2.0f * x * y, &s2, &c2);
    const float cphi = ct * c2 + st * s2;   // cos(theta - 2xy)
    const float sphi = st * c2 - ct * s2;   // sin(theta - 2xy)
    const float rden = __builtin_amdgcn_rcpf(1.0f + 2.0f * sg * rho * ct + rho * rho);
    const float g = 2.0f * M * rden;
    const float cr = g * (sg * cphi + rho * c2);
    const float ci = g * (sg * sphi - rho * s2);
    // u = 1/(1 + sg p) = (1 + sg rho cos theta - i sg rho sin theta) rden
    const float ur = (1.0f + sg * rho * ct) * rden, ui = -sg * rho * st * rden;
    const float cur = cr * ur - ci * ui, cui = cr * ui + ci * ur;
    K += cr;
    Kx += -2.0f * (x * cr - y * ci) - a * cui;
    Ky += 2.0f * (x * ci + y * cr) - a * cur;
  }
}

__device__ inline void kgrad_mixed(float x, float y, float& K, float& Kx, float& Ky) {
  const float r = x * x + y * y;
  if (r >= kF32Gh2R2) {
    const float t2s[1] = {(float)(kGh2T * kGh2T)};
    const float cs[1] = {(float)(2.0 * kGh2W * kInvPi)};
    kgrad_pairs32<1>(x, y, t2s, cs, K, Kx, Ky);
  } else if (r >= kF32Gh4R2) {
    const float t2s[2] = {(float)(kGh4T[0] * kGh4T[0]), (float)(kGh4T[1] * kGh4T[1])};
    const float cs[2] = {(float)(2.0 * kGh4W[0] * kInvPi), (float)(2.0 * kGh4W[1] * kInvPi)};
    kgrad_pairs32<2>(x, y, t2s, cs, K, Kx, Ky);
  } else {
    kgrad_cr32(x, y, K, Kx, Ky);
  }
}

}  // namespace lbl

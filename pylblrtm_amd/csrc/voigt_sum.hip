// Tiled Voigt line sum with the radiation term, layer sum and exp(-tau/mu)
// fused in (pylblrtm/forward.py:295-315 and :151).
//
// Work decomposition.  The reference evaluates every reference tile (a run
// of tile_samples grid points) against every line whose shifted centre lies
// in [tile first point - cutoff, tile last point + cutoff] (forward.py:298-301).
// Here one workgroup owns kBlock * P consecutive points of one tile (P
// points per thread, contiguous) and walks the layers in order; for each
// layer it streams that tile's candidate records through LDS in chunks of
// kBlock, keeps only those inside the tile window (exact FP64 compare), and
// every thread accumulates its P points.  At the end of a layer each point's
// sum is multiplied by R(nu, T_k) (forward.py:312), checked for finiteness
// (forward.py:314-315) and added to the column total; after the last layer
// the kernel writes exp(-tau/mu) for every requested mu.
//
// Region split.  Almost every evaluation is far from its line centre, where
// a short rational form is accurate: the shifted one-node form (4 VALU
// operations per point in mixed mode) and, closer in, its second-order
// refinement or (mixed, broad lines) the two-node Gauss-Hermite form, then in
// FP64 the four-node form (faddeeva.h gives each form's error).  A thread decides
// once per record which form holds for all of its P points (a compare of |x|
// at the middle of its run with a per-record threshold); pairs (record,
// thread) that need the full region dispatch of faddeeva.h go to an LDS work
// queue that the workgroup drains densely after each chunk of records.
//
// Precision.  FP64 mode: every operation in FP64.  Mixed mode: nu - nu_c is
// formed in FP64 from the exact grid point and the FP64 centre, rounded to
// FP32 once per (thread, record); the P points then step by step * s in
// FP32.  Sums are FP32 over at most 32 records, folded into a compensated
// (Kahan) per-layer sum, and the per-layer values into a compensated column
// total.
#include <math.h>

#include "faddeeva.h"
#include "internal.h"
#include "surrogate.h"

namespace lbl {
namespace {

constexpr int kChunk = kBlock;  // records staged per LDS round
constexpr int kFoldEvery = 32;  // records between compensated folds (mixed)
constexpr int kExCap = 128;     // far-field exception list entries per wavefront
constexpr int kPointsFp64 = 8;  // grid points per thread of the FP64 line sums
// Mixed near-centre drain: queued points nearer than the six-node form's
// radius (gh6_core_r2, 3.6 to 7 by y) take the trapezoid + residue form; a
// wavefront's list of such points holds kCoreCap entries.
constexpr int kCoreCap = 512;   // >= the 64 x 8 points one pass-A round can add
constexpr int kQueueW = 448;     // queue entries per wavefront (the loops drain past kQueueW - 64)
constexpr int kAccStride = 256;  // s_acc row length of the mixed line sums

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;
typedef __attribute__((address_space(3))) MixedRec lds_mixed_rec;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ __forceinline__ void kahan_add(float& sum, float& comp, float v) {
  const float y = v - comp;
  const float t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}


// Work counting variant (kCount): per-wavefront totals added once to the
// WorkCount slots of a.work (internal.h).
__device__ __forceinline__ void count_add(unsigned long long* w, int slot, uint64_t v) {
  if (v) atomicAdd(w + slot, (unsigned long long)v);
}

__device__ __forceinline__ void flag_od(uint32_t* err, int k) {
  atomicMin(err, err_numeric(k, 1, 0));
}

// ------------------------------------------------------------------ FP64 --

// The four-node form (|z| >= 35, faddeeva.h) holds at every point within Hu
// of a middle point d = nu_m - c away.  Shared by the series kernel's levels
// and the near field's test for the pairs the group level leaves to it, so
// both reach the same verdict bit for bit.
__device__ __forceinline__ bool gh4_holds_within(double d, double sv, double y2, double Hu) {
#pragma clang fp contract(off)
  const double xn = sv * (fabs(d) - Hu);
  return fma(xn, xn, y2) >= kFp64Gh4R2;
}

// The full FP64 dispatch, called (not inlined) from the rare near-centre
// paths so that its registers do not crowd the far-field loop.
__device__ __noinline__ double K_fp64_call(double x, double y) { return K_fp64(x, y); }
__device__ __noinline__ float K_mixed_call(float x, float y) { return K_mixed(x, y); }

// Record of the FP64 line sum as staged in LDS (one per (layer, line)), with
// everything the tiers and the queue drain read.  hw2, hw4: |x| at the
// middle of a thread's run above which all P points are in the second-order
// shifted and four-node regions.
struct alignas(16) Rec64 {
  double c, s;        // shifted centre, sqrt(ln2)/alpha_D
  double y, hs;       // Lorentz/Doppler ratio; step*s
  double am, ay;      // the line's weight A; A y/sqrt(pi)
  double hw2, hw4;    // sqrt(max(s2_r2_min64(y^2) - y^2, 0)), sqrt(max(35^2 - y^2, 0)), each + (P-1)/2 hs
};
static_assert(sizeof(Rec64) == 64, "four 16-byte LDS reads per record");
typedef __attribute__((address_space(3))) Rec64 lds_rec64;
// 4 workgroups per CU need under 40 KB of LDS each: 16 KB of records, 16 KB
// of FP64 partial sums, and per wavefront a queue of (record, lane) entries
// and a list of core points.
constexpr int kQueueW64 = 256;    // queue entries per wavefront (drained when more than kQueueW64 - 64 are pending)
constexpr int kCoreCap64 = 512;   // >= the 64 x 8 points one pass-A round can add

// FP64 line sum.  Same tiling as the mixed kernel: per record and thread
// the middle x of the thread's P points decides, for all of them at once,
// between the second-order shifted form (faddeeva.h, error <= 1e-11), the
// 4-node form (|z| >= 35) and the near-centre queue.  The wavefront drains
// its queue when it could not take one more record, in two dense passes:
// pass A evaluates every queued point with the 10-node form and delivers it
// to the owner's sums with plain LDS read-modify-writes; the points it
// cannot hold (|z| < 7, the core) go to a list that pass B evaluates one
// point per lane with trapezoid + residue.
template <int P, bool kCount>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 3)))
void k_voigt_sum_fp64(const SumArgs a) {
  // every wavefront works alone on its own near range (k_far_series_fp64 has
  // summed the rest of the tile): its own LDS stage of kStage records, its
  // own queue and core list drained by itself, no workgroup barriers
  constexpr int kWaves = kBlock / 64;
  constexpr int kStage = 64;
  __shared__ Rec64 s_rec[kWaves][kStage];
  __shared__ uint16_t s_q[kWaves][kQueueW64];      // (record << 6) | lane
  __shared__ uint16_t s_ci[kWaves][kCoreCap64];    // (record << 9) | (lane << 3) | point
  __shared__ double s_acc[kBlock * P];
  __shared__ uint8_t s_claim[kWaves][64];          // claim byte per owner lane
  __shared__ uint8_t s_gmask[kWaves][kStage];      // phase 2: the groups that take a staged record

  // (layer group, tile slice), group-major so that an XCD holds neighbouring
  // tiles of one layer range
  const int64_t lw = xcd_remap(blockIdx.x, a.n_wg * a.layer_groups);
  const int grp = (int)(lw / a.n_wg);
  const int64_t wg = lw - (int64_t)grp * a.n_wg;
  const int k_begin = (int)((int64_t)a.n_layers * grp / a.layer_groups);
  const int k_end = (int)((int64_t)a.n_layers * (grp + 1) / a.layer_groups);
  const int64_t tile = wg / a.slices_per_tile;
  const int slice = (int)(wg - tile * a.slices_per_tile);
  const int64_t t0 = tile * a.ts;
  const int64_t t1 = min(t0 + a.ts, a.n_nu);
  const int64_t s0 = t0 + (int64_t)slice * (kBlock * P);
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  // a wavefront wholly past the end of a short (last) tile has no points
  if (s0 + (int64_t)wave * kWavePoints >= t1) return;
  const int64_t i0 = s0 + (int64_t)threadIdx.x * P;
  const double lo_b = __dsub_rn(grid_nu(a.nu_min, a.step, t0), a.cutoff);
  const double hi_b = __dadd_rn(grid_nu(a.nu_min, a.step, t1 - 1), a.cutoff);
  const int32_t rlo = a.lo[tile], rhi = a.hi[tile];
  const double nu_m = __dadd_rn(grid_nu(a.nu_min, a.step, i0), 0.5 * (P - 1) * a.step);
  const bool live = i0 < t1;
  const double* rs = static_cast<const double*>(a.rec.s);
  const double* ry = static_cast<const double*>(a.rec.y);
  const double* ra = static_cast<const double*>(a.rec.a);
  const double half = 0.5 * (P - 1);
  const double t0sq = kGh4T[0] * kGh4T[0], t1sq = kGh4T[1] * kGh4T[1];
  const int32_t nlo = min(max(a.wlo[wg * kWaves + wave], rlo), rhi);
  const int32_t nhi = max(min(a.whi[wg * kWaves + wave], rhi), nlo);
  // s_acc[j][thread]: a wavefront's lanes touch consecutive words (no bank conflicts)
  double* const acc = s_acc + threadIdx.x;
  for (int j = 0; j < P; ++j) acc[j * kBlock] = 0.0;

  // column totals live in tau_part[grp], read and rewritten once per layer,
  // so that no registers hold them through the record loops and the drains
  double* const tau_acc = static_cast<double*>(a.tau_part) + (int64_t)grp * a.n_nu;
  uint64_t n_evals = 0, n_queued = 0, n_core = 0;   // kCount: point evaluations, queued pairs, core points
  // the wavefront's groups of kGroupPoints points (8 lanes each) and their
  // own near ranges (records within kSeriesRatio Hg of the group's middle)
  const int64_t w_first = s0 + (int64_t)wave * kWavePoints;
  const double Hg = 0.5 * kGroupPoints * a.step;
  const int32_t* const gl = a.glo + (wg * kWaves + wave) * kGroupsPerWave;
  const int32_t* const gh = a.ghi + (wg * kWaves + wave) * kGroupsPerWave;
  const uint32_t my_group = (uint32_t)lane >> 3;
  const int64_t g_first = w_first + (int64_t)my_group * kGroupPoints;
  const bool live_g = g_first < t1;
  const uint64_t g_pts = (uint64_t)max((int64_t)0, min((int64_t)kGroupPoints, t1 - g_first));
  const int32_t my_lo = min(max(gl[my_group], nlo), nhi);
  const int32_t my_hi = max(min(gh[my_group], nhi), my_lo);
  int32_t n_own = 0;   // the longest of the eight (scalar loads)
#pragma unroll
  for (int g = 0; g < kGroupsPerWave; ++g) {
    const int32_t g_lo = min(max(gl[g], nlo), nhi), g_hi = max(min(gh[g], nhi), g_lo);
    n_own = max(n_own, g_hi - g_lo);
  }
  const volatile lds_u32x4* const rb = (const volatile lds_u32x4*)(lds_rec64*)s_rec[wave];
  auto f64 = [](unsigned lo, unsigned hi) { return __hiloint2double((int)hi, (int)lo); };

  for (int k = k_begin; k < k_end; ++k) {
    double lay[P];
#pragma unroll
    for (int j = 0; j < P; ++j) lay[j] = 0.0;
    const int64_t kb = (int64_t)k * a.n_lines;
    int nq = 0;   // queue entries not yet drained (uniform)
    // the staged record of line r of this layer: the tier thresholds, and y
    // and A for the drain; a record outside the tile window (or skipped)
    // adds 0 and is never queued
    auto stage = [&](const int slot, const int32_t r, const bool valid) {
      double c = 0.0, sv = 0.0, y = 1.0e3, am = 0.0;
      bool inside = false;
      if (valid) {
        const double rc_ = a.rec.c[kb + r];
        inside = rc_ >= lo_b && rc_ <= hi_b;
        if (inside) { c = rc_; sv = rs[kb + r]; y = ry[kb + r]; am = ra[kb + r]; }
      }
      Rec64 R;
      const double y2 = y * y, hs = a.step * sv;
      R.c = c; R.s = sv; R.hs = hs;
      R.hw2 = fma(half, hs, sqrt(fmax(s2_r2_min64(y2) - y2, 0.0)));
      R.hw4 = fma(half, hs, sqrt(fmax(kFp64Gh4R2 - y2, 0.0)));
      R.y = y; R.am = am; R.ay = am * y * kInvSqrtPi;
      s_rec[wave][slot] = R;
      return inside;
    };
    // Records are read from LDS as 16-byte pieces: (c, s) and (y, hs) one
    // iteration ahead, (A, ay) and (hw2, hw4) at the record.  The group
    // level summed every (group, record) pair outside the group's own near
    // range where the four-node form holds at all its points, so the
    // one-node form (|z| >= 630 for narrow lines) is never reached here.
    auto record = [&](const int q, const u32x4 h0, const u32x4 h1,
                      const volatile lds_u32x4* const more, const bool direct) {
      const double rc = f64(h0.x, h0.y), rsv = f64(h0.z, h0.w);
      const double y = f64(h1.x, h1.y), hs = f64(h1.z, h1.w);
      const double xm = (nu_m - rc) * rsv;
      const double ax = fabs(xm);
      const u32x4 h2 = more[0], h3 = more[1];
      const double y2 = y * y, ay = direct ? f64(h2.z, h2.w) : 0.0;
      const double hw2 = f64(h3.x, h3.y), hw4 = f64(h3.z, h3.w);
      // One tier for the whole wavefront, the lowest any lane needs (each
      // form holds wherever the ones above it do): no lane pays for two
      // tiers.  Lanes that need the queue run the four-node form at a safe x
      // with weight 0 and queue the pair.
      if (__ballot(direct && ax < hw2) == 0) {
        // second-order shifted form (faddeeva.h), A K = ay r / D with
        // r = x^2 + y^2, D = r^2 - 3/2 r + 2 y^2 - 3/2 >= 6.9e7; the quotients
        // by batch inversion in two batches of four (a product of four D
        // ~ |z|^16 stays inside the FP64 range for |z| < 1e19)
        const double cc = fma(2.0, y2, -1.5);
        const double x0 = fma(-half, hs, xm);
        double rr[P], D[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const double x = fma((double)j, hs, x0);
          rr[j] = fma(x, x, y2);
          D[j] = fma(rr[j], rr[j] - 1.5, cc);
        }
#pragma unroll
        for (int h = 0; h < P; h += 4) {
          const double pa = D[h] * D[h + 1], pb = D[h + 2] * D[h + 3];
          const double r = ay * rcp64_1nr(pa * pb);
          const double ra_ = r * pb, rb_ = r * pa;
          lay[h] = fma(ra_ * D[h + 1], rr[h], lay[h]);
          lay[h + 1] = fma(ra_ * D[h], rr[h + 1], lay[h + 1]);
          lay[h + 2] = fma(rb_ * D[h + 3], rr[h + 2], lay[h + 2]);
          lay[h + 3] = fma(rb_ * D[h + 2], rr[h + 3], lay[h + 3]);
        }
        return;
      }
      // four-node form (|z| >= 35 at every point of the run; at the safe
      // middle x = hw4 for lanes whose run reaches nearer, with weight 0);
      // A (2y/pi) W = ay (2/sqrt(pi)) W; the two points of a pair share one
      // v_rcp_f64 (a product of two D0 D1 ~ |z|^16)
      const bool near = direct && ax < hw4;
      {
        const double fy2 = 4.0 * y2;
        const double ayw = near ? 0.0 : ay;
        const double w0 = ayw * (2.0 * kInvSqrtPi * kGh4W[0]);
        const double w1 = ayw * (2.0 * kInvSqrtPi * kGh4W[1]);
        const double x0 = fma(-half, hs, near ? hw4 : xm);
#pragma unroll
        for (int j = 0; j < P; j += 2) {
          double Np[2], Dp[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const double x = fma((double)(j + u), hs, x0);
            const double x2 = x * x;
            const double r = x2 + y2, xmy = x2 - y2, fxy = x2 * fy2;
            const double u0 = xmy - t0sq, u1 = xmy - t1sq;
            const double D0 = fma(u0, u0, fxy), D1 = fma(u1, u1, fxy);
            Np[u] = w0 * (r + t0sq) * D1 + w1 * (r + t1sq) * D0;
            Dp[u] = D0 * D1;
          }
          const double rp = rcp64_1nr(Dp[0] * Dp[1]);
          lay[j] = fma(Np[0] * Dp[1], rp, lay[j]);
          lay[j + 1] = fma(Np[1] * Dp[0], rp, lay[j + 1]);
        }
      }
      // append the queued lanes' (stage slot, lane) entries; nq advances in
      // every lane (the loops drain before the queue could overflow)
      const bool queued = near && live;
      const uint64_t msk = __ballot(queued);
      if (msk) {   // uniform
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(msk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)msk, 0u));
        if (queued) s_q[wave][nq + (int)rank] = (uint16_t)((q << 6) | lane);
        nq += (int)__popcll(msk);
      }
    };
    // The queue, drained by this wavefront in two dense passes.  Pass A: one
    // entry per lane, its P points at their exact grid values in the
    // 10-node form (<= 1e-11 for |z| >= 7, y >= 1e-8); core points get
    // weight 0 and go to the core list.  Pass B: the core list, one point
    // per lane, trapezoid + residue.
    auto drain = [&](const int n) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the queue writes
      __builtin_amdgcn_wave_barrier();
      int ncore = 0;   // uniform
      auto core_pass = [&](const int m) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the list writes
        __builtin_amdgcn_wave_barrier();
        for (int b = lane; b < m; b += 64) {
          const uint32_t ci = s_ci[wave][b];
          const int qq = (int)(ci >> 9), t = (int)((ci >> 3) & 63), j = (int)(ci & (P - 1));
          const u32x4 h0 = rb[4 * qq], h1 = rb[4 * qq + 1], h2 = rb[4 * qq + 2];   // (c, s), (y, hs), (A, ay)
          const double y = f64(h1.x, h1.y), am = f64(h2.x, h2.y);
          const double x = (grid_nu(a.nu_min, a.step, s0 + (int64_t)(wave * 64 + t) * P + j) -
                            f64(h0.x, h0.y)) * f64(h0.z, h0.w);
          atomicAdd(&s_acc[j * kBlock + wave * 64 + t], am * K_cr64(x, y));
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // before the list is reused
        __builtin_amdgcn_wave_barrier();
      };
      for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        const bool has = e < n;
        const uint32_t qi = has ? s_q[wave][e] : 0u;
        const int qq = (int)(qi >> 6), t = (int)(qi & 63);
        const u32x4 h0 = rb[4 * qq], h1 = rb[4 * qq + 1], h2 = rb[4 * qq + 2];   // (c, s), (y, hs), (A, ay)
        const double rc = f64(h0.x, h0.y), rsv = f64(h0.z, h0.w);
        // A (2y/pi) = (2/sqrt(pi)) ay; idle lanes add 0
        const double cw = has ? 2.0 * kInvSqrtPi * f64(h2.z, h2.w) : 0.0;
        const double y = has ? f64(h1.x, h1.y) : 1.0;
        const double y2 = y * y;
        const double rcore = y >= kFp64Gh10Ymin ? kFp64Gh10R2 : kFp64Gh8R2;
        const int64_t ib = s0 + (int64_t)(wave * 64 + t) * P;
        uint32_t cmask = 0;
        double v[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const double x = (grid_nu(a.nu_min, a.step, ib + j) - rc) * rsv;
          const double x2 = x * x;
          const double r = x2 + y2, xmy = x2 - y2, fxy = 4.0 * x2 * y2;
          const bool core = has && r < rcore;
          // (a select: at a core point the node sum may be infinite)
          const double g = gh_sum64<5>(r, xmy, fxy, kGh10T, kGh10W);
          v[j] = core ? 0.0 : cw * g;
          cmask |= core ? (1u << j) : 0u;
        }
        // Deliver the P values to the owner's sums with plain LDS reads and
        // writes: in each round every pending lane writes its id to its
        // owner's claim byte, the lanes that read their own id back hold
        // distinct owners and update, the rest try again.
        bool pending = has;
        while (__ballot(pending)) {   // uniform
          if (pending) s_claim[wave][t] = (uint8_t)lane;
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          if (pending && s_claim[wave][t] == (uint8_t)lane) {
            double* const dst = s_acc + wave * 64 + t;
#pragma unroll
            for (int j = 0; j < P; ++j) dst[j * kBlock] += v[j];
            pending = false;
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
        // this lane's core points go to the list at an exclusive prefix sum
        // of the lanes' counts (bit planes of the 0..8 count, one ballot each)
        const uint32_t cnt8 = __builtin_popcount(cmask);
        uint32_t pre = 0, total = 0;
#pragma unroll
        for (int bit = 0; bit < 4; ++bit) {
          const uint64_t m = __ballot((cnt8 >> bit) & 1u);
          pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << bit;
          total += (uint32_t)__popcll(m) << bit;
        }
        if (total) {   // uniform
          if (ncore + (int)total > kCoreCap64) {
            core_pass(ncore);
            ncore = 0;
          }
          int off = ncore + (int)pre;
#pragma unroll
          for (int j = 0; j < P; ++j) {
            if (cmask & (1u << j)) {
              s_ci[wave][off] = (uint16_t)((uint32_t)qq << 9 | (uint32_t)t << 3 | (uint32_t)j);
              ++off;
            }
          }
          ncore += (int)total;
          if constexpr (kCount) n_core += total;
        }
      }
      if (ncore > 0) core_pass(ncore);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int j = 0; j < P; ++j) {
        lay[j] += acc[j * kBlock];
        acc[j * kBlock] = 0.0;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    };

    // Phase 1: every group's own near range, eight records per group per
    // stage (slots 8g .. 8g + 7, one loaded by each lane of the group); in
    // iteration i every lane evaluates its group's record 8g + i, so each
    // iteration holds eight records and no lane idles while its group has
    // records left.  Phase 2: the pairs the group level left to this kernel
    // (outside the group's own range the four-node form fails at some point
    // of the group: gh4_holds_within, the test k_far_series_fp64 makes, in
    // the same arithmetic), only where the series kernel counted some for
    // this (layer, wavefront): lane l tests record base + l of the
    // wavefront's near range for all eight groups, the records some group
    // needs are packed into the stage with that mask, and every lane takes
    // each, with weight 0 where its group is not in it.  One loop runs the
    // stages of both phases, so that staging, the record and the drain each
    // have a single call site.
    const uint32_t n_left = a.deferred[(int64_t)k * (a.n_wg * kWaves) + wg * kWaves + wave];
    int32_t p1 = 0;     // phase 1: offset of the next stage in each group's range
    int32_t p2 = nlo;   // phase 2: next record of the near range to test
    for (;;) {
      bool ph2;
      int slot;
      int32_t r;
      bool write, valid;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the previous stage's readers
      if (p1 < n_own) {   // uniform
        ph2 = false;
        slot = lane;      // 8 g + (lane & 7)
        r = my_lo + p1 + (lane & 7);
        valid = r < my_hi;
        write = true;
      } else if (n_left != 0 && p2 < nhi) {   // uniform
        ph2 = true;
        r = p2 + lane;
        uint32_t gmask = 0;
        if (r < nhi) {
          const double c = a.rec.c[kb + r];
          if (c >= lo_b && c <= hi_b) {
            const double sv = rs[kb + r], y = ry[kb + r];
#pragma unroll
            for (int g = 0; g < kGroupsPerWave; ++g) {
              const int32_t g_lo = min(max(gl[g], nlo), nhi), g_hi = max(min(gh[g], nhi), g_lo);
              const double nu_g = group_middle(a.nu_min, a.step, w_first + (int64_t)g * kGroupPoints);
              const bool left = (r < g_lo || r >= g_hi) && !gh4_holds_within(nu_g - c, sv, y * y, Hg);
              gmask |= left ? (1u << g) : 0u;
            }
          }
        }
        const uint64_t keep = __ballot(gmask != 0);
        slot = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(keep >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)keep, 0u));
        write = valid = gmask != 0;
        if (write) s_gmask[wave][slot] = (uint8_t)gmask;
      } else {
        break;
      }
      const bool inside = write && stage(slot, r, valid);
      if constexpr (kCount) n_evals += !ph2 && inside && live_g ? g_pts : 0u;
      const int n_it = ph2 ? (int)__popcll(__ballot(write))
                           : __builtin_amdgcn_readfirstlane(min(8, n_own - p1));
      if (ph2) p2 += kStage; else p1 += 8;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // the queue is drained when it could not take one more record, and at
      // the stage's end (it refers to the stage's slots)
      int it = 0;
      for (;;) {
        for (; it < n_it; ++it) {
          const int q = ph2 ? it : (int)(my_group * 8) + it;
          bool direct = true;
          if (ph2) {   // uniform
            const uint32_t gm = __builtin_amdgcn_readfirstlane((uint32_t)s_gmask[wave][it]);
            direct = (gm >> my_group) & 1u;
            if constexpr (kCount) n_evals += direct && live_g && (lane & 7) == 0 ? g_pts : 0u;
          }
          record(q, rb[4 * q], rb[4 * q + 1], rb + 4 * q + 2, direct);
          if (nq > kQueueW64 - 64) {   // uniform
            ++it;
            break;
          }
        }
        if (nq > 0) {   // uniform
          if constexpr (kCount) n_queued += (uint64_t)nq;
          drain(nq);
          nq = 0;
        }
        if (it >= n_it) break;
      }
    }
    const double T = a.t_k[k];
    bool bad = false;
    const double* const far_k = static_cast<const double*>(a.far_sum) + (int64_t)k * a.n_nu;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int64_t i = i0 + j;
      if (i < t1) {
        const double nu = grid_nu(a.nu_min, a.step, i);
        const double v = (lay[j] + far_k[i]) * (nu * tanh(0.5 * kC2 * nu / T));
        bad |= !isfinite(v);
        tau_acc[i] = (k > k_begin ? tau_acc[i] : 0.0) + v;
        if (a.tau_layers) static_cast<double*>(a.tau_layers)[(int64_t)k * a.n_nu + i] = v;
      }
    }
    if (bad) flag_od(a.err, k);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int64_t i = i0 + j;
    if (i < t1) {
      if (k_begin == k_end) tau_acc[i] = 0.0;   // no layers in this group
      if (a.layer_groups > 1) continue;   // k_tau_epilogue64 adds the groups
      const double t = tau_acc[i];
      for (int m = 0; m < a.n_mu; ++m)
        static_cast<double*>(a.trans)[(int64_t)m * a.n_nu + i] = exp(-t * a.inv_mu[m]);
      if (a.tau) static_cast<double*>(a.tau)[i] = t;
    }
  }
  if constexpr (kCount) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) n_evals += __shfl_xor(n_evals, off);
    if (lane == 0) {
      count_add(a.work, kWorkNearEvals, n_evals);
      count_add(a.work, kWorkQueued, n_queued);
      count_add(a.work, kWorkCorePoints, n_core);
    }
  }
}

// ----------------------------------------------------------------- mixed --

// Surrogate mode: a queued point goes to the network iff it lies in the
// trained domain and its optical-depth contribution bound A c K_ref(x, y)
// (R(nu, T) <= nu ~ c near the centre, K <= 1.6 K_ref over the domain) is
// below the gate; everything else keeps the exact FP32 dispatch.
__device__ __forceinline__ bool surrogate_point(float x, float y, float ac, float gate) {
  if (!(fabsf(x) <= mlp::kXMax)) return false;
  const float ly = __log10f(y);
  if (!(ly >= mlp::kLyMin && ly <= mlp::kLyMax)) return false;
  return ac * mlp::kref(x, y) <= gate;
}

// FP64 far-field series (FP64 mode): as k_far_series_mixed, two levels, with
// the second-order shifted form, whose FP64 region (error <= 1e-11) starts
// much nearer than the one-node form's:
//   A K = ay r / D, r = (xm + sh tau)^2 + y^2 = r0 + a1 tau + a2 tau^2,
//   D = r^2 - 3/2 r + 2 y^2 - 3/2 = d0 + d1 tau + ... + d4 tau^4,
// whose Taylor coefficients follow from D Q = ay r:
//   Q_n = (ay r_n + e1 Q_(n-1) + ... + e4 Q_(n-4)) / d0 ... with e_k = -d_k / d0.
// The poles lie at |tau| >= 1/0.345 (kSeriesRatio half-spans), so
// kSeriesTerms64 terms leave <= 1e-11 of each record's value.  A record is
// summed about a middle point of half-span Hu iff the form holds at all the
// points covered: with x_n = s (|nu_m - c| - Hu),
// (x_n^2 + y^2)^3 >= kFp64S2Coef (9 y^2 + 21/4) (the form's threshold,
// cubed); a record that fails at the slice level goes on to the wavefront
// level, and one that fails there is evaluated point by point with the full
// FP64 dispatch.
template <int P, bool kCount>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 3)))
void k_far_series_fp64(const SumArgs a) {
  static_assert(P == kPointsFp64 && 64 * P == kWavePoints && kBlock * P == kSlicePoints,
                "slice and wavefront geometry");
  constexpr int kWaves = kBlock / 64;
  constexpr int kSliceExCap = 1024;
  constexpr int kN = kSeriesTerms64;
  __shared__ int32_t s_ex[kWaves][kExCap];
  __shared__ int32_t s_sx[kSliceExCap];
  __shared__ int s_sxn;
  __shared__ double s_coef[kWaves][kN];
  // the slice series plus the wavefront level's exceptions at each thread's
  // own points (written and read by that thread only)
  __shared__ double s_exv[kBlock * P];
  const int64_t lw = xcd_remap(blockIdx.x, a.n_wg * a.n_layers);
  const int k = (int)(lw / a.n_wg);
  const int64_t wg = lw - (int64_t)k * a.n_wg;
  const int64_t tile = wg / a.slices_per_tile;
  const int slice = (int)(wg - tile * a.slices_per_tile);
  const int64_t t0 = tile * a.ts;
  const int64_t t1 = min(t0 + a.ts, a.n_nu);
  const int64_t s0 = t0 + (int64_t)slice * kSlicePoints;
  if (s0 >= t1) return;   // the whole workgroup: no points (before the barriers)
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const int64_t w_first = s0 + (int64_t)wave * kWavePoints;
  const bool wave_live = w_first < t1;
  const int64_t i0 = w_first + (int64_t)lane * P;
  const double nu_s = slice_middle(a.nu_min, a.step, s0);
  const double nu_w = wave_middle(a.nu_min, a.step, w_first);
  const double H = 0.5 * kWavePoints * a.step, Hs = 0.5 * kSlicePoints * a.step;
  const double lo_b = __dsub_rn(grid_nu(a.nu_min, a.step, t0), a.cutoff);
  const double hi_b = __dadd_rn(grid_nu(a.nu_min, a.step, t1 - 1), a.cutoff);
  const int32_t rlo = a.lo[tile], rhi = a.hi[tile];
  const int32_t slo = min(max(a.slo[wg], rlo), rhi);
  const int32_t shi = max(min(a.shi[wg], rhi), slo);
  const int32_t nlo = min(max(a.wlo[wg * kWaves + wave], slo), shi);
  const int32_t nhi = max(min(a.whi[wg * kWaves + wave], shi), nlo);
  const int64_t kb = (int64_t)k * a.n_lines;
  const double* __restrict__ rc = a.rec.c + kb;
  const double* __restrict__ rs = static_cast<const double*>(a.rec.s) + kb;
  const double* __restrict__ ry = static_cast<const double*>(a.rec.y) + kb;
  const double* __restrict__ ra = static_cast<const double*>(a.rec.a) + kb;
  // the form holds at every point within Hu of the middle point d away
  auto holds = [&](double d, double sv, double y2, double Hu) {
    const double xn = sv * (fabs(d) - Hu);
    const double rn = fma(xn, xn, y2);
    return rn * rn * rn >= kFp64S2Coef * fma(9.0, y2, 5.25);
  };
  // adds the Taylor coefficients of (n0 + n1 tau + n2 tau^2) / (d0 + ... + d4 tau^4)
  // to C: D Q = N term by term
  auto rational = [&](double (&C)[kN], double n0, double n1, double n2, double d0, double d1,
                      double d2, double d3, double d4, int nt) {
    const double inv = rcp64_1nr(d0);
    const double e1 = -d1 * inv, e2 = -d2 * inv, e3 = -d3 * inv, e4 = -d4 * inv;
    double q0 = n0 * inv;
    double q1 = fma(e1, q0, n1 * inv);
    double q2 = fma(e1, q1, fma(e2, q0, n2 * inv));
    double q3 = fma(e1, q2, fma(e2, q1, e3 * q0));
    C[0] += q0;
    C[1] += q1;
    C[2] += q2;
    C[3] += q3;
#pragma unroll
    for (int n = 4; n < kN; ++n) {
      if ((n == 12 || n == 14 || n == 16 || n == 18 || n == 20 || n == 24) && n >= nt) break;
      const double qn = fma(e1, q3, fma(e2, q2, fma(e3, q1, e4 * q0)));
      C[n] += qn;
      q0 = q1; q1 = q2; q2 = q3; q3 = qn;
    }
  };
  // second-order shifted form about a middle point d = nu_m - c away:
  // ay r / (r^2 - 3/2 r + 2 y^2 - 3/2), r = X + y^2, X = x^2 = X0 + a1 tau + a2 tau^2
  auto series = [&](double (&C)[kN], double d, double sv, double y, double am, double Hu, int nt) {
    const double y2 = y * y;
    const double xm = d * sv, sh = sv * Hu;
    const double a1 = 2.0 * xm * sh, a2 = sh * sh;
    const double r0 = fma(xm, xm, y2);
    const double t = fma(2.0, r0, -1.5);
    const double g = am * y * kInvSqrtPi;
    rational(C, g * r0, g * a1, g * a2, fma(r0, r0 - 1.5, fma(2.0, y2, -1.5)), t * a1,
             fma(a1, a1, t * a2), 2.0 * a1 * a2, a2 * a2, nt);
  };
  // four-node form (|z| >= 35, faddeeva.h) as two pairs of poles, each
  // w_k (X + b) / (X^2 + (2 y^2 - 2 t_k^2) X + b^2), b = y^2 + t_k^2
  auto series_gh4 = [&](double (&C)[kN], double d, double sv, double y, double am, double Hu,
                        int nt) {
    const double y2 = y * y;
    const double xm = d * sv, sh = sv * Hu;
    const double a1 = 2.0 * xm * sh, a2 = sh * sh;
    const double X0 = xm * xm;
    const double ay2 = am * y * kInvSqrtPi * (2.0 * kInvSqrtPi);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const double t2 = kGh4T[p] * kGh4T[p], w = ay2 * kGh4W[p];
      const double b = y2 + t2, c = 2.0 * (y2 - t2);
      const double t = fma(2.0, X0, c);
      rational(C, w * (X0 + b), w * a1, w * a2, fma(X0, X0 + c, b * b), t * a1,
               fma(a1, a1, t * a2), 2.0 * a1 * a2, a2 * a2, nt);
    }
  };
  // the term count of a round (wavefront-uniform): the fewest terms whose
  // truncation stays <= 1e-11 for every record the round sums, each at R =
  // |nu_m - c| / Hu unit half-spans (>= kSeriesRatio) from the middle point:
  // 28 terms from R = 3, 24 from 3.5, 20 from 4.5, 18 from 5, 16 from 6, 14
  // from 8, 12 from 12 (the second-order form and the four-node pole pairs
  // alike; tests/test_far_series.py)
  auto series_terms64 = [&](bool use, double R) -> int {
    if (__ballot(use && R < 3.5)) return kN;
    if (__ballot(use && R < 4.5)) return 24;
    if (__ballot(use && R < 5.0)) return 20;
    if (__ballot(use && R < 6.0)) return 18;
    if (__ballot(use && R < 8.0)) return 16;
    if (__ballot(use && R < 12.0)) return 14;
    return 12;
  };
  static_assert(kN == 28, "series_terms64's table ends at 28 terms");
  uint64_t n_terms = 0, n_terms2 = 0;   // kCount: terms summed (slice, wavefront level)
  auto holds_gh4 = [&](double d, double sv, double y2, double Hu) {
    return gh4_holds_within(d, sv, y2, Hu);
  };
  if (threadIdx.x == 0) s_sxn = 0;
  double* const exv = s_exv + threadIdx.x * P;
  __syncthreads();

  // ---- slice level: records outside [slo, shi), one per thread
  {
    double C[kN];
#pragma unroll
    for (int n = 0; n < kN; ++n) C[n] = 0.0;
    const int32_t n_below = slo - rlo;
    const int32_t n_sl = n_below + (rhi - shi);
    auto sl_index = [&](int32_t v) -> int64_t {
      return v < n_below ? (int64_t)rlo + v : (int64_t)shi + (v - n_below);
    };
    double nc = __builtin_nan(""), ns = 0.0, ny = 1.0, na = 0.0;
    uint64_t n_sser = 0, n_sexc = 0;   // kCount
    if ((int)threadIdx.x < n_sl) {
      const int64_t i = sl_index(threadIdx.x);
      nc = rc[i]; ns = rs[i]; ny = ry[i]; na = ra[i];
    }
    const int32_t n_it = (n_sl + kBlock - 1) / kBlock;
    for (int32_t it = 0; it < n_it; ++it) {
      const double c = nc, sv = ns, y = ny, am = na;
      const int32_t v = it * kBlock + (int)threadIdx.x;
      if (v + kBlock < n_sl) {
        const int64_t i = sl_index(v + kBlock);
        nc = rc[i]; ns = rs[i]; ny = ry[i]; na = ra[i];
      }
      const double d = nu_s - c;
      const bool inw = (v < n_sl) & (c >= lo_b) & (c <= hi_b);
      const bool ser = inw && holds(d, sv, y * y, Hs);
      const int nt = series_terms64(ser, fabs(d) / Hs);
      if (ser) series(C, d, sv, y, am, Hs, nt);
      if constexpr (kCount) n_terms += ser ? (uint64_t)nt : 0u;
      if (inw & !ser) {   // on to the wavefront level
        const int pos = atomicAdd(&s_sxn, 1);
        if (pos < kSliceExCap) s_sx[pos] = (int32_t)sl_index(v);
      }
      if constexpr (kCount) {
        n_sser += (uint64_t)__popcll(__ballot(ser));
        n_sexc += (uint64_t)__popcll(__ballot(inw & !ser));
      }
    }
    if constexpr (kCount) {
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) n_terms += __shfl_xor(n_terms, off);
      if (lane == 0) {
        count_add(a.work, kWorkSliceSeries, n_sser);
        count_add(a.work, kWorkSliceExceptions, n_sexc);
        count_add(a.work, kWorkSeriesTerms, n_terms);
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
      for (int n = 0; n < kN; ++n) C[n] += __shfl_xor(C[n], off);
    }
    if (lane == 0) {
#pragma unroll
      for (int n = 0; n < kN; ++n) s_coef[wave][n] = C[n];
    }
  }
  __syncthreads();
  if (!wave_live) return;   // the last barrier is behind every wavefront

  // the slice series goes to this thread's LDS slots at once (no registers
  // held through the wavefront loop); the wavefront level's exceptions add to
  // it there, and far_sum is written once at the end
  double* const out = static_cast<double*>(a.far_sum) + (int64_t)k * a.n_nu;
  {
    double T[kN];
#pragma unroll
    for (int n = 0; n < kN; ++n) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) t += s_coef[w][n];
      T[n] = t;
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const double t = ((double)(wave * kWavePoints + lane * P + j) - 0.5 * (double)(kSlicePoints - 1)) *
                       (2.0 / (double)kSlicePoints);
      double f = T[kN - 1];
#pragma unroll
      for (int n = kN - 2; n >= 0; --n) f = fma(f, t, T[n]);
      exv[j] = f;
    }
  }

  // ---- wavefront level
  double C[kN];
#pragma unroll
  for (int n = 0; n < kN; ++n) C[n] = 0.0;
  const int n_sx = __builtin_amdgcn_readfirstlane(s_sxn);
  const bool rescan = n_sx > kSliceExCap;
  // after an overflow the rescan (nD) covers every listed record too: the
  // list is not read then, or its records would be summed twice
  const int32_t nA = nlo - slo, nB = shi - nhi, nC = rescan ? 0 : n_sx;
  const int32_t nD = rescan ? (slo - rlo) + (rhi - shi) : 0;
  const int32_t n_far = nA + nB + nC + nD;
  auto far_index = [&](int32_t v) -> int64_t {
    if (v < nA) return (int64_t)slo + v;
    v -= nA;
    if (v < nB) return (int64_t)nhi + v;
    v -= nB;
    if (v < nC) return (int64_t)s_sx[v];
    v -= nC;
    return v < slo - rlo ? (int64_t)rlo + v : (int64_t)shi + (v - (slo - rlo));
  };
  // records of s_ex[wave][0, n) at every lane's own points, full dispatch
  auto far_exceptions = [&](const int n) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the lanes' list writes
    for (int e = 0; e < n; ++e) {
      const int32_t idx = __builtin_amdgcn_readfirstlane(s_ex[wave][e]);
      const double c = rc[idx], sv = rs[idx], y = ry[idx], am = ra[idx];
      // four-node form where |z| >= 35 (faddeeva.h), else the full dispatch
      const double y2 = y * y, fy2 = 4.0 * y2;
      const double w0 = am * y * kInvSqrtPi * (2.0 * kInvSqrtPi * kGh4W[0]);
      const double w1 = am * y * kInvSqrtPi * (2.0 * kInvSqrtPi * kGh4W[1]);
      const double t0sq = kGh4T[0] * kGh4T[0], t1sq = kGh4T[1] * kGh4T[1];
#pragma unroll 1
      for (int j = 0; j < P; ++j) {
        const double x = (grid_nu(a.nu_min, a.step, i0 + j) - c) * sv;
        const double x2 = x * x;
        double v;
        if (x2 + y2 >= kFp64Gh4R2) {
          const double r = x2 + y2, xmy = x2 - y2, fxy = x2 * fy2;
          const double u0 = xmy - t0sq, u1 = xmy - t1sq;
          const double D0 = fma(u0, u0, fxy), D1 = fma(u1, u1, fxy);
          const double N = w0 * (r + t0sq) * D1 + w1 * (r + t1sq) * D0;
          v = N * rcp64_1nr(D0 * D1);
        } else {
          v = am * K_fp64_call(x, y);
        }
        exv[j] += v;   // this lane's own point
      }
    }
  };
  int n_ex = 0;
  uint64_t n_wser = 0, n_wexc = 0, n_wlist = 0;   // kCount
  double nc = __builtin_nan(""), ns = 0.0, ny = 1.0, na = 0.0;
  if (lane < n_far) {
    const int64_t i = far_index(lane);
    nc = rc[i]; ns = rs[i]; ny = ry[i]; na = ra[i];
  }
  const int32_t n_it = (n_far + 63) >> 6;
  for (int32_t it = 0; it < n_it; ++it) {
    const double c = nc, sv = ns, y = ny, am = na;
    const int32_t v = it * 64 + lane;
    if (v + 64 < n_far) {
      const int64_t i = far_index(v + 64);
      nc = rc[i]; ns = rs[i]; ny = ry[i]; na = ra[i];
    }
    const double d = nu_w - c;
    const double y2 = y * y;
    bool inw = (v < n_far) & (c >= lo_b) & (c <= hi_b);
    // the rescan skips what the slice level summed
    if (v >= nA + nB + nC) inw = inw && !holds(nu_s - c, sv, y2, Hs);
    // the four-node form here: it holds much nearer than the second-order
    // form at FP64 accuracy, so broad lines need no point-by-point pass
    const bool ser = inw && holds_gh4(d, sv, y2, H);
    const int nt = series_terms64(ser, fabs(d) / H);
    if (ser) series_gh4(C, d, sv, y, am, H, nt);
    if constexpr (kCount) n_terms2 += ser ? (uint64_t)nt : 0u;
    const bool exc = inw & !ser;
    const uint64_t ex = __ballot(exc);
    if constexpr (kCount) {
      n_wser += (uint64_t)__popcll(__ballot(ser));
      n_wexc += (uint64_t)__popcll(ex);
      n_wlist += (uint64_t)__popcll(__ballot(inw & (v >= nA + nB) & (v < nA + nB + nC)));
    }
    if (ex) {   // uniform
      if (n_ex + 64 > kExCap) {
        far_exceptions(n_ex);
        n_ex = 0;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(ex >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ex, 0u));
      if (exc) s_ex[wave][n_ex + (int)rank] = (int32_t)far_index(v);
      n_ex += (int)__popcll(ex);
    }
  }
  far_exceptions(n_ex);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int n = 0; n < kN; ++n) C[n] += __shfl_xor(C[n], off);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const double t = ((double)(lane * P + j) - 0.5 * (double)(kWavePoints - 1)) *
                     (2.0 / (double)kWavePoints);
    double f = C[kN - 1];
#pragma unroll
    for (int n = kN - 2; n >= 0; --n) f = fma(f, t, C[n]);
    exv[j] += f;   // (this lane's own slots) far_sum is written after the group level
  }

  // ---- group level: the records of the wavefront's near range [nlo, nhi)
  // outside a group's own near range [g_lo, g_hi) (records within
  // kSeriesRatio Hg of its middle), about the group's middle (Hu = Hg = H/8).
  // Lane = (group lane/8, record slot lane%8): eight records per round, each
  // tested for all eight groups.  The four-node form's series (two pole
  // pairs, as at the wavefront level) where that form holds at every point of
  // the group (holds_gh4); a pair where it does not is left to the near
  // field, which applies this same test in the same arithmetic, so each
  // (group, record) pair is taken exactly once.
  uint32_t n_gdef = 0;                              // pairs left to the near field
  uint64_t n_gser = 0, n_gdef_pts = 0;              // kCount
  {
    const int grp = lane >> 3, r8 = lane & 7;
    const int64_t g_first = w_first + (int64_t)grp * kGroupPoints;
    const bool g_live = g_first < t1;
    const uint64_t g_pts = (uint64_t)max((int64_t)0, min((int64_t)kGroupPoints, t1 - g_first));
    const double nu_g = group_middle(a.nu_min, a.step, g_first);
    const double Hg = 0.5 * kGroupPoints * a.step;
    const int32_t* const gl = a.glo + (wg * kWaves + wave) * kGroupsPerWave;
    const int32_t* const gh = a.ghi + (wg * kWaves + wave) * kGroupsPerWave;
    const int32_t g_lo = min(max(gl[grp], nlo), nhi), g_hi = max(min(gh[grp], nhi), g_lo);
    double G[kN];
#pragma unroll
    for (int n = 0; n < kN; ++n) G[n] = 0.0;
    auto candidate = [&](int32_t r) { return g_live && r < nhi && (r < g_lo || r >= g_hi); };
    double gc = __builtin_nan(""), gs = 0.0, gy = 1.0, ga = 0.0;
    if (candidate(nlo + r8)) {
      gc = rc[nlo + r8]; gs = rs[nlo + r8]; gy = ry[nlo + r8]; ga = ra[nlo + r8];
    }
    const int32_t n_rounds = (nhi - nlo + 7) >> 3;
    for (int32_t it = 0; it < n_rounds; ++it) {
      const int32_t r = nlo + it * 8 + r8;
      const double c = gc, sv = gs, y = gy, am = ga;
      if (candidate(r + 8)) {   // the next round's record, ahead
        gc = rc[r + 8]; gs = rs[r + 8]; gy = ry[r + 8]; ga = ra[r + 8];
      }
      const bool act = candidate(r) && (c >= lo_b) && (c <= hi_b);
      const double d = nu_g - c;
      const bool valid = holds_gh4(d, sv, y * y, Hg);
      const bool ser = act && valid;
      const int nt = series_terms64(ser, fabs(d) / Hg);
      if (ser) series_gh4(G, d, sv, y, am, Hg, nt);
      n_gdef += act && !valid ? 1u : 0u;
      if constexpr (kCount) {
        n_terms2 += ser ? (uint64_t)nt : 0u;
        n_gser += ser ? 1u : 0u;
        n_gdef_pts += act && !valid ? g_pts : 0u;
      }
    }
    // totals over the group's eight lanes, in each of them
#pragma unroll
    for (int off = 4; off >= 1; off >>= 1) {
#pragma unroll
      for (int n = 0; n < kN; ++n) G[n] += __shfl_xor(G[n], off);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      // tau of point (lane % 8) * P + j of the group (exact in FP64)
      const double tg = ((double)(r8 * P + j) - 0.5 * (double)(kGroupPoints - 1)) *
                        (2.0 / (double)kGroupPoints);
      double f = G[kN - 1];
#pragma unroll
      for (int n = kN - 2; n >= 0; --n) f = fma(f, tg, G[n]);
      const int64_t i = i0 + j;
      if (i < t1) out[i] = exv[j] + f;
    }
  }
  // the pairs this wavefront left to the near field (it skips its scan for
  // them when there are none)
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) n_gdef += __shfl_xor(n_gdef, off);
  if (lane == 0) a.deferred[(int64_t)k * (a.n_wg * kWaves) + wg * kWaves + wave] = n_gdef;
  if constexpr (kCount) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      n_terms2 += __shfl_xor(n_terms2, off);
      n_gser += __shfl_xor(n_gser, off);
      n_gdef_pts += __shfl_xor(n_gdef_pts, off);
    }
    if (lane == 0) {
      count_add(a.work, kWorkGroupSeries, n_gser);
      count_add(a.work, kWorkGroupDeferred, n_gdef);
      count_add(a.work, kWorkGroupDeferredEvals, n_gdef_pts);
      const uint64_t w_pts = (uint64_t)min((int64_t)kWavePoints, t1 - w_first);
      count_add(a.work, kWorkWaveSeries, n_wser);
      count_add(a.work, kWorkSeriesTerms2, n_terms2);
      count_add(a.work, kWorkWaveExceptions, n_wexc);
      count_add(a.work, kWorkExceptionEvals, n_wexc * w_pts);
      count_add(a.work, kWorkWaveFromSliceList, n_wlist);
      if (wave == 0 && rescan) count_add(a.work, kWorkRescans, 1);
    }
  }
}

// Mixed-precision line sums: kMixedBlock threads x P points per workgroup.
// Records arrive precomputed (MixedRec); staging is a copy plus the tile's
// exact inclusion test.
constexpr int kMixedBlock = 256;

// Far-field series of the mixed line sums (see kSeriesRatio and
// kSlicePoints): one workgroup per (slice, layer).  Writes far_sum[k][i], the
// layer-k sum over every record of the tile outside the near range of the
// wavefront that owns point i, compensated; the line sum adds it before the
// radiation term.
//
// Each record's one-node form is summed as its Taylor series in the
// normalised offset tau from a middle point nu_m of half-span Hu:
//   ay / ((xm + sh tau)^2 + c1) = sum_n Q_n tau^n,  xm = s (nu_m - c), sh = s Hu,
//   Q_0 = ay / D0, Q_1 = al Q_0, Q_n = al Q_(n-1) + be Q_(n-2),
//   D0 = xm^2 + c1, al = -2 xm sh / D0, be = -sh^2 / D0.
// |nu_m - c| >= kSeriesRatio Hu puts both poles at |tau| >= 1/0.345 (0.345:
// y^2 < 3/2, |xm| >= 25 + sh), so the terms dropped after kSeriesTerms weigh
// <= 2.6e-7 of the record's value at any |tau| < 1.  Two levels: the records
// outside the slice's near range [slo, shi) about the slice's middle
// (Hu = Hs, all 256 threads, one record per thread), then those between the
// slice's and the wavefront's near ranges about the wavefront's middle
// (Hu = H, one record per lane).  The form must hold at every point the
// series covers: |nu_m - c| >= ds1 + (Hu - H).  A record that fails this at
// the slice level goes on to the wavefront level; one that fails it there is
// evaluated at every lane's own points with the near field's tiers.
template <int P, bool kCount>
__global__ __launch_bounds__(kMixedBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_far_series_mixed(const SumArgs a) {
  static_assert(P == kPointsMixed && 64 * P == kWavePoints && kMixedBlock * P == kSlicePoints,
                "slice and wavefront geometry");
  constexpr int kWaves = kMixedBlock / 64;
  constexpr int kSliceExCap = 1024;
  constexpr int kN = kSeriesTerms;
  // per wavefront: records nearer than their ds1, evaluated point by point
  // after (or, when the list fills, during) its loop
  __shared__ int32_t s_ex[kWaves][kExCap];
  // slice level: records nearer than ds1 + (Hs - H), for the wavefront level
  __shared__ int32_t s_sx[kSliceExCap];
  __shared__ int s_sxn;
  // slice-level coefficients per wavefront (Q_0 and Q_1 with compensation)
  __shared__ float s_coef[kWaves][kN + 2];
  // the lane's compensated layer sums, parked in LDS through the group level
  // (whose series coefficients need the registers)
  __shared__ float s_lay[2 * P][kMixedBlock];
  // (layer, slice), layer-major so that an XCD holds neighbouring tiles of a layer
  const int64_t lw = xcd_remap(blockIdx.x, a.n_wg * a.n_layers);
  const int k = (int)(lw / a.n_wg);
  const int64_t wg = lw - (int64_t)k * a.n_wg;
  const int64_t tile = wg / a.slices_per_tile;
  const int slice = (int)(wg - tile * a.slices_per_tile);
  const int64_t t0 = tile * a.ts;
  const int64_t t1 = min(t0 + a.ts, a.n_nu);
  const int64_t s0 = t0 + (int64_t)slice * kSlicePoints;
  if (s0 >= t1) return;   // the whole workgroup: no points (before the barriers)
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const int64_t w_first = s0 + (int64_t)wave * kWavePoints;
  const bool wave_live = w_first < t1;   // uniform across the wavefront
  const int64_t i0 = w_first + (int64_t)lane * P;
  const double nu_s = slice_middle(a.nu_min, a.step, s0);
  const double nu_w = wave_middle(a.nu_min, a.step, w_first);
  const double nu_m = __dadd_rn(grid_nu(a.nu_min, a.step, i0), 0.5 * (P - 1) * a.step);
  const double lo_b = __dsub_rn(grid_nu(a.nu_min, a.step, t0), a.cutoff);
  const double hi_b = __dadd_rn(grid_nu(a.nu_min, a.step, t1 - 1), a.cutoff);
  const int32_t rlo = a.lo[tile], rhi = a.hi[tile];
  const int32_t slo = min(max(a.slo[wg], rlo), rhi);
  const int32_t shi = max(min(a.shi[wg], rhi), slo);
  const int32_t nlo = min(max(a.wlo[wg * kWaves + wave], slo), shi);
  const int32_t nhi = max(min(a.whi[wg * kWaves + wave], shi), nlo);
  const float dS = 0.5f * (float)(kSlicePoints - kWavePoints) * (float)a.step;   // Hs - H
  const int64_t kb = (int64_t)k * a.n_lines;
  const MixedRec* __restrict__ lrec = a.rec.mix + kb;
  // the fields the series needs, prefetched one iteration ahead
  struct FarRec { double c; float s, c1, ay, ds1, dm; };
  auto far_load = [&](int64_t i) {
    const MixedRec* const p = lrec + i;
    FarRec f;
    f.c = p->c; f.s = p->s; f.c1 = p->c1; f.ay = p->ay; f.ds1 = p->ds1; f.dm = p->dm;
    return f;
  };
  const float Hw = wave_half_span32(a.step);   // the series' unit at the wavefront level is s Hw
  // adds the first nt terms of record m's series about a middle point at
  // offset d = nu_m - c (units of the level: sh scaled by the caller) to C
  // (Q_0, Q_1 into C[0], C[1]); nt is wavefront-uniform (series_terms32)
  auto series = [&](float (&C)[kN], const FarRec& m, float d, float sh, int nt) {
    const float xm = d * m.s;
    const float r = __builtin_amdgcn_rcpf(fmaf(xm, xm, m.c1));
    const float u = sh * r;
    const float al = -2.0f * xm * u, be = -sh * u;
    float qa = m.ay * r, qb = al * qa;
    C[0] += qa;
    C[1] += qb;
#pragma unroll
    for (int n = 2; n < kN; ++n) {
      if ((n == 8 || n == 10 || n == 12 || n == 14) && n >= nt) break;
      const float qn = fmaf(al, qb, be * qa);
      C[n] += qn;
      qa = qb;
      qb = qn;
    }
  };
  // the term count of a round (wavefront-uniform): the fewest terms whose
  // truncation stays <= 2.6e-7 for every record the round sums, each at
  // R = |nu_m - c| / Hu unit half-spans (>= kSeriesRatio) from the middle
  // point: 18 terms from R = 3, 14 from 4, 12 from 5, 10 from 7, 8 from 10
  // (the one-node and the middle-tier forms alike; tests/test_far_series.py)
  auto series_terms32 = [&](bool use, float R) -> int {
    if (__ballot(use && R < 4.0f)) return kN;
    if (__ballot(use && R < 5.0f)) return 14;
    if (__ballot(use && R < 7.0f)) return 12;
    if (__ballot(use && R < 10.0f)) return 10;
    return 8;
  };
  static_assert(kN == 18, "series_terms32's table ends at 18 terms");
  // 1 / Hu of the three levels (slice 4 Hw, wavefront Hw, group Hw / 8)
  const float inv_hw = 1.0f / Hw;
  // adds record m's middle-tier form (faddeeva.h: the second-order shifted
  // form for y^2 < kF32S2Y2, else the two-node form) about a middle point at
  // offset d (unit sh) to C.  Both forms are u = x^2 + A, f = (ay u + Cn) /
  // (u^2 + B) (A = y^2 - 3/4, B = 2 y^2 - 33/16, Cn = 3/4 ay, or A = y^2 -
  // 1/2, B = 2 y^2, Cn = ay; the near field's middle tier); with x = xm +
  // sh tau, u = u0 + u1 tau + u2 tau^2 and D = u^2 + B = sum_(i<=4) d_i tau^i,
  // and D Q = N gives Q_n = (N_n - d1 Q_(n-1) - d2 Q_(n-2) - d3 Q_(n-3) -
  // d4 Q_(n-4)) / d0 (N_n = 0 for n > 2).  Used at the group level only,
  // where the record is at least kSeriesRatio Hu away and the form holds at
  // every point: the poles then lie at |tau| >= 2.69 and kN terms keep the
  // truncation <= 1e-7 (FP32 arithmetic adds <= 1.1e-6 of the value); a
  // round whose nearest record is farther sums fewer (nt, series_terms32's
  // table, which holds for this form too).
  auto series_mid = [&](float (&C)[kN], const FarRec& m, float d, float sh, int nt) {
    const float xm = d * m.s;
    const bool s2 = m.c1 < kF32S2Y2 - 1.5f;
    const float A = s2 ? m.c1 + 0.75f : m.c1 + 1.0f;
    const float B = s2 ? fmaf(2.0f, m.c1, 0.9375f) : fmaf(2.0f, m.c1, 3.0f);
    const float Cn = s2 ? 0.75f * m.ay : m.ay;
    const float u0 = fmaf(xm, xm, A), u1 = 2.0f * xm * sh, u2 = sh * sh;
    const float r = __builtin_amdgcn_rcpf(fmaf(u0, u0, B));
    const float e1 = 2.0f * u0 * u1 * r, e2 = fmaf(u1, u1, 2.0f * u0 * u2) * r;
    const float e3 = 2.0f * u1 * u2 * r, e4 = u2 * u2 * r;
    float q0 = fmaf(m.ay, u0, Cn) * r;
    float q1 = fmaf(-e1, q0, m.ay * u1 * r);
    float q2 = fmaf(-e2, q0, fmaf(-e1, q1, m.ay * u2 * r));
    float q3 = -fmaf(e1, q2, fmaf(e2, q1, e3 * q0));
    C[0] += q0;
    C[1] += q1;
    C[2] += q2;
    C[3] += q3;
#pragma unroll
    for (int n = 4; n < kN; ++n) {
      if ((n == 8 || n == 10 || n == 12 || n == 14) && n >= nt) break;
      const float qn = -fmaf(e1, q3, fmaf(e2, q2, fmaf(e3, q1, e4 * q0)));
      C[n] += qn;
      q0 = q1;
      q1 = q2;
      q2 = q3;
      q3 = qn;
    }
  };
  if (threadIdx.x == 0) s_sxn = 0;
  __syncthreads();

  // kCount: terms summed by this lane's one-node series and by its
  // middle-tier group series (the FLOP they take)
  uint64_t n_terms = 0, n_terms2 = 0;
  auto count_terms = [&]() {
    if constexpr (kCount) {
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) n_terms += __shfl_xor(n_terms, off);
      if (lane == 0) count_add(a.work, kWorkSeriesTerms, n_terms);
      n_terms = 0;
    }
  };
  // ---- slice level: records outside [slo, shi), one per thread
  {
    float C[kN];
#pragma unroll
    for (int n = 0; n < kN; ++n) C[n] = 0.0f;
    float S0 = 0.0f, S0c = 0.0f, S1 = 0.0f, S1c = 0.0f;   // compensated Q_0, Q_1 sums
    const int32_t n_below = slo - rlo;
    const int32_t n_sl = n_below + (rhi - shi);
    auto sl_index = [&](int32_t v) -> int64_t {
      return v < n_below ? (int64_t)rlo + v : (int64_t)shi + (v - n_below);
    };
    FarRec nx;
    nx.c = __builtin_nan("");
    uint64_t n_sser = 0, n_sexc = 0;   // kCount
    if ((int)threadIdx.x < n_sl) nx = far_load(sl_index(threadIdx.x));
    const int32_t n_it = (n_sl + kMixedBlock - 1) / kMixedBlock;
    for (int32_t it = 0; it < n_it; ++it) {
      const FarRec m = nx;
      const int32_t v = it * kMixedBlock + (int)threadIdx.x;
      if (v + kMixedBlock < n_sl) nx = far_load(sl_index(v + kMixedBlock));
      const float d = (float)(nu_s - m.c);
      const bool inw = (v < n_sl) & (m.c >= lo_b) & (m.c <= hi_b);
      const bool ser = inw & (fabsf(d) >= m.ds1 + dS);
      const int nt = series_terms32(ser, fabsf(d) * (0.25f * inv_hw));
      if (ser) series(C, m, d, 4.0f * (m.s * Hw), nt);
      if constexpr (kCount) n_terms += ser ? (uint64_t)nt : 0u;
      if (inw & !ser) {   // rare: on to the wavefront level
        const int pos = atomicAdd(&s_sxn, 1);
        if (pos < kSliceExCap) s_sx[pos] = (int32_t)sl_index(v);
      }
      if constexpr (kCount) {
        n_sser += (uint64_t)__popcll(__ballot(ser));
        n_sexc += (uint64_t)__popcll(__ballot(inw & !ser));
      }
      if ((it & 15) == 15) {
        kahan_add(S0, S0c, C[0]);
        kahan_add(S1, S1c, C[1]);
        C[0] = C[1] = 0.0f;
      }
    }
    kahan_add(S0, S0c, C[0]);
    kahan_add(S1, S1c, C[1]);
    if constexpr (kCount) {
      if (lane == 0) {
        count_add(a.work, kWorkSliceSeries, n_sser);
        count_add(a.work, kWorkSliceExceptions, n_sexc);
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      S0 += __shfl_xor(S0, off);
      S0c += __shfl_xor(S0c, off);
      S1 += __shfl_xor(S1, off);
      S1c += __shfl_xor(S1c, off);
#pragma unroll
      for (int n = 2; n < kN; ++n) C[n] += __shfl_xor(C[n], off);
    }
    if (lane == 0) {
      s_coef[wave][0] = S0;
      s_coef[wave][1] = S0c;
      s_coef[wave][2] = S1;
      s_coef[wave][3] = S1c;
#pragma unroll
      for (int n = 2; n < kN; ++n) s_coef[wave][n + 2] = C[n];
    }
  }
  __syncthreads();
  count_terms();   // (the slice level's, before a wavefront past the tile end leaves)
  if (!wave_live) return;   // the last barrier is behind every wavefront

  float lay[P], layc[P];
#pragma unroll
  for (int j = 0; j < P; ++j) lay[j] = layc[j] = 0.0f;
  {
    // the slice series at this lane's points
    float T[kN + 2];
#pragma unroll
    for (int n = 0; n < kN + 2; ++n) {
      float t = 0.0f;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) t += s_coef[w][n];
      T[n] = t;
    }
    const float s1t = T[2] - T[3];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      // tau of point wave*kWavePoints + lane*P + j of the slice (exact in FP32)
      const float t = ((float)(wave * kWavePoints + lane * P + j) - 0.5f * (float)(kSlicePoints - 1)) *
                      (2.0f / (float)kSlicePoints);
      float f = T[kN + 1];
#pragma unroll
      for (int n = kN - 2; n >= 2; --n) f = fmaf(f, t, T[n + 2]);
      f = fmaf(f, t, s1t) * t;
      kahan_add(lay[j], layc[j], T[0]);
      kahan_add(lay[j], layc[j], f - T[1]);
    }
  }

  // ---- wavefront level: [slo, nlo) and [nhi, shi), then the slice level's
  // exceptions; if their list overflowed, the slice-level range again with
  // every record that passed there skipped
  float C[kN];
#pragma unroll
  for (int n = 0; n < kN; ++n) C[n] = 0.0f;
  float S0 = 0.0f, S0c = 0.0f, S1 = 0.0f, S1c = 0.0f;
  const int n_sx = __builtin_amdgcn_readfirstlane(s_sxn);
  const bool rescan = n_sx > kSliceExCap;
  // after an overflow the rescan (nD) covers every listed record too: the
  // list is not read then, or its records would be summed twice
  const int32_t nA = nlo - slo, nB = shi - nhi, nC = rescan ? 0 : n_sx;
  const int32_t nD = rescan ? (slo - rlo) + (rhi - shi) : 0;
  const int32_t n_far = nA + nB + nC + nD;
  auto far_index = [&](int32_t v) -> int64_t {
    if (v < nA) return (int64_t)slo + v;
    v -= nA;
    if (v < nB) return (int64_t)nhi + v;
    v -= nB;
    if (v < nC) return (int64_t)s_sx[v];
    v -= nC;
    return v < slo - rlo ? (int64_t)rlo + v : (int64_t)shi + (v - (slo - rlo));
  };
  // records of s_ex[wave][0, n) at every lane's own points (record index
  // uniform across the wavefront), into the compensated layer sum
  auto far_exceptions = [&](const int n) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the lanes' list writes
    for (int e = 0; e < n; ++e) {
      const int32_t idx = __builtin_amdgcn_readfirstlane(s_ex[wave][e]);
      const MixedRec R = lrec[idx];
      const float xm = (float)(nu_m - R.c) * R.s;
      float v[P];
      if (fabsf(xm) >= R.hw1) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const float x = fmaf((float)j - 0.5f * (float)(P - 1), R.hs, xm);
          v[j] = R.ay * __builtin_amdgcn_rcpf(fmaf(x, x, R.c1));
        }
      } else if (fabsf(xm) >= R.hw2) {
        const bool s2 = R.c1 < kF32S2Y2 - 1.5f;
        const float y2 = R.c1 + 1.5f, cc = fmaf(2.0f, R.c1, 1.5f);
        const float cw = R.c1 + 1.0f, d2 = fmaf(2.0f, R.c1, 3.0f);
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const float x = fmaf((float)j - 0.5f * (float)(P - 1), R.hs, xm);
          if (s2) {   // second-order shifted form
            const float r = fmaf(x, x, y2);
            v[j] = R.ay * r * __builtin_amdgcn_rcpf(fmaf(r, r - 1.5f, cc));
          } else {    // two-node form
            const float w = fmaf(x, x, cw);
            v[j] = fmaf(R.ay, w, R.ay) * __builtin_amdgcn_rcpf(fmaf(w, w, d2));
          }
        }
      } else {
#pragma unroll
        for (int j = 0; j < P; ++j)
          v[j] = R.a * K_mixed_call(fmaf((float)j - 0.5f * (float)(P - 1), R.hs, xm), R.y);
      }
#pragma unroll
      for (int j = 0; j < P; ++j) kahan_add(lay[j], layc[j], v[j]);
    }
  };
  int n_ex = 0;
  uint64_t n_wser = 0, n_wexc = 0, n_wlist = 0;   // kCount
  FarRec nx;
  nx.c = __builtin_nan("");
  if (lane < n_far) nx = far_load(far_index(lane));
  const int32_t n_it = (n_far + 63) >> 6;
  for (int32_t it = 0; it < n_it; ++it) {
    const FarRec m = nx;
    const int32_t v = it * 64 + lane;
    if (v + 64 < n_far) nx = far_load(far_index(v + 64));
    const float d = (float)(nu_w - m.c);
    bool inw = (v < n_far) & (m.c >= lo_b) & (m.c <= hi_b);
    // the rescan skips what the slice level summed
    if (v >= nA + nB + nC) inw &= !(fabsf((float)(nu_s - m.c)) >= m.ds1 + dS);
    const bool ser = inw & (fabsf(d) >= m.ds1);
    const int nt = series_terms32(ser, fabsf(d) * inv_hw);
    if (ser) series(C, m, d, m.s * Hw, nt);
    if constexpr (kCount) n_terms += ser ? (uint64_t)nt : 0u;
    const bool exc = inw & !ser;
    const uint64_t ex = __ballot(exc);
    if constexpr (kCount) {
      n_wser += (uint64_t)__popcll(__ballot(ser));
      n_wexc += (uint64_t)__popcll(ex);
      n_wlist += (uint64_t)__popcll(__ballot(inw & (v >= nA + nB) & (v < nA + nB + nC)));
    }
    if (ex) {   // uniform
      if (n_ex + 64 > kExCap) {
        far_exceptions(n_ex);
        n_ex = 0;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(ex >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ex, 0u));
      if (exc) s_ex[wave][n_ex + (int)rank] = (int32_t)far_index(v);
      n_ex += (int)__popcll(ex);
    }
    if ((it & 15) == 15) {
      kahan_add(S0, S0c, C[0]);
      kahan_add(S1, S1c, C[1]);
      C[0] = C[1] = 0.0f;
    }
  }
  far_exceptions(n_ex);
  kahan_add(S0, S0c, C[0]);
  kahan_add(S1, S1c, C[1]);
  // totals over the wavefront's 64 lanes, in every lane
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    S0 += __shfl_xor(S0, off);
    S0c += __shfl_xor(S0c, off);
    S1 += __shfl_xor(S1, off);
    S1c += __shfl_xor(S1c, off);
#pragma unroll
    for (int n = 2; n < kN; ++n) C[n] += __shfl_xor(C[n], off);
  }
  const float s1t = S1 - S1c;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    // tau of point lane*P + j of the wavefront (exact in FP32)
    const float t = ((float)(lane * P + j) - 0.5f * (float)(kWavePoints - 1)) *
                    (2.0f / (float)kWavePoints);
    float f = C[kN - 1];
#pragma unroll
    for (int n = kN - 2; n >= 2; --n) f = fmaf(f, t, C[n]);
    f = fmaf(f, t, s1t) * t;
    kahan_add(lay[j], layc[j], S0);
    kahan_add(lay[j], layc[j], f - S0c);
  }

  // ---- group level: the records of the wavefront's near range [nlo, nhi)
  // outside a group's own near range [g_lo, g_hi), about the group's middle
  // (Hu = Hg = H/8, unit sh/8).  Lane = (group lane/8, record slot lane%8):
  // eight records per round, each tested for all eight groups.  The one-node
  // form's series where that form holds at every point of the group, else
  // the middle-tier form's where that holds; a pair for which neither does
  // is left to the near field, which applies these same tests.
  uint32_t n_gdef = 0;                             // pairs left to the near field
  uint64_t n_gser = 0, n_gmid = 0, n_gdef_pts = 0;   // kCount
#pragma unroll
  for (int j = 0; j < P; ++j) {
    s_lay[j][threadIdx.x] = lay[j];
    s_lay[P + j][threadIdx.x] = layc[j];
  }
  {
    const int grp = lane >> 3, r8 = lane & 7;
    const int64_t g_first = w_first + (int64_t)grp * kGroupPoints;
    const bool g_live = g_first < t1;
    const uint64_t g_pts = (uint64_t)max((int64_t)0, min((int64_t)kGroupPoints, t1 - g_first));
    const double nu_g = group_middle(a.nu_min, a.step, g_first);
    const int32_t* const gl = a.glo + (wg * kWaves + wave) * kGroupsPerWave;
    const int32_t* const gh = a.ghi + (wg * kWaves + wave) * kGroupsPerWave;
    const int32_t g_lo = min(max(gl[grp], nlo), nhi), g_hi = max(min(gh[grp], nhi), g_lo);
    const GroupTest gt = group_test(a.nu_min, a.step, w_first);
    const float goff = group_offset(gt, grp);
    float G[kN];
#pragma unroll
    for (int n = 0; n < kN; ++n) G[n] = 0.0f;
    float G0 = 0.0f, G0c = 0.0f, G1 = 0.0f, G1c = 0.0f;
    auto candidate = [&](int32_t r) { return g_live && r < nhi && (r < g_lo || r >= g_hi); };
    FarRec gx;
    gx.c = __builtin_nan("");
    gx.s = gx.c1 = gx.ay = gx.ds1 = gx.dm = 0.0f;
    if (candidate(nlo + r8)) gx = far_load(nlo + r8);
    const int32_t n_rounds = (nhi - nlo + 7) >> 3;
    for (int32_t it = 0; it < n_rounds; ++it) {
      const int32_t r = nlo + it * 8 + r8;
      const FarRec m = gx;
      if (candidate(r + 8)) gx = far_load(r + 8);   // the next round's record, ahead
      const bool act = candidate(r) && (m.c >= lo_b) && (m.c <= hi_b);
      const float d = (float)(nu_g - m.c);
      const float d0 = group_d0(gt, m.c);
      const bool valid = group_one_node_holds(d0, group_thr(gt, m.ds1), goff);
      const bool ser = act && valid;
      const int nt = series_terms32(ser, fabsf(d) * (8.0f * inv_hw));
      if (ser) series(G, m, d, 0.125f * (m.s * Hw), nt);
      if constexpr (kCount) n_terms += ser ? (uint64_t)nt : 0u;
      const bool mid = act && !valid && group_one_node_holds(d0, group_mid_thr(gt, m.dm), goff);
      if (__ballot(mid)) {   // uniform
        const int ntm = series_terms32(mid, fabsf(d) * (8.0f * inv_hw));
        if (mid) series_mid(G, m, d, 0.125f * (m.s * Hw), ntm);
        if constexpr (kCount) n_terms2 += mid ? (uint64_t)ntm : 0u;
      }
      const bool left = act && !valid && !mid;
      n_gdef += left ? 1u : 0u;
      if constexpr (kCount) {
        n_gser += ser || mid ? 1u : 0u;
        n_gmid += mid ? 1u : 0u;
        n_gdef_pts += left ? g_pts : 0u;
      }
      if ((it & 15) == 15) {
        kahan_add(G0, G0c, G[0]);
        kahan_add(G1, G1c, G[1]);
        G[0] = G[1] = 0.0f;
      }
    }
    kahan_add(G0, G0c, G[0]);
    kahan_add(G1, G1c, G[1]);
    // totals over the group's eight lanes, in each of them
#pragma unroll
    for (int off = 4; off >= 1; off >>= 1) {
      G0 += __shfl_xor(G0, off);
      G0c += __shfl_xor(G0c, off);
      G1 += __shfl_xor(G1, off);
      G1c += __shfl_xor(G1c, off);
#pragma unroll
      for (int n = 2; n < kN; ++n) G[n] += __shfl_xor(G[n], off);
    }
    const float g1t = G1 - G1c;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      // tau of point (lane % 8) * P + j of the group (exact in FP32)
      const float t = ((float)(r8 * P + j) - 0.5f * (float)(kGroupPoints - 1)) *
                      (2.0f / (float)kGroupPoints);
      float f = G[kN - 1];
#pragma unroll
      for (int n = kN - 2; n >= 2; --n) f = fmaf(f, t, G[n]);
      f = fmaf(f, t, g1t) * t;
      lay[j] = s_lay[j][threadIdx.x];   // (the lane's own words: no barrier)
      layc[j] = s_lay[P + j][threadIdx.x];
      kahan_add(lay[j], layc[j], G0);
      kahan_add(lay[j], layc[j], f - G0c);
    }
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int64_t i = i0 + j;
    if (i < t1) static_cast<float*>(a.far_sum)[(int64_t)k * a.n_nu + i] = lay[j] - layc[j];
  }
  // the pairs this wavefront left to the near field (it skips its scan for
  // them when there are none)
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) n_gdef += __shfl_xor(n_gdef, off);
  if (lane == 0) a.deferred[(int64_t)k * (a.n_wg * kWaves) + wg * kWaves + wave] = n_gdef;
  if constexpr (kCount) {
    if (lane == 0) {
      const uint64_t w_pts = (uint64_t)min((int64_t)kWavePoints, t1 - w_first);
      count_add(a.work, kWorkWaveSeries, n_wser);
      count_add(a.work, kWorkWaveExceptions, n_wexc);
      count_add(a.work, kWorkExceptionEvals, n_wexc * w_pts);
      count_add(a.work, kWorkWaveFromSliceList, n_wlist);
      if (wave == 0 && rescan) count_add(a.work, kWorkRescans, 1);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      n_gser += __shfl_xor(n_gser, off);
      n_gmid += __shfl_xor(n_gmid, off);
      n_gdef_pts += __shfl_xor(n_gdef_pts, off);
      n_terms2 += __shfl_xor(n_terms2, off);
    }
    if (lane == 0) {
      count_add(a.work, kWorkGroupSeries, n_gser);
      count_add(a.work, kWorkGroupMiddle, n_gmid);
      count_add(a.work, kWorkSeriesTerms2, n_terms2);
    }
    count_terms();
    if (lane == 0) {
      count_add(a.work, kWorkGroupDeferred, n_gdef);
      count_add(a.work, kWorkGroupDeferredEvals, n_gdef_pts);
    }
  }
}

template <int P, bool kSurrogate, bool kCount>
__device__ __forceinline__ void voigt_sum_mixed_body(const SumArgs& a) {
  // Every wavefront works alone (no workgroup barriers): it streams its own
  // near range through its own LDS stage of 64 records, and (record, lane)
  // pairs whose P points are not all in one of the far-field regions go to
  // its own queue, drained by the same wavefront after each stage with the
  // full region dispatch; results come back through s_acc (LDS float
  // atomics) into the owning lane's partial sums.
  constexpr int kWaves = kMixedBlock / 64;
  constexpr int kStage = 64;                        // records per wavefront stage
  // phase 1 stages kStageG records of each group's own range (one per lane of the group)
  constexpr int kStageG = kStage / kGroupsPerWave;
  static_assert(kStageG * kGroupsPerWave == 64, "one record per lane per stage");
  __shared__ MixedRec s_rec[kWaves][kStage];
  // queue: x at the lane's first point, (record << 6) | lane
  __shared__ float s_qx[kWaves][kQueueW];
  __shared__ uint16_t s_qi[kWaves][kQueueW];
  __shared__ float s_acc[kAccStride * P];
  // queued points in the core (gh6_core_r2): (queue entry << 3) | point
  // (queue entry << 3) | point, overwritten in place by the point's value
  __shared__ uint32_t s_ci[kWaves][kSurrogate ? 1 : kCoreCap];
  // claim byte per owner lane (the drain's plain read-modify-write rounds)
  __shared__ uint8_t s_claim[kWaves][64];

  // (layer group, tile slice), group-major so that an XCD holds neighbouring
  // tiles of one layer range
  const int64_t lw = xcd_remap(blockIdx.x, a.n_wg * a.layer_groups);
  const int grp = (int)(lw / a.n_wg);
  const int64_t wg = lw - (int64_t)grp * a.n_wg;
  const int k_begin = (int)((int64_t)a.n_layers * grp / a.layer_groups);
  const int k_end = (int)((int64_t)a.n_layers * (grp + 1) / a.layer_groups);
  const int64_t tile = wg / a.slices_per_tile;
  const int slice = (int)(wg - tile * a.slices_per_tile);
  const int64_t t0 = tile * a.ts;
  const int64_t t1 = min(t0 + a.ts, a.n_nu);
  const int64_t s0 = t0 + (int64_t)slice * (kMixedBlock * P);   // first point of the slice
  // the wavefront index read into a scalar register, so that the per-wave
  // tables (near and group ranges) are read with scalar loads
  const int lane = (int)threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  // a wavefront wholly past the end of a short (last) tile has no points
  if (s0 + (int64_t)wave * kWavePoints >= t1) return;
  const int64_t i0 = s0 + (int64_t)threadIdx.x * P;
  const double lo_b = __dsub_rn(grid_nu(a.nu_min, a.step, t0), a.cutoff);
  const double hi_b = __dadd_rn(grid_nu(a.nu_min, a.step, t1 - 1), a.cutoff);
  const int32_t rlo = a.lo[tile], rhi = a.hi[tile];
  static_assert(P == kPointsMixed, "records carry hw for kPointsMixed points per thread");
  // middle of this thread's run of P points; point j sits at x_mid + (j - (P-1)/2) hs
  const double nu_f = grid_nu(a.nu_min, a.step, i0);   // this thread's first point
  const double nu_m = __dadd_rn(nu_f, 0.5 * (P - 1) * a.step);
  const bool live = i0 < t1;   // lanes wholly past the tile end never write
  const MixedRec* __restrict__ recs = a.rec.mix;
  // This wavefront's near range: records within kSeriesRatio H of its middle
  // point by nu0 (k_far_series_mixed has summed all others of the tile), and
  // in it each group's: records within kSeriesRatio Hg of the group's middle.
  // The series kernel has summed, per group, the records outside the group's
  // range where the one-node form holds at all of the group's points; per
  // lane this kernel evaluates the records in its group's range and those
  // the group series could not take, and skips a record no group needs.
  const int32_t nlo = min(max(a.wlo[wg * kWaves + wave], rlo), rhi);
  const int32_t nhi = max(min(a.whi[wg * kWaves + wave], rhi), nlo);
  const int32_t* const gl = a.glo + (wg * kWaves + wave) * kGroupsPerWave;
  const int32_t* const gh = a.ghi + (wg * kWaves + wave) * kGroupsPerWave;
  const int64_t w_first = s0 + (int64_t)wave * kWavePoints;
  const GroupTest gt = group_test(a.nu_min, a.step, w_first);
  const uint32_t my_group = (uint32_t)lane >> 3;
  // points of this lane (0 for a lane wholly past the tile end)
  const int pts_lane = (int)max((int64_t)0, min((int64_t)P, t1 - i0));
  // this wavefront's stage as an LDS-qualified pointer, so record reads stay ds_read_b128
  lds_mixed_rec* const s_rec_lds = (lds_mixed_rec*)s_rec[wave];
  // s_acc[j][thread]: a wavefront's lanes touch consecutive words
  float* const acc = s_acc + threadIdx.x;
#pragma unroll
  for (int j = 0; j < P; ++j) acc[j * kAccStride] = 0.0f;

  // column totals (compensated) live in tau_part[grp], read and rewritten
  // once per layer, so that no registers hold them through the record loops
  float2* const tau_acc = static_cast<float2*>(a.tau_part) + (int64_t)grp * a.n_nu;
  // kCount: point evaluations of this lane, queued (record, lane) pairs, core points
  uint64_t n_pts = 0, n_queued = 0, n_core = 0;

  // this wavefront's record stage; a record's header (c, s, hw1) and first
  // body word (hs, c1, ay, hw2) are read as one ds_read_b128 each (volatile
  // keeps them single reads)
  const volatile lds_u32x4* const rb = (const volatile lds_u32x4*)s_rec_lds;
  // this lane's group's near range (clamped to the wavefront's, as the
  // staging of the deferred pairs tests it) and the longest of the eight
  const int32_t my_lo = min(max(gl[my_group], nlo), nhi);
  const int32_t my_hi = max(min(gh[my_group], nhi), my_lo);
  int32_t n_own = 0;   // uniform (scalar loads of the per-wave tables)
#pragma unroll
  for (int g = 0; g < kGroupsPerWave; ++g) {
    const int32_t g_lo = min(max(gl[g], nlo), nhi), g_hi = max(min(gh[g], nhi), g_lo);
    n_own = max(n_own, g_hi - g_lo);
  }
  // the record that fills a stage slot with nothing to evaluate: weight 0 at
  // finite x in every tier, never the one-node tier's obstacle, never queued
  const u32x4 null0 = {0u, 0u, 0u, __float_as_uint(-1.0f)};   // c = 0, s = 0, hw1 = -1
  const u32x4 null1 = {0u, __float_as_uint(1.0f), 0u, __float_as_uint(-1.0f)};   // hs, c1, ay, hw2
  const u32x4 null2 = {__float_as_uint(1.0f), 0u, 0u, 0u};   // y, a, rho, mask

  for (int k = k_begin; k < k_end; ++k) {
    float lay[P], layc[P], part[P];
#pragma unroll
    for (int j = 0; j < P; ++j) lay[j] = layc[j] = part[j] = 0.0f;
    const int64_t kb = (int64_t)k * a.n_lines;
    int nq = 0;   // queue entries not yet drained (uniform)
    // One record (stage slot q) at this lane's P points; a lane that does not
    // evaluate it (direct false) adds weight 0.  Three tiers per (record,
    // lane): the shifted one-node form, the middle tier (second-order shifted
    // form for y^2 < kF32S2Y2, else two-node) and the work queue.  One tier
    // for the whole wavefront: the middle tier holds wherever the one-node
    // form does, so if some lane needs it every lane runs it, and lanes that
    // need the queue instead run it too, at a safe x (the run's middle at
    // |x| = hw2) with weight 0, so that no branch splits the wavefront.
    auto record = [&](const int q, const u32x4 h0, const u32x4 h1, const bool direct) {
      const double c = __hiloint2double((int)h0.y, (int)h0.x);
      const float sv = __uint_as_float(h0.z);
      const float xm = (float)(nu_m - c) * sv;
      const float hs = __uint_as_float(h1.x), c1 = __uint_as_float(h1.y);
      const float ay = direct ? __uint_as_float(h1.z) : 0.0f;
      if constexpr (kCount) n_pts += direct && live ? (uint64_t)pts_lane : 0u;
      bool queued = false;
      if (__ballot(direct && fabsf(xm) < __uint_as_float(h0.w)) == 0) {
        // shifted one-node form (faddeeva.h): K A = (A y/sqrt(pi)) / (x^2 + y^2 - 3/2)
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const float x = fmaf((float)j - 0.5f * (float)(P - 1), hs, xm);
          const float D = fmaf(x, x, c1);
          part[j] = fmaf(ay, __builtin_amdgcn_rcpf(D), part[j]);
        }
      } else {
        const float hw2 = __uint_as_float(h1.w);
        const bool mid = fabsf(xm) >= hw2;
        queued = !mid && live && direct;
        const float xe = mid ? xm : hw2, aye = mid ? ay : 0.0f;
        // Both middle forms are u = x^2 + A, K A = (aye u + C) / (u^2 + B):
        // the second-order shifted form, K A = (A y/sqrt(pi)) r / (r^2 - 3/2 r
        // + 2 y^2 - 3/2) written about u = r - 3/4 (A = y^2 - 3/4, B = 2 y^2
        // - 33/16, C = 3/4 aye), and the two-node form, K A = (A y/sqrt(pi))
        // (w + 1) / (w^2 + 2 y^2) with w = r - 1/2 (A = y^2 - 1/2, B = 2 y^2,
        // C = aye); the record's c1 picks one per lane (6 VALU per point).
        const bool s2 = c1 < kF32S2Y2 - 1.5f;
        const float A = s2 ? c1 + 0.75f : c1 + 1.0f;
        const float B = s2 ? fmaf(2.0f, c1, 0.9375f) : fmaf(2.0f, c1, 3.0f);
        const float C = s2 ? 0.75f * aye : aye;
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const float x = fmaf((float)j - 0.5f * (float)(P - 1), hs, xe);
          const float u = fmaf(x, x, A);
          const float D = fmaf(u, u, B);
          part[j] = fmaf(fmaf(aye, u, C), __builtin_amdgcn_rcpf(D), part[j]);
        }
      }
      // append the queued lanes' entries (x at the lane's first point, from
      // the FP64 offset as the far test forms it; stage slot and lane) to
      // this wavefront's queue; the count nq advances in every lane
      const uint64_t msk = __ballot(queued);
      if (msk) {   // uniform
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(msk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)msk, 0u));
        if (queued) {
          const int pos = nq + (int)rank;   // < kQueueW: the loops drain before it fills
          s_qx[wave][pos] = (float)(nu_f - c) * sv;
          s_qi[wave][pos] = (uint16_t)((q << 6) | lane);
        }
        nq += (int)__popcll(msk);
      }
    };
    // The queue, drained by this wavefront.  kSurrogate: every queued point
    // goes to the network (inside its domain and gate) or to the full FP32
    // dispatch.  Otherwise two passes per round of 64 entries, each dense and
    // free of divergent regions.  Pass A: one entry per lane, all P points
    // in the six-node form (<= 3.2e-6 outside gh6_core_r2, faddeeva.h);
    // points inside (the core) go to the wavefront's core list (x and the
    // target kept in the entry).  Pass B: the round's core list, two points
    // per lane per round, trapezoid + residue form, each value written over
    // its list slot, from where the entry's lane takes it back.  The
    // entry's P values then reach the owning lane's sums through claimed
    // plain LDS updates.
    auto drain = [&](const int n) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the queue writes
      __builtin_amdgcn_wave_barrier();
      if constexpr (!kSurrogate) {
        auto core_pass = [&](const int m) {
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the list writes
          __builtin_amdgcn_wave_barrier();
          // two points per lane per round, so the list, queue and record
          // reads of both are in flight together; x as pass A formed it
          for (int b = lane; b < m; b += 128) {
            const bool two = b + 64 < m;
            const uint32_t ca = s_ci[wave][b], cb = two ? s_ci[wave][b + 64] : ca;
            const float xa0 = s_qx[wave][ca >> 3], xb0 = s_qx[wave][cb >> 3];
            const uint32_t qa = s_qi[wave][ca >> 3], qb = s_qi[wave][cb >> 3];
            const u32x4 ha1 = rb[3 * (qa >> 6) + 1], hb1 = rb[3 * (qb >> 6) + 1];   // hs, c1, ay, hw2
            const u32x4 ha2 = rb[3 * (qa >> 6) + 2], hb2 = rb[3 * (qb >> 6) + 2];   // y, a, rho, mask
            const float xa = fmaf((float)(ca & (P - 1)), __uint_as_float(ha1.x), xa0);
            const float xb = fmaf((float)(cb & (P - 1)), __uint_as_float(hb1.x), xb0);
            const float va = __uint_as_float(ha2.y) *
                             K_cr32(xa, __uint_as_float(ha2.x), __uint_as_float(ha2.z));
            const float vb = __uint_as_float(hb2.y) *
                             K_cr32(xb, __uint_as_float(hb2.x), __uint_as_float(hb2.z));
            s_ci[wave][b] = __float_as_uint(va);   // (slots b, b + 64: this lane's own)
            if (two) s_ci[wave][b + 64] = __float_as_uint(vb);
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");   // the values
          __builtin_amdgcn_wave_barrier();
        };
        for (int e0 = 0; e0 < n; e0 += 64) {
          const int e = e0 + lane;
          const bool has = e < n;
          float x0 = 0.0f;
          uint32_t qi = 0;
          if (has) {
            x0 = s_qx[wave][e];
            qi = s_qi[wave][e];
          }
          const int t = (int)(qi & 63);
          const u32x4 h1 = rb[3 * (qi >> 6) + 1];   // hs, c1, ay, hw2
          const float hs = __uint_as_float(h1.x), y2 = __uint_as_float(h1.y) + 1.5f;
          // six-node form: A K = (2 A y / pi) sum_k W_k (r + t_k^2) / D_k, and
          // A y = sqrt(pi) ay
          const float ay = has ? __uint_as_float(h1.z) : 0.0f;
          const float c2 = 2.0f * (float)kInvSqrtPi * ay;
          const float fy2 = 4.0f * y2;
          // core radius of the six-node form; y^2 = c1 + 3/2 carries y^2 to
          // 1.2e-7, which only ever picks the larger radius
          const float rcore = gh6_core_r2(y2);
          // a product of three D ~ |z|^12 overflows FP32 beyond |z| = 1e3
          // (broad lines, y > 1e3, or grids much coarser than a Doppler
          // width): such points take the shifted one-node form, within 2e-6
          // there
          const float xf = fmaxf(fabsf(x0), fabsf(fmaf((float)(P - 1), hs, x0)));
          const bool far_pts = __ballot(has && fmaf(xf, xf, y2) > 1e6f) != 0;
          uint32_t cmask = 0;
          float v[P];
#pragma unroll
          for (int j = 0; j < P; ++j) {
            const float x = fmaf((float)j, hs, x0);
            const float x2 = x * x;
            const float r = fmaf(x, x, y2), xmy = fmaf(x, x, -y2), fxy = x2 * fy2;
            const bool core = has && r < rcore;
            // (a select: at a core point the node sum may be infinite)
            const float g = c2 * gh6_sum32(r, xmy, fxy);
            v[j] = core ? 0.0f : g;
            cmask |= core ? (1u << j) : 0u;
          }
          // (a branch of its own, so that the common case does not also
          // compute the far form; the same r as above, so core points never
          // take it: gh6_core_r2 < 1e6)
          if (far_pts) {   // uniform
#pragma unroll
            for (int j = 0; j < P; ++j) {
              const float x = fmaf((float)j, hs, x0);
              const float r = fmaf(x, x, y2);
              if (r > 1e6f) v[j] = ay * __builtin_amdgcn_rcpf(r - 1.5f);
            }
          }
          // this lane's core points go to the list at an exclusive prefix sum
          // of the lanes' counts (bit planes of the 0..8 count, one ballot
          // each); pass B evaluates them and the lane reads its values back
          const uint32_t cnt = __builtin_popcount(cmask);
          uint32_t pre = 0, total = 0;
#pragma unroll
          for (int bit = 0; bit < 4; ++bit) {
            const uint64_t m = __ballot((cnt >> bit) & 1u);
            pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << bit;
            total += (uint32_t)__popcll(m) << bit;
          }
          if (total) {   // uniform (<= 64 P <= kCoreCap)
            int off = (int)pre;
#pragma unroll
            for (int j = 0; j < P; ++j) {
              if (cmask & (1u << j)) {
                s_ci[wave][off] = (uint32_t)e << 3 | (uint32_t)j;
                ++off;
              }
            }
            core_pass((int)total);
            off = (int)pre;
#pragma unroll
            for (int j = 0; j < P; ++j) {
              if (cmask & (1u << j)) {
                v[j] = __uint_as_float(s_ci[wave][off]);
                ++off;
              }
            }
            if constexpr (kCount) n_core += total;
          }
          // Deliver the P values to the owner's accumulators with plain LDS
          // reads and writes (LDS float atomics cost far more): in each
          // round every pending lane writes its id to its owner's claim
          // byte, the lanes that read their own id back hold distinct
          // owners and update, the rest try again (entries of several
          // records for one owner take one round each).
          bool pending = has;
          while (__ballot(pending)) {   // uniform
            if (pending) s_claim[wave][t] = (uint8_t)lane;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (pending && s_claim[wave][t] == (uint8_t)lane) {
              // all P reads first, then all P writes: one LDS round trip
              float* const dst = s_acc + wave * 64 + t;
              float cur[P];
#pragma unroll
              for (int j = 0; j < P; ++j) cur[j] = dst[j * kAccStride];
#pragma unroll
              for (int j = 0; j < P; ++j) dst[j * kAccStride] = cur[j] + v[j];
              pending = false;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          }
        }
      } else {
        for (int e = lane; e < n; e += 64) {
          const float x0 = s_qx[wave][e];
          const uint32_t qi = s_qi[wave][e];
          const MixedRec& R = s_rec[wave][qi >> 6];
          const int tt = wave * 64 + (int)(qi & 63);   // owning thread
          const float ac = R.a * (float)R.c;
          int exact = 0;
          for (int j = 0; j < P; ++j) {
            const float x = fmaf((float)j, R.hs, x0);
            if (surrogate_point(x, R.y, ac, a.mlp_gate)) continue;
            atomicAdd(&s_acc[j * kAccStride + tt], R.a * K_mixed(x, R.y));
            ++exact;
          }
          if (a.exact_count) atomicAdd(a.exact_count, (unsigned long long)exact);
        }
        // network pass: 64 / P queue entries (64 points) per round (lane =
        // entry (l / P), point (l % P)); all lanes stay active through the
        // MFMA evaluation.
        constexpr int kPer = 64 / P;
        for (int g = 0; g * kPer < n; ++g) {
          const int e = g * kPer + lane / P, j = lane % P;
          bool use = false;
          float x = 0.0f, y = 1.0f, am = 0.0f;
          int tt = 0;
          if (e < n) {
            const uint32_t qi = s_qi[wave][e];
            tt = wave * 64 + (int)(qi & 63);
            const MixedRec& R = s_rec[wave][qi >> 6];
            x = fmaf((float)j, R.hs, s_qx[wave][e]);
            y = R.y;
            am = R.a;
            use = surrogate_point(x, y, am * (float)R.c, a.mlp_gate);
          }
          const uint64_t users = __ballot(use);
          if (users == 0) continue;   // wavefront-uniform: nothing for the network here
          if (a.mlp_count && lane == 0) atomicAdd(a.mlp_count, (unsigned long long)__popcll(users));
          const float u = use ? mlp::input_u(x) : 0.0f;
          const float v = use ? mlp::input_v(__log10f(y)) : 0.0f;
          const float tv = mlp::wave_eval(u, v, a.mlp_params, a.mlp_wbf);
          if (use) atomicAdd(&s_acc[j * kAccStride + tt], am * mlp::kref(x, y) * __expf(tv));
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int j = 0; j < P; ++j) {
        part[j] += acc[j * kAccStride];
        acc[j * kAccStride] = 0.0f;
      }
    };
    auto fold = [&]() {
#pragma unroll
      for (int j = 0; j < P; ++j) {
        kahan_add(lay[j], layc[j], part[j]);
        part[j] = 0.0f;
      }
    };

    // ---- Phase 1: every group's own near range, eight records per group
    // per stage.  The eight lanes of group g load the next eight records of
    // its range (slots 8g .. 8g + 7), apply the exact FP64 tile-window test
    // and stage them; in iteration i every lane evaluates its group's record
    // 8g + i, so each iteration holds eight records and no lane idles while
    // its group has records left.
    for (int32_t base = 0; base < n_own; base += kStageG) {
      const int n_it = __builtin_amdgcn_readfirstlane(min(kStageG, n_own - base));
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the previous stage's readers
      const int slot = (int)(my_group * kStageG) + (lane & (kStageG - 1));
      {
        const int32_t r = my_lo + base + (lane & (kStageG - 1));
        u32x4 w0 = null0, w1 = null1, w2 = null2;
        if (r < my_hi) {
          const u32x4* const src = (const u32x4*)(recs + kb + r);
          const u32x4 v0 = src[0], v1 = src[1], v2 = src[2];
          const double c = __hiloint2double((int)v0.y, (int)v0.x);
          if (c >= lo_b && c <= hi_b) {   // (a skipped record's NaN centre fails)
            w0 = v0;
            w1 = v1;
            // the LDS copy carries the core form's rho in ds1
            w2 = v2;
            w2.z = __float_as_uint(cr32_rho(__uint_as_float(v2.x)));
          }
        }
        lds_u32x4* const dst = (lds_u32x4*)&s_rec_lds[slot];
        dst[0] = w0;
        dst[1] = w1;
        dst[2] = w2;
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // header and body of the next record read one iteration ahead
      const volatile lds_u32x4* const pg = rb + 3 * (my_group * kStageG);
      u32x4 h0 = pg[0], h1 = pg[1];
      for (int it = 0; it < n_it; ++it) {
        u32x4 n0 = h0, n1 = h1;
        if (it + 1 < n_it) {
          n0 = pg[3 * (it + 1)];
          n1 = pg[3 * (it + 1) + 1];
        }
        // a null slot (past the group's range or outside the tile window)
        // has hw2 = -1; such a lane is not counted and never queued
        record((int)(my_group * kStageG) + it, h0, h1, __uint_as_float(h1.w) >= 0.0f);
        h0 = n0;
        h1 = n1;
        if (nq > kQueueW - 64) {   // uniform: one more record could not be taken
          if constexpr (kCount) n_queued += (uint64_t)nq;
          drain(nq);
          nq = 0;
          if (it + 1 < n_it) {   // not held in registers through the drain
            h0 = pg[3 * (it + 1)];
            h1 = pg[3 * (it + 1) + 1];
          }
        }
      }
      if (nq > 0) {   // uniform: the stage is overwritten next
        if constexpr (kCount) n_queued += (uint64_t)nq;
        drain(nq);
        nq = 0;
      }
      if (((base + kStageG) & (kFoldEvery - 1)) == 0) fold();
    }
    fold();

    // ---- Phase 2: pairs the group series left to this kernel.  Outside a
    // group's own range a record is evaluated point by point where neither
    // the one-node nor the middle-tier form holds at every point of the group
    // (group_one_node_holds with group_thr and group_mid_thr, the same tests
    // in the same arithmetic as k_far_series_mixed's group level, so each
    // (group, record) pair is taken exactly once); the series kernel counts
    // them per (layer, wavefront), and the scan is skipped where there are
    // none.  Lane l tests record base + l of the wavefront's near range for
    // all eight groups; records some group needs are packed into the stage
    // with that mask, and every lane takes each staged record, with weight 0
    // where its group is not in the mask.
    const uint32_t n_left = a.deferred[(int64_t)k * (a.n_wg * kWaves) + wg * kWaves + wave];
    for (int32_t base = nlo; n_left != 0 && base < nhi; base += kStage) {
      const int n_here = __builtin_amdgcn_readfirstlane(min(kStage, nhi - base));
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      // the record as three 16-byte words in registers: (c, s, hw1),
      // (hs, c1, ay, hw2), (y, a, ds1, dm)
      u32x4 w0 = {0u, 0u, 0u, 0u}, w1 = w0, w2 = w0;
      uint32_t gmask = 0;
      if (lane < n_here) {
        const u32x4* const src = (const u32x4*)(recs + kb + base + lane);
        w0 = src[0];
        w1 = src[1];
        w2 = src[2];
        const double c = __hiloint2double((int)w0.y, (int)w0.x);
        if (c >= lo_b && c <= hi_b) {
          const int32_t r = base + lane;
          const float d0 = group_d0(gt, c), thr = group_thr(gt, __uint_as_float(w2.z));
          const float thr2 = group_mid_thr(gt, __uint_as_float(w2.w));
#pragma unroll
          for (int grp = 0; grp < kGroupsPerWave; ++grp) {
            const int32_t g_lo = min(max(gl[grp], nlo), nhi), g_hi = max(min(gh[grp], nhi), g_lo);
            const float goff = group_offset(gt, grp);
            const bool deferred = (r < g_lo || r >= g_hi) && !group_one_node_holds(d0, thr, goff) &&
                                  !group_one_node_holds(d0, thr2, goff);
            gmask |= deferred ? (1u << grp) : 0u;
          }
        }
      }
      const uint64_t keep = __ballot(gmask != 0);
      const int cnt = (int)__popcll(keep);   // records in the stage (uniform)
      if (cnt == 0) continue;                // (the stage was not written)
      if (gmask != 0) {
        // the LDS copy carries the core form's rho in ds1 and the mask in dm
        w2.z = __float_as_uint(cr32_rho(__uint_as_float(w2.x)));
        w2.w = gmask;
        lds_u32x4* const dst = (lds_u32x4*)&s_rec_lds[__builtin_amdgcn_mbcnt_hi(
            (uint32_t)(keep >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)keep, 0u))];
        dst[0] = w0;
        dst[1] = w1;
        dst[2] = w2;
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (int q = 0; q < cnt; ++q) {
        const volatile lds_u32x4* const body = rb + 3 * q + 1;
        const uint32_t gm = __builtin_amdgcn_readfirstlane(((const volatile lds_u32*)body)[7]);
        record(q, rb[3 * q], *body, (gm >> my_group) & 1u);
        if (((q + 1) & (kFoldEvery - 1)) == 0) fold();
        if (nq > kQueueW - 64) {   // uniform
          if constexpr (kCount) n_queued += (uint64_t)nq;
          drain(nq);
          nq = 0;
        }
      }
      if (nq > 0) {   // uniform
        if constexpr (kCount) n_queued += (uint64_t)nq;
        drain(nq);
        nq = 0;
      }
      fold();
    }
    const double T = a.t_k[k];
    bool bad = false;
    const float* const far_k = static_cast<const float*>(a.far_sum) + (int64_t)k * a.n_nu;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int64_t i = i0 + j;
      if (i < t1) {
        kahan_add(lay[j], layc[j], far_k[i]);   // the far field of this layer
        // R(nu, T) in FP32 (relative error ~1e-7, far inside the mixed budget)
        const float nu = (float)grid_nu(a.nu_min, a.step, i);
        const float R = nu * tanhf(0.5f * (float)kC2 * nu / (float)T);
        const float v = (lay[j] - layc[j]) * R;
        bad |= !isfinite(v);
        float2 acc2 = make_float2(0.0f, 0.0f);
        if (k > k_begin) acc2 = tau_acc[i];
        kahan_add(acc2.x, acc2.y, v);
        tau_acc[i] = acc2;
        if (a.tau_layers) static_cast<float*>(a.tau_layers)[(int64_t)k * a.n_nu + i] = v;
      }
    }
    if (bad) flag_od(a.err, k);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int64_t i = i0 + j;
    if (i < t1) {
      if (k_begin == k_end) tau_acc[i] = make_float2(0.0f, 0.0f);   // no layers in this group
      if (a.layer_groups > 1) continue;   // k_tau_epilogue32 adds the groups
      const float2 acc2 = tau_acc[i];
      const float t = acc2.x - acc2.y;
      for (int m = 0; m < a.n_mu; ++m)
        static_cast<float*>(a.trans)[(int64_t)m * a.n_nu + i] = expf(-t * (float)a.inv_mu[m]);
      if (a.tau) static_cast<float*>(a.tau)[i] = t;
    }
  }
  if constexpr (kCount) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) n_pts += __shfl_xor(n_pts, off);
    if (lane == 0) {
      count_add(a.work, kWorkNearEvals, n_pts);
      count_add(a.work, kWorkQueued, n_queued);
      count_add(a.work, kWorkCorePoints, n_core);
    }
  }
}

// Four waves per SIMD (<= 128 VGPRs): every tile of a 2 cm^-1 x 0.001 grid
// (1001 workgroups) is resident at once on 256 CUs.  (Measured: 3 waves/SIMD
// left a second partial round, +12%; 16 points x 128 threads at 2 waves/SIMD
// hid latency worse, +20%.)
template <int P, bool kCount>
__global__ __launch_bounds__(kMixedBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_voigt_sum_mixed(const SumArgs a) {
  voigt_sum_mixed_body<P, false, kCount>(a);
}

// Same register budget: the network's MFMA tiles spill around the (rarely
// taken) network pass instead of halving the occupancy of the main loop.
template <int P>
__global__ __launch_bounds__(kMixedBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_voigt_sum_surrogate(const SumArgs a) {
  voigt_sum_mixed_body<P, true, false>(a);
}

// --------------------------------------------------------- test entry --

__global__ void k_voigt_eval64(int64_t n, const double* x, const double* y, double* k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) k[i] = K_fp64(x[i], y[i]);
}

__global__ void k_voigt_eval32(int64_t n, const float* x, const float* y, float* k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) k[i] = K_mixed(x[i], y[i]);
}

// Layer-group epilogue: tau = sum over groups of the partial taus (in group,
// i.e. layer, order), T = exp(-tau/mu); one thread per grid point.
__global__ void k_tau_epilogue64(const SumArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_nu) return;
  const double* part = static_cast<const double*>(a.tau_part);
  double t = 0.0;
  for (int g = 0; g < a.layer_groups; ++g) t += part[(int64_t)g * a.n_nu + i];
  for (int m = 0; m < a.n_mu; ++m)
    static_cast<double*>(a.trans)[(int64_t)m * a.n_nu + i] = exp(-t * a.inv_mu[m]);
  if (a.tau) static_cast<double*>(a.tau)[i] = t;
}

__global__ void k_tau_epilogue32(const SumArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_nu) return;
  const float2* part = static_cast<const float2*>(a.tau_part);
  float t = 0.0f, c = 0.0f;
  for (int g = 0; g < a.layer_groups; ++g) {
    const float2 p = part[(int64_t)g * a.n_nu + i];
    kahan_add(t, c, p.x - p.y);
  }
  t -= c;
  for (int m = 0; m < a.n_mu; ++m)
    static_cast<float*>(a.trans)[(int64_t)m * a.n_nu + i] = expf(-t * (float)a.inv_mu[m]);
  if (a.tau) static_cast<float*>(a.tau)[i] = t;
}


}  // namespace

int sum_points_per_thread(int precision) {
  return precision == LBL_FP64 ? kPointsFp64 : kPointsMixed;
}

int sum_block(int precision) { return precision == LBL_FP64 ? kBlock : kMixedBlock; }

hipError_t launch_far_series(const SumArgs& a, int precision, hipStream_t s) {
  if (a.n_wg <= 0 || a.n_layers <= 0) return hipSuccess;
  const int64_t n_far = a.n_wg * a.n_layers;
  if (n_far * kBlock > (int64_t)UINT32_MAX) return hipErrorInvalidConfiguration;
  // a.work set: the counting variants (same results, per-unit work counts)
  if (precision == LBL_FP64) {
    if (a.work)
      k_far_series_fp64<kPointsFp64, true><<<(unsigned)n_far, kBlock, 0, s>>>(a);
    else
      k_far_series_fp64<kPointsFp64, false><<<(unsigned)n_far, kBlock, 0, s>>>(a);
  } else if (a.work) {
    k_far_series_mixed<kPointsMixed, true><<<(unsigned)n_far, kMixedBlock, 0, s>>>(a);
  } else {
    k_far_series_mixed<kPointsMixed, false><<<(unsigned)n_far, kMixedBlock, 0, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_voigt_sum(const SumArgs& a, int precision, hipStream_t s) {
  if (a.n_wg <= 0) return hipSuccess;
  const int64_t n_blocks = a.n_wg * a.layer_groups;
  if (a.layer_groups < 1 || n_blocks * kBlock > (int64_t)UINT32_MAX) return hipErrorInvalidConfiguration;
  if (precision == LBL_FP64) {
    if (a.work)
      k_voigt_sum_fp64<kPointsFp64, true><<<(unsigned)n_blocks, kBlock, 0, s>>>(a);
    else
      k_voigt_sum_fp64<kPointsFp64, false><<<(unsigned)n_blocks, kBlock, 0, s>>>(a);
  } else if (precision == LBL_SURROGATE) {
    k_voigt_sum_surrogate<kPointsMixed><<<(unsigned)n_blocks, kMixedBlock, 0, s>>>(a);
  } else if (a.work) {
    k_voigt_sum_mixed<kPointsMixed, true><<<(unsigned)n_blocks, kMixedBlock, 0, s>>>(a);
  } else {
    k_voigt_sum_mixed<kPointsMixed, false><<<(unsigned)n_blocks, kMixedBlock, 0, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_tau_epilogue(const SumArgs& a, int precision, hipStream_t s) {
  if (a.layer_groups <= 1 || a.n_nu <= 0) return hipSuccess;
  const unsigned g = (unsigned)((a.n_nu + 255) / 256);
  if (precision == LBL_FP64)
    k_tau_epilogue64<<<g, 256, 0, s>>>(a);
  else
    k_tau_epilogue32<<<g, 256, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_voigt_eval(int64_t n, const void* x, const void* y, void* k, int precision,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned g = (unsigned)((n + 255) / 256);
  if (precision == LBL_FP64)
    k_voigt_eval64<<<g, 256, 0, s>>>(n, static_cast<const double*>(x),
                                     static_cast<const double*>(y), static_cast<double*>(k));
  else
    k_voigt_eval32<<<g, 256, 0, s>>>(n, static_cast<const float*>(x),
                                     static_cast<const float*>(y), static_cast<float*>(k));
  return hipGetLastError();
}

}  // namespace lbl

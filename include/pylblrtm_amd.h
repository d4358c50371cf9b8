/*
 * pylblrtm_amd — C ABI of the MI355X line-by-line optical-depth library
 * (libpylblrtm_amd.so, HIP kernels built for gfx950).
 *
 * The library replaces the body of PyLBLRTM's private seam
 *
 *     _build_optical_depth(nu, sd, layers, line_cutoff_cm1, tile_width_cm1)
 *         -> od[n_layers, n_nu]                       (pylblrtm/forward.py:157-317)
 *
 * and the column/transmittance assembly its two callers do with it
 *
 *     T = exp(-od.sum(axis=0) / mu_view)              (pylblrtm/forward.py:151,
 *                                                      pylblrtm/config.py:266-273)
 *
 * The host keeps the reference's prelude (grouping by (mol_id, iso_id) in
 * sorted order, warnings, the band mask of forward.py:215-220, molar-mass
 * fallback) and hands the device the surviving lines.
 *
 * Conventions
 *  - Every array pointer in the structs below is DEVICE memory owned by the
 *    caller; the library never allocates or frees caller memory.  Scratch
 *    lives in a caller-provided workspace sized by lbl_workspace_bytes().
 *  - Calls are host-synchronous (they return after the work on `stream`,
 *    a hipStream_t, has finished), like the reference function they replace.
 *    Calls on the same stream from several threads must be serialised.
 *  - Return codes: LBL_OK, or
 *      LBL_EINVAL     invalid argument                  -> ValueError
 *      LBL_ESTRENGTH  non-finite / negative strength     -> FloatingPointError
 *                     (forward.py:280-283; err_layer/err_group say where)
 *      LBL_ENONFINITE non-finite optical depth           -> FloatingPointError
 *                     (forward.py:314-315; err_layer says where)
 *      LBL_EHIP       HIP runtime error                  -> RuntimeError
 *    lbl_last_error() returns a thread-local message for the last failure.
 */
#ifndef PYLBLRTM_AMD_H
#define PYLBLRTM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LBL_OK = 0,
  LBL_EINVAL = 1,
  LBL_ESTRENGTH = 2,
  LBL_ENONFINITE = 3,
  LBL_EHIP = 4
};

/* Arithmetic of the Voigt line sum. */
enum {
  LBL_FP64 = 0,      /* FP64 throughout (tau to ~1e-12 relative of the FP64 oracle) */
  LBL_MIXED = 1,     /* FP32 line sum, FP64 centre offset + Boltzmann exponent */
  LBL_SURROGATE = 2  /* LBL_MIXED + MLP surrogate for K(x,y) in its trained domain */
};

/* Lines that survive the host prelude, SORTED BY nu0 ASCENDING.
 * Fields are those of SpectralData.lines (pylblrtm/spectral_data.py:21-26,
 * pylblrtm/hitran_manager.py:314-324); `group` replaces (mol_id, iso_id) by
 * the group's position in the host's sorted group list.  Device-side indices
 * are checked by the kernels: a group outside [0, n_groups), a gas row
 * outside [0, n_gases) or an empty TIPS range makes the call return
 * LBL_EINVAL (every kernel still exits normally). */
typedef struct lbl_lines {
  int64_t n_lines;
  const double *nu0, *sref, *g_air, *g_self, *elow, *n_air, *shift;
  const int32_t *group;
} lbl_lines_t;

/* One entry per (mol_id, iso_id) group, in the reference's sorted order
 * (forward.py:194).  TIPS tables (SpectralData.tips) are concatenated:
 * group g owns tips_t/tips_q[tips_offset[g] .. tips_offset[g+1]). */
typedef struct lbl_groups {
  int32_t n_groups;
  const int32_t *gas;          /* row of lbl_layers.vmr / .col used by the group */
  const double *molar_mass;    /* g/mol (28.97 fallback already applied) */
  const int32_t *tips_offset;  /* n_groups + 1 entries */
  const double *tips_t, *tips_q;
} lbl_groups_t;

/* Layer averages of AtmosphericProfile.build_layers (atmosphere.py:110-141). */
typedef struct lbl_layers {
  int32_t n_layers, n_gases;
  const double *p_atm, *t_k;   /* n_layers */
  const double *vmr, *col;     /* n_gases x n_layers, row-major by gas */
} lbl_layers_t;

/* The wavenumber grid nu_i = nu_min + i*step (multiply then add, as
 * forward.py:141-142) and the reference tiling (forward.py:175-176,295-301). */
typedef struct lbl_grid {
  double nu_min, step;
  int64_t n_nu;
  int64_t tile_samples;        /* max(1, floor(tile_width/(nu[1]-nu[0]) + 1e-8)) */
  double cutoff;               /* line_cutoff_cm1 */
  double shift_bound;          /* >= max over lines of |shift| * max_k P_k, the
                                  bound of forward.py:190-191 (widens the candidate
                                  line range per tile; inclusion stays exact) */
} lbl_grid_t;

/* The grid of the seam's own argument: _build_optical_depth receives the
 * wavenumber array nu itself (forward.py:157-163) and takes its tile size
 * from nu[1] - nu[0] (forward.py:175-176).  Given that array (HOST memory,
 * n_nu points), fills grid with nu_min = nu[0], n_nu, tile_samples =
 * max(1, floor(tile_width_cm1 / (nu[1] - nu[0]) + 1e-8)) (step 0.001 when
 * n_nu == 1, as the reference), cutoff, shift_bound = 0 (it depends on the
 * lines; the caller sets it), and a step for which nu[0] + i*step (multiply,
 * then add) equals every nu[i] bit for bit, so the kernels' grid is the
 * caller's.  The arrays compute_transmittance and run_forward build
 * (forward.py:141-142, config.py:246-247) always have one; an array that is
 * not of that form (unevenly spaced, decreasing, non-finite) or longer than
 * 2^40 points returns LBL_EINVAL.  (tile_samples is capped at 2^62; the
 * entry points below treat any tile longer than the grid as the whole
 * grid.)  Host code only. */
int lbl_grid_from_nu(const double *nu, int64_t n_nu, double tile_width_cm1,
                     double line_cutoff_cm1, lbl_grid_t *grid);

/* Outputs.  `transmittance` is n_mu x n_nu; optional outputs may be NULL.
 * Element type: double for LBL_FP64, float otherwise. */
typedef struct lbl_outputs {
  int32_t n_mu;
  const double *mu;            /* HOST array of n_mu viewing cosines in (0, 1] */
  void *transmittance;         /* exp(-sum_k od_k / mu_m)            (forward.py:151) */
  void *optical_depth;         /* sum_k od_k (vertical), or NULL     (config.py:270) */
  void *optical_depth_layers;  /* od[n_layers][n_nu], or NULL        (config.py:273) */
} lbl_outputs_t;

/* Bytes of scratch lbl_optical_depth() needs for this problem. */
int lbl_workspace_bytes(const lbl_lines_t *lines, const lbl_layers_t *layers,
                        const lbl_grid_t *grid, int precision, size_t *bytes);

/* The whole path: per-(layer, line) scaling, tile binning, Voigt line sum,
 * radiation term, layer sum and exp(-tau/mu).  Replaces
 * _build_optical_depth (forward.py:157-317) plus forward.py:151.
 * On LBL_ESTRENGTH / LBL_ENONFINITE, *err_layer and *err_group (host ints,
 * may be NULL) receive the layer index and group index of the first failure
 * in the reference's loop order (layer, then group). */
int lbl_optical_depth(const lbl_lines_t *lines, const lbl_groups_t *groups,
                      const lbl_layers_t *layers, const lbl_grid_t *grid, int precision,
                      const lbl_outputs_t *out, void *workspace, size_t workspace_bytes,
                      int32_t *err_layer, int32_t *err_group, void *stream);

/* lbl_optical_depth without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * scratch (per-(layer, line) records, tile ranges, partial sums, error word)
 * lives in a fixed device array of LBL_NOALLOC_SCRATCH_BYTES inside the
 * library, one per GPU, held until the process ends; the call allocates and
 * frees nothing.  Precision LBL_FP64 or LBL_MIXED; results, outputs and
 * errors as lbl_optical_depth.  A problem whose lbl_workspace_bytes exceed
 * the limit is rejected with LBL_EINVAL before any device work (the CO2
 * window of 5,000 lines x 20 layers x 50,001 points fits; the five-gas
 * workload does not).  The device current on the calling thread must be the
 * stream's device.  Every call on one GPU uses that GPU's one array: calls
 * that overlap on one GPU (other threads, other streams) race, and must be
 * serialised by the caller.  Kernel timing (lbl_enable_timing) is not
 * recorded by this entry point. */
#define LBL_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int lbl_optical_depth_noalloc(const lbl_lines_t *lines, const lbl_groups_t *groups,
                              const lbl_layers_t *layers, const lbl_grid_t *grid, int precision,
                              const lbl_outputs_t *out, int32_t *err_layer, int32_t *err_group,
                              void *stream);

/* Network of the learned Voigt surrogate (see lbl_mlp_* below) and its gate:
 * a near-centre point uses the network iff |x| <= 50, -4 <= log10 y <= 2 and
 * A * nu_c * K_ref(x, y) <= gate (a bound on its optical-depth contribution
 * divided by the network's relative error budget); all other points keep the
 * exact FP32 Faddeeva dispatch. */
typedef struct lbl_surrogate {
  const float *params;           /* device, lbl_mlp_num_params() floats */
  const unsigned short *wbf16;   /* device, 3*64*64 BF16 hidden weights */
  double gate;
  uint64_t *n_network;           /* device counter, incremented by network evaluations (nullable) */
  uint64_t *n_exact_near;        /* device counter, exact near-centre evaluations (nullable) */
} lbl_surrogate_t;

/* lbl_optical_depth with precision LBL_SURROGATE (no counterpart in the
 * reference): the mixed line sum with the network fused into the near-centre
 * evaluations.  Outputs are float. */
int lbl_optical_depth_surrogate(const lbl_lines_t *lines, const lbl_groups_t *groups,
                                const lbl_layers_t *layers, const lbl_grid_t *grid,
                                const lbl_outputs_t *out, const lbl_surrogate_t *sur,
                                void *workspace, size_t workspace_bytes, int32_t *err_layer,
                                int32_t *err_group, void *stream);

/* Evaluate the device Voigt function K(x, y) = Re w(x + iy) on n device
 * points with the region dispatch the line sum uses (precision LBL_FP64:
 * double in/out; LBL_MIXED: float in/out).  Stands in for
 * scipy.special.wofz(z).real at forward.py:62 in known-answer tests. */
int lbl_voigt_eval(int64_t n, const void *x, const void *y, void *k_out, int precision,
                   void *stream);

/* Adjoint of the transmittance (no counterpart in the reference: the
 * gradients retrievals need).  Given the forward transmittance T_m(nu) and
 * per-layer optical depth od_k(nu) (float, n_mu x n_nu and n_layers x n_nu,
 * both from one lbl_optical_depth call with LBL_MIXED on the same inputs) and
 * an upstream gradient G_m(nu) = dL/dT_m(nu) (float, n_mu x n_nu), returns
 *   grad_t[k]          = dL/dT_k            (double, n_layers)
 *   grad_vmr[g][k]     = dL/dvmr_{k,g}      (double, n_gases x n_layers)
 * with the column of gas g in layer k taken as vmr_{k,g} * (air column of
 * layer k), i.e. lbl_layers.col must equal vmr * air.  The Voigt values are
 * recomputed (FP32), nothing of size lines x grid is stored.  Where a
 * mixing ratio is exactly 0 the forward skips that gas in that layer
 * (forward.py:253-255) and this entry point leaves grad_vmr there at 0; the
 * forward is continuous there (strength ~ vmr * air, forward.py:270-278), and
 * the one-sided derivative is sum_nu W(nu) od_unit_k(nu), W = -sum_m G_m T_m
 * / mu_m, od_unit the gas's layer optical depth from one lbl_optical_depth
 * call with col = air_k for it and 0 for every other (gas, layer).  The
 * package's host layer (pylblrtm_amd.device.OpticalDepthProblem.vjp, used by
 * the autograd functions) adds exactly that. */
int lbl_vjp_workspace_bytes(const lbl_lines_t *lines, const lbl_layers_t *layers,
                            const lbl_grid_t *grid, size_t *bytes);
int lbl_transmittance_vjp(const lbl_lines_t *lines, const lbl_groups_t *groups,
                          const lbl_layers_t *layers, const lbl_grid_t *grid, int32_t n_mu,
                          const double *mu, const float *transmittance,
                          const float *optical_depth_layers, const float *grad_transmittance,
                          double *grad_t, double *grad_vmr, void *workspace,
                          size_t workspace_bytes, void *stream);

/* Learned Voigt surrogate (no counterpart in the reference; its Voigt is
 * scipy.special.wofz at forward.py:62).  MLP 2-64-64-64-64-1 (SiLU) on
 * u = 2 asinh(|x|)/asinh(50) - 1 (K is even in x), v = (log10 y + 1)/3
 * predicting t = log(K / K_ref),
 * K_ref = exp(-x^2)/(1 + sqrt(pi) y) + y/(sqrt(pi)(x^2 + y^2 + 1/2)).
 * Parameters: FP32 vector of lbl_mlp_num_params() entries (layout in
 * pylblrtm_amd/csrc/surrogate.h); wbf16 = BF16 copies of the three hidden
 * 64x64 weight matrices (3*64*64 entries) used by the MFMA kernels. */
int lbl_mlp_num_params(void);
/* n seeded samples x ~ U[-50,50], log10 y ~ U[-4,2], target t from the FP64
 * device Faddeeva function. */
int lbl_mlp_make_data(int64_t n, uint64_t seed, float *x, float *log10y, float *target,
                      void *stream);
int lbl_mlp_pack(const float *params, unsigned short *wbf16, void *stream);
int lbl_mlp_train_workspace_bytes(int batch, size_t *bytes);
/* One step on samples idx(j) = (perm_a*j + perm_b) mod n_data, j = base..base+batch-1
 * (n_data a power of two, perm_a odd): gradient of mean (out - t)^2 into
 * grad (device, num_params floats), loss sum into loss (device float), then,
 * if lr > 0, Adam (beta 0.9/0.999, eps 1e-8, step adam_t >= 1) on params and
 * a fresh wbf16.  Asynchronous on `stream`. */
int lbl_mlp_train_step(float *params, float *adam_m, float *adam_v, unsigned short *wbf16,
                       const float *x, const float *log10y, const float *target, int64_t n_data,
                       uint64_t perm_a, uint64_t perm_b, int64_t base, int batch, float lr,
                       int adam_t, float *grad, float *loss, void *workspace,
                       size_t workspace_bytes, void *stream);
/* Network output t for n points (BF16 MFMA inference). */
int lbl_mlp_eval(int64_t n, const float *x, const float *log10y, const float *params,
                 const unsigned short *wbf16, float *t_out, void *stream);

/* HITRAN-2004 .par line-list reader (host code; replaces the line loop of
 * par_to_dict, pylblrtm/hitran_manager.py:816-872).  Text lines are split as
 * Python's universal-newline mode splits them (\n, \r\n, \r).  Call once
 * with capacity 0 to get *n_text_lines, then with arrays of that length:
 * status[i] = -1 for a line shorter than 155 characters (skipped), -2 for a
 * line holding non-ASCII bytes (left to the caller, which decodes UTF-8 with
 * replacement as Python does), else the number 0..9 of leading fields
 * mol_id, iso_id, nu0, Sref, g_air, g_self, Elow, n_air, shift that convert
 * under Python's int()/float() rules (the reference appends field by field,
 * so a malformed line keeps the fields before the bad one); ids[2i..2i+1]
 * and values[7i..7i+6] hold them.  LBL_EINVAL if the file cannot be read. */
int lbl_par_read(const char *path, int64_t capacity, int64_t *n_text_lines, int8_t *status,
                 int32_t *ids, double *values);

/* Kernel timing (no counterpart in the reference; for benchmarks).  When
 * enabled, lbl_optical_depth records HIP events on its stream around the
 * per-(layer, line) kernel and the line-sum kernel; lbl_last_timings returns
 * their durations in ms for the calling thread's last call. */
void lbl_enable_timing(int on);
int lbl_last_timings(double *ms_line_params, double *ms_voigt_sum);
/* The line-sum time split (no counterpart in the reference; for
 * benchmarks): every precision runs a far-field series kernel
 * (k_far_series_*) before the near-field line sum (k_voigt_sum_*, with the
 * layer-group epilogue); these are their durations in ms. */
int lbl_last_timings_split(double *ms_far_series, double *ms_near_sum);
/* The part of ms_near_sum spent in the layer-group epilogue (the sum of the
 * groups' partial optical depths and T = exp(-tau/mu) over the grid; 0 when
 * the line sum writes T itself).  No counterpart in the reference. */
int lbl_last_timing_epilogue(double *ms_epilogue);

/* Work counting (no counterpart in the reference; for measurement and for
 * tests of the work counts).  When enabled, lbl_optical_depth (LBL_FP64 and
 * LBL_MIXED) runs the counting variants of its line-sum kernels, which give
 * the same results and add up what they evaluate; lbl_last_work_count then
 * returns, for the calling thread's last call, up to n of these counts:
 *   [0] near-field point evaluations (records inside the tile window x the
 *       points of each wavefront (LBL_FP64) or group of 64 points
 *       (LBL_MIXED) that takes them point by point)
 *   [1] (record, slice) series pairs      [2] (record, wavefront) series pairs
 *   [3] records the slice level passes on (its form does not hold there)
 *   [4] records the wavefront level evaluates point by point
 *   [5] their point evaluations           [6] queued near-centre (record, lane) pairs
 *   [7] queued core points (LBL_MIXED)    [8] slice-list overflows
 *   [9] wavefront-level candidates taken from the slice level's list [3]
 *       (one per record and live wavefront of its slice)
 *   group level (records of a wavefront's near range outside a group's
 *   own): [10] (record, group) series pairs, [11] pairs left to the near
 *   field (no series form of the level holds at every point of the group),
 *   [12] the near field's point evaluations of [11] (part of [0]),
 *   [13] LBL_MIXED: the pairs of [10] summed in the middle-tier form
 *   [14] series terms summed (each pair takes the terms its distance
 *   needs): LBL_MIXED over the one-node pairs of [1], [2] and [10], LBL_FP64
 *   over the pairs of [1]   [15] LBL_MIXED: over the pairs of [13];
 *   LBL_FP64: over the pairs of [2] and [10] (each pole pair's terms once)
 * All counts are 0 after a call made without counting. */
void lbl_enable_work_count(int on);
int lbl_last_work_count(uint64_t *counts, int32_t n);

/* Message of the last failure on the calling thread ("" if none). */
const char *lbl_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PYLBLRTM_AMD_H */

"""Parity of the HIP path against the CPU oracle and the reference's golden
vectors.  Needs a ROCm GPU; everything is called through the C ABI."""

import json
import os
import sys
import warnings

import numpy as np
import pytest
from scipy.special import wofz

from oracle import lbl_oracle as O
from pylblrtm_amd import AtmosphericProfile, SpectralData, compute_transmittance, run_forward, synth
from pylblrtm_amd.device import OpticalDepthProblem, voigt_eval

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

TAU_RTOL_FP64 = 1e-9     # FP64 mode: tau relative to the FP64 oracle
DT_ATOL_MIXED = 1e-5     # mixed mode: max |dT| against FP64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ── device Faddeeva ──────────────────────────────────────────────────────────

def _plane(n=200_000, seed=11):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.uniform(-10, 10, n // 2), rng.uniform(-120, 120, n // 2)])
    y = 10.0 ** rng.uniform(-6, 2.5, n)
    return x, y


def test_voigt_fp64_known_answers():
    say("case: voigt fp64 KAT")
    kat = json.load(open(os.path.join(HERE, "golden", "wofz_kat.json")))
    z = np.array([complex(*v) for v in kat["z"]])
    w = np.array([complex(*v) for v in kat["w"]])
    K = voigt_eval(z.real, z.imag, "fp64")
    np.testing.assert_allclose(K, w.real, rtol=2e-11, atol=0)


def test_voigt_fp64_against_scipy_plane():
    say("case: voigt fp64 plane")
    x, y = _plane()
    K = voigt_eval(x, y, "fp64")
    ref = wofz(x + 1j * y).real
    rel = np.abs(K - ref) / ref
    assert rel.max() < 2e-11, (rel.max(), x[rel.argmax()], y[rel.argmax()])


def test_voigt_fp64_real_axis_and_imag_axis():
    say("case: voigt fp64 axes")
    from scipy.special import erfcx
    x = np.linspace(-5.5, 5.5, 1001)
    np.testing.assert_allclose(voigt_eval(x, np.zeros_like(x), "fp64"), np.exp(-x * x),
                               rtol=1e-12, atol=1e-300)
    y = np.geomspace(1e-6, 1e3, 500)       # y >= 700 takes the 2-node form (<= 1.04e-11)
    np.testing.assert_allclose(voigt_eval(np.zeros_like(y), y, "fp64"), erfcx(y), rtol=2e-11)


def test_voigt_mixed_against_scipy_plane():
    say("case: voigt fp32 plane")
    x, y = _plane(seed=12)
    K = voigt_eval(x, y, "mixed").astype(np.float64)
    ref = wofz(x.astype(np.float32).astype(np.float64) + 1j * y.astype(np.float32).astype(np.float64)).real
    rel = np.abs(K - ref) / ref
    assert rel.max() < 1e-5, (rel.max(), x[rel.argmax()], y[rel.argmax()])


# ── golden vectors of the reference ──────────────────────────────────────────

def _golden(name):
    from golden_io import unpack_case
    z = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    lines, tips, mm, p, t, vmr, meta = unpack_case(z)
    sd = SpectralData(lines=lines, tips=tips, molar_masses=mm)
    prof = AtmosphericProfile(p, t, vmr)
    return z, sd, prof, meta


@pytest.mark.parametrize("name", ["co2_mini", "three_gas_edge", "tile1"])
def test_golden_fp64(name):
    say(f"case: golden fp64 {name}")
    z, sd, prof, meta = _golden(name)
    lo, hi, step = meta["grid"]
    for j, mu in enumerate(meta["mus"]):
        cfg = {"spectral": {"nu_min_cm1": lo, "nu_max_cm1": hi, "step_cm1": step},
               "calculation": {"mu_view": mu, "line_cutoff_cm1": meta["cutoff"],
                               "tile_width_cm1": meta["tile"]},
               "output": {"transmittance": True, "optical_depth": True, "optical_depth_layers": True}}
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            res = run_forward(sd, prof, cfg)
        assert sorted({str(x.message) for x in w}) == meta["warnings"]
        np.testing.assert_array_equal(res.wavenumber_cm1, z["out_nu"])
        ref = z["out_od_layers"]
        np.testing.assert_allclose(res.optical_depth_layers, ref, rtol=TAU_RTOL_FP64, atol=0)
        np.testing.assert_allclose(res.optical_depth, ref.sum(0), rtol=TAU_RTOL_FP64, atol=0)
        np.testing.assert_allclose(res.transmittance, z[f"out_T_mu{j}"], rtol=1e-8, atol=1e-300)


@pytest.mark.parametrize("name", ["co2_mini", "three_gas_edge", "tile1"])
def test_golden_through_the_seam(name):
    """The reference's seam called the way its callers call it
    (forward.py:148, config.py:253-259): the golden grid array itself, the
    layers, cutoff and tile width.  Same shape as the reference's od, same
    values to 1e-9, same warnings."""
    from pylblrtm_amd.seam import build_optical_depth
    say(f"case: golden seam {name}")
    z, sd, prof, meta = _golden(name)
    nu = z["out_nu"]
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        od = build_optical_depth(nu, sd, prof.build_layers(), meta["cutoff"], meta["tile"])
    assert sorted({str(x.message) for x in w}) == meta["warnings"]
    ref = z["out_od_layers"]
    assert od.shape == ref.shape == (prof.build_layers().n_layers, nu.size)
    np.testing.assert_allclose(od, ref, rtol=TAU_RTOL_FP64, atol=0)


def test_golden_mixed_transmittance():
    say("case: golden mixed co2_mini")
    z, sd, prof, meta = _golden("co2_mini")
    lo, hi, step = meta["grid"]
    for j, mu in enumerate(meta["mus"]):
        nu, T = compute_transmittance(sd, prof, lo, hi, step_cm1=step, line_cutoff_cm1=meta["cutoff"],
                                      tile_width_cm1=meta["tile"], mu_view=mu, precision="mixed")
        assert np.max(np.abs(T - z[f"out_T_mu{j}"])) <= DT_ATOL_MIXED


def test_golden_error_message():
    say("case: golden nan strength")
    z, sd, prof, meta = _golden("nan_strength")
    lo, hi, step = meta["grid"]
    for prec in ("fp64", "mixed"):
        with pytest.raises(FloatingPointError) as e:
            compute_transmittance(sd, prof, lo, hi, step_cm1=step, line_cutoff_cm1=meta["cutoff"],
                                  tile_width_cm1=meta["tile"], precision=prec)
        assert ["FloatingPointError", str(e.value)] == meta["error"]


# ── seeded workloads against the oracle ──────────────────────────────────────

def _oracle_tau(sd, layers, nu_min, nu_max, step, cutoff, tile, tiles=None):
    nu = O.grid(nu_min, nu_max, step)
    od = O.optical_depth(nu, sd.lines, sd.tips, sd.molar_masses, layers.pressure_atm,
                         layers.temperature_k, layers.vmr, layers.column_cm2, cutoff, tile,
                         tiles=tiles)
    return nu, od


def test_co2_window_subset_vs_oracle():
    say("case: co2 window 500 lines, 20 layers, 715-720 @ 0.001")
    sd, prof = synth.co2_window(n_lines=500)
    lay = prof.build_layers()
    nu, od = _oracle_tau(sd, lay, 715.0, 720.0, 0.001, 25.0, 2.0)
    tau_ref = od.sum(0)
    prob = OpticalDepthProblem(sd, lay, 715.0, 720.0, 0.001, 25.0, 2.0)
    o64 = prob.run("fp64", mus=(1.0, 0.5), want_tau=True, want_layers=True)
    tau = o64["optical_depth"].cpu().numpy()
    assert np.max(np.abs(tau - tau_ref) / tau_ref) < TAU_RTOL_FP64
    np.testing.assert_allclose(o64["optical_depth_layers"].cpu().numpy(), od, rtol=TAU_RTOL_FP64)
    T1 = o64["transmittance"][0].cpu().numpy()
    T2 = o64["transmittance"][1].cpu().numpy()
    np.testing.assert_allclose(T1, np.exp(-tau_ref), rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose(T2, T1 * T1, rtol=1e-12, atol=1e-300)   # mu = 0.5 doubles the path
    om = prob.run("mixed", mus=(1.0, 0.5), want_tau=True)
    for m, mu in enumerate((1.0, 0.5)):
        Tm = om["transmittance"][m].double().cpu().numpy()
        assert np.max(np.abs(Tm - np.exp(-tau_ref / mu))) <= DT_ATOL_MIXED


def test_five_gas_tiles_vs_oracle():
    say("case: five gas 2000 lines/gas, 50 layers, 500-2500 @ 0.001, tiles 0/437/1000")
    sd, prof = synth.five_gas(lines_per_gas=2000)
    lay = prof.build_layers()
    tiles = [0, 437, 1000]
    nu, od = _oracle_tau(sd, lay, 500.0, 2500.0, 0.001, 25.0, 2.0, tiles=tiles)
    tau_ref = od.sum(0)
    m = ~np.isnan(tau_ref)
    prob = OpticalDepthProblem(sd, lay, 500.0, 2500.0, 0.001, 25.0, 2.0)
    o64 = prob.run("fp64", mus=(1.0,), want_tau=True)
    tau = o64["optical_depth"].cpu().numpy()
    rel = np.abs(tau[m] - tau_ref[m]) / tau_ref[m]
    assert rel.max() < TAU_RTOL_FP64, rel.max()
    om = prob.run("mixed", mus=(1.0, 0.5), want_tau=True)
    T64 = o64["transmittance"][0].cpu().numpy()
    Tm = om["transmittance"][0].double().cpu().numpy()
    assert np.max(np.abs(Tm - T64)) <= DT_ATOL_MIXED     # whole 2,000,001-point grid
    Tm5 = om["transmittance"][1].double().cpu().numpy()
    assert np.max(np.abs(Tm5 - np.exp(-2.0 * tau))) <= DT_ATOL_MIXED


def test_co2_window_full_size_tiles_vs_oracle():
    say("case: co2 window full size (5000 lines, 20 layers, 700-750 @ 0.001), tiles 0/12/25")
    sd, prof = synth.co2_window()
    lay = prof.build_layers()
    tiles = [0, 12, 25]
    nu, od = _oracle_tau(sd, lay, 700.0, 750.0, 0.001, 25.0, 2.0, tiles=tiles)
    m = ~np.isnan(od[0])
    prob = OpticalDepthProblem(sd, lay, 700.0, 750.0, 0.001, 25.0, 2.0)
    assert prob.n_nu == 50_001 and prob.tile_samples == 2000
    o64 = prob.run("fp64", mus=(1.0,), want_tau=True, want_layers=True)
    odl = o64["optical_depth_layers"].cpu().numpy()
    rel = np.abs(odl[:, m] - od[:, m]) / od[:, m]
    assert rel.max() < TAU_RTOL_FP64, rel.max()
    om = prob.run("mixed", mus=(1.0,), want_tau=True, want_layers=True)
    relm = np.abs(om["optical_depth_layers"].double().cpu().numpy()[:, m] - od[:, m]) / od[:, m]
    assert relm.max() < 2e-5, relm.max()


def test_dense_lines_queue_overflow_vs_oracle():
    """400 lines within 0.01 cm-1 on a 1e-5 grid, 2 layers: nearly every
    thread of both workgroups is near-centre for nearly every line, tens of
    thousands of (record, thread) pairs per chunk against queue capacities of
    8192 (mixed) and 3840 (FP64), so the queue-full path evaluates in line."""
    say("case: 400 lines in 0.01 cm-1, step 1e-5, 2 layers (near-centre queue overflow)")
    rng = np.random.default_rng(17)
    lines = synth.random_lines(rng, 400, 700.000, 700.010, 2)
    sd = SpectralData(lines=lines, tips={(2, 1): synth.tips_table()}, molar_masses={(2, 1): 43.99})
    p, t, _ = synth.levels(3)
    lay = AtmosphericProfile(p, t, {"CO2": np.full(3, 4.15e-4)}).build_layers()
    nu, od = _oracle_tau(sd, lay, 699.99, 700.02, 1e-5, 25.0, 2.0)
    prob = OpticalDepthProblem(sd, lay, 699.99, 700.02, 1e-5, 25.0, 2.0)
    assert prob.n_nu == 3001 and prob.tile_samples == 200_000
    o64 = prob.run("fp64", mus=(1.0,), want_layers=True)["optical_depth_layers"].cpu().numpy()
    rel = np.abs(o64 - od) / od
    say(f"  fp64 max rel tau {rel.max():.2e}")
    assert rel.max() < TAU_RTOL_FP64
    om = prob.run("mixed", mus=(1.0,), want_layers=True)["optical_depth_layers"]
    relm = np.abs(om.double().cpu().numpy() - od) / od
    say(f"  mixed max rel tau {relm.max():.2e}")
    assert relm.max() < 2e-5


def test_extreme_line_widths_vs_oracle():
    """Lorentz widths from 1e-3x to 100x the usual (y from 1e-4 to 1e4 over
    10 layers): every far-field tier of both precisions (shifted one-node,
    the two-node band of broad lines, the FP64 second-order and four-node
    tiers, the work queue) and each boundary between them falls inside the
    window for some line."""
    say("case: 300 lines, widths x1e-3..x100, 10 layers, 700-702 @ 0.001")
    rng = np.random.default_rng(23)
    n = 300
    lines = synth.random_lines(rng, n, 675.0, 727.0, 2)
    mult = rng.choice([1e-3, 0.1, 1.0, 3.0, 10.0, 100.0], n)
    lines["g_air"] = lines["g_air"] * mult
    lines["g_self"] = lines["g_self"] * mult
    sd = SpectralData(lines=lines, tips={(2, 1): synth.tips_table()}, molar_masses={(2, 1): 43.99})
    p, t, _ = synth.levels(11)
    lay = AtmosphericProfile(p, t, {"CO2": np.full(11, 4.15e-4)}).build_layers()
    nu, od = _oracle_tau(sd, lay, 700.0, 702.0, 0.001, 25.0, 2.0)
    prob = OpticalDepthProblem(sd, lay, 700.0, 702.0, 0.001, 25.0, 2.0)
    o64 = prob.run("fp64", mus=(1.0,), want_layers=True)["optical_depth_layers"].cpu().numpy()
    rel = np.abs(o64 - od) / od
    say(f"  fp64 max rel tau per layer {rel.max():.2e}")
    assert rel.max() < TAU_RTOL_FP64
    om = prob.run("mixed", mus=(1.0,), want_layers=True)["optical_depth_layers"]
    relm = np.abs(om.double().cpu().numpy() - od) / od
    say(f"  mixed max rel tau per layer {relm.max():.2e}")
    assert relm.max() < 2e-5


@pytest.mark.parametrize("lo,hi,step,ts", [(2400.0, 2406.0, 0.001, 1999),
                                           (2400.0, 2401.0, 1e-4, 19999)])
def test_grid_origin_tile_sizes_vs_oracle(lo, hi, step, ts):
    """Tile size depends on the grid origin (forward.py:175-176): 1999 points
    at 2400 cm-1 and step 0.001 (tiles just under one 2048-point slice), 19999
    at step 1e-4 (ten slices per tile, the last one partial)."""
    say(f"case: 400 lines, {lo}-{hi} @ {step}, 5 layers (tile_samples {ts})")
    rng = np.random.default_rng(29)
    lines = synth.random_lines(rng, 400, lo - 27.0, hi + 27.0, 2)
    sd = SpectralData(lines=lines, tips={(2, 1): synth.tips_table()}, molar_masses={(2, 1): 43.99})
    p, t, _ = synth.levels(6)
    lay = AtmosphericProfile(p, t, {"CO2": np.full(6, 4.15e-4)}).build_layers()
    nu, od = _oracle_tau(sd, lay, lo, hi, step, 25.0, 2.0)
    prob = OpticalDepthProblem(sd, lay, lo, hi, step, 25.0, 2.0)
    assert prob.tile_samples == ts and prob.n_nu == len(nu)
    o64 = prob.run("fp64", mus=(1.0,), want_layers=True)["optical_depth_layers"].cpu().numpy()
    rel = np.abs(o64 - od) / od
    say(f"  fp64 max rel tau per layer {rel.max():.2e}")
    assert rel.max() < TAU_RTOL_FP64
    om = prob.run("mixed", mus=(1.0,), want_layers=True)["optical_depth_layers"]
    relm = np.abs(om.double().cpu().numpy() - od) / od
    say(f"  mixed max rel tau per layer {relm.max():.2e}")
    assert relm.max() < 2e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_window_without_lines_and_zero_columns(precision):
    """No line reaches the window (the reference returns od = 0, forward.py:245-246),
    and a gas whose mixing ratio is 0 in every layer (each group skipped,
    forward.py:253-255): T = 1 exactly."""
    say(f"case: empty window and zero columns ({precision})")
    sd, prof = synth.co2_window(n_lines=200)
    nu, T = compute_transmittance(sd, prof, 900.0, 901.0, precision=precision)
    assert nu.size == 1001 and np.all(T == 1.0)
    zero = AtmosphericProfile(prof.pressure_hpa, prof.temperature_k,
                              {"CO2": np.zeros_like(prof.pressure_hpa)})
    nu, T = compute_transmittance(sd, zero, 720.0, 722.0, precision=precision)
    assert nu.size == 2001 and np.all(T == 1.0)


def test_many_layers_per_workgroup_vs_oracle():
    """200 layers on a 1,000,001-point grid: 500 tile slices, so the mixed
    sum runs 132 layer groups and the FP64 sum 33 (workgroups loop over 1-2
    and 6-7 layers, carrying their column totals from layer to layer).  FP64
    against the oracle on two tiles, mixed against FP64 on the whole grid."""
    say("case: 500 lines, 200 layers, 700-1700 @ 0.001, tiles 0/250")
    rng = np.random.default_rng(31)
    lines = synth.random_lines(rng, 500, 680.0, 1720.0, 2)
    sd = SpectralData(lines=lines, tips={(2, 1): synth.tips_table()}, molar_masses={(2, 1): 43.99})
    p, t, _ = synth.levels(201)
    lay = AtmosphericProfile(p, t, {"CO2": np.full(201, 4.15e-4)}).build_layers()
    tiles = [0, 250]
    nu, od = _oracle_tau(sd, lay, 700.0, 1700.0, 0.001, 25.0, 2.0, tiles=tiles)
    m = ~np.isnan(od[0])
    prob = OpticalDepthProblem(sd, lay, 700.0, 1700.0, 0.001, 25.0, 2.0)
    assert prob.n_nu == 1_000_001 and prob.n_layers == 200
    o64 = prob.run("fp64", mus=(1.0,), want_tau=True, want_layers=True)
    odl = o64["optical_depth_layers"].cpu().numpy()
    rel = np.abs(odl[:, m] - od[:, m]) / od[:, m]
    say(f"  fp64 max rel tau per layer {rel.max():.2e}")
    assert rel.max() < TAU_RTOL_FP64
    tau64 = o64["optical_depth"].cpu().numpy()
    np.testing.assert_allclose(tau64[m], od[:, m].sum(0), rtol=TAU_RTOL_FP64)
    om = prob.run("mixed", mus=(1.0, 0.5), want_tau=True)
    taum = om["optical_depth"].double().cpu().numpy()
    relm = np.abs(taum - tau64) / tau64
    say(f"  mixed max rel column tau (whole grid) {relm.max():.2e}")
    assert relm.max() < 2e-5
    for k, mu in enumerate((1.0, 0.5)):
        Tm = om["transmittance"][k].double().cpu().numpy()
        assert np.max(np.abs(Tm - np.exp(-tau64 / mu))) <= DT_ATOL_MIXED


def test_cutoff_below_near_field_radius_vs_oracle():
    """Cutoff 0.5 cm-1, below the mixed sum's near-field radius (3 x 256 grid
    steps = 0.768 cm-1 at step 0.001): near-field records reach the edges of
    the tile window, whose exact inclusion test must still hold."""
    say("case: 400 lines, 700-704 @ 0.001, cutoff 0.5, 5 layers")
    rng = np.random.default_rng(37)
    lines = synth.random_lines(rng, 400, 699.0, 705.0, 2)
    sd = SpectralData(lines=lines, tips={(2, 1): synth.tips_table()}, molar_masses={(2, 1): 43.99})
    p, t, _ = synth.levels(6)
    lay = AtmosphericProfile(p, t, {"CO2": np.full(6, 4.15e-4)}).build_layers()
    nu, od = _oracle_tau(sd, lay, 700.0, 704.0, 0.001, 0.5, 2.0)
    prob = OpticalDepthProblem(sd, lay, 700.0, 704.0, 0.001, 0.5, 2.0)
    pos = od > 0
    say(f"  points outside every line's reach: {np.count_nonzero(~pos)} of {od.size}")
    o64 = prob.run("fp64", mus=(1.0,), want_layers=True)["optical_depth_layers"].cpu().numpy()
    assert np.all(o64[~pos] == 0.0)
    rel = np.abs(o64[pos] - od[pos]) / od[pos]
    say(f"  fp64 max rel tau per layer {rel.max():.2e}")
    assert rel.max() < TAU_RTOL_FP64
    om = prob.run("mixed", mus=(1.0,), want_layers=True)["optical_depth_layers"].double().cpu().numpy()
    assert np.all(om[~pos] == 0.0)
    big = od > 1e-30   # above FP32's normal range
    relm = np.abs(om[big] - od[big]) / od[big]
    say(f"  mixed max rel tau per layer {relm.max():.2e}")
    assert relm.max() < 2e-5


# ── device-side index checks ─────────────────────────────────────────────────

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_out_of_range_indices_return_einval(precision):
    """A line's group outside [0, n_groups), a group's gas row outside
    [0, n_gases) and an empty TIPS range are flagged by the per-(layer, line)
    kernel, which then reads nothing through them; the call returns
    LBL_EINVAL (ValueError) and the same problem runs again once repaired."""
    say(f"case: invalid device indices {precision}")
    sd, prof = synth.co2_window(n_lines=300, n_levels=5)
    prob = OpticalDepthProblem(sd, prof.build_layers(), 720.0, 721.0, 0.001, 25.0, 2.0)
    good = prob.run(precision, mus=(1.0,), want_tau=True)["optical_depth"].double().cpu()
    cases = [("group", 17, 99, "group index"), ("group", 3, -1, "group index"),
             ("gas", 0, 5, "gas index"), ("toff", 1, 0, "TIPS")]
    for key, pos, bad, msg in cases:
        keep = int(prob._t[key][pos].item())
        prob._t[key][pos] = bad
        with pytest.raises(ValueError, match=msg):
            prob.run(precision, mus=(1.0,))
        prob._t[key][pos] = keep
        again = prob.run(precision, mus=(1.0,), want_tau=True)["optical_depth"].double().cpu()
        assert float(((again - good).abs() / good).max()) <= 1e-6


def test_out_of_range_group_in_vjp_returns_einval():
    import torch
    say("case: invalid device indices vjp")
    sd, prof = synth.co2_window(n_lines=300, n_levels=5)
    prob = OpticalDepthProblem(sd, prof.build_layers(), 720.0, 721.0, 0.001, 25.0, 2.0)
    fw = prob.run("mixed", mus=(1.0,), want_layers=True)
    G = torch.ones_like(fw["transmittance"])
    prob._t["group"][5] = 42
    with pytest.raises(ValueError, match="group index"):
        prob.vjp(fw["transmittance"], fw["optical_depth_layers"], G, (1.0,))
    prob._t["group"][5] = 0
    gt, gv = prob.vjp(fw["transmittance"], fw["optical_depth_layers"], G, (1.0,))
    assert torch.isfinite(gt).all() and torch.isfinite(gv).all()

"""The work counts behind the reported rates.

* ``OpticalDepthProblem.count_evaluations`` (the reference-equivalent
  evaluation count, the numerator of the headline rate) equals the oracle's
  own loop count ``stats['evals']`` (itself pinned by brute force on the CPU,
  tests/test_oracle.py).
* ``OpticalDepthProblem.count_series_work`` (the host's prediction of how the
  line-sum kernels split that work: point evaluations of the near ranges,
  series pairs per level, (line, group) pairs of the mixed group level)
  equals what the kernels count in their counting variant, up to the pairs
  the kernels themselves find the one-node form cannot take (counted
  separately), and the counting variant leaves the results unchanged.

Needs a ROCm GPU."""

import sys

import numpy as np
import pytest

from oracle import lbl_oracle as O
from pylblrtm_amd import AtmosphericProfile, SpectralData, synth
from pylblrtm_amd.device import OpticalDepthProblem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _golden(name):
    import os
    from golden_io import unpack_case
    here = os.path.dirname(os.path.abspath(__file__))
    z = np.load(os.path.join(here, "golden", f"{name}.npz"))
    lines, tips, mm, p, t, vmr, meta = unpack_case(z)
    return SpectralData(lines=lines, tips=tips, molar_masses=mm), AtmosphericProfile(p, t, vmr), meta


@pytest.mark.parametrize("name", ["co2_mini", "three_gas_edge", "tile1"])
def test_evaluation_count_equals_oracle_loop_count_golden(name):
    import warnings
    say(f"case: evaluation count, golden {name}")
    sd, prof, meta = _golden(name)
    lay = prof.build_layers()
    lo, hi, step = meta["grid"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        prob = OpticalDepthProblem(sd, lay, lo, hi, step, meta["cutoff"], meta["tile"])
        st = {}
        O.optical_depth(O.grid(lo, hi, step), sd.lines, sd.tips, sd.molar_masses, lay.pressure_atm,
                        lay.temperature_k, lay.vmr, lay.column_cm2, meta["cutoff"], meta["tile"],
                        stats=st)
    assert prob.count_evaluations() == st["evals"] > 0


def test_evaluation_count_equals_oracle_loop_count_five_gas_sample():
    say("case: evaluation count, five-gas sample (2,000 lines/gas; tiles 0, 500, 1000; layers 20-29)")
    sd, prof = synth.five_gas(lines_per_gas=2000)
    lay = prof.build_layers()
    prob = OpticalDepthProblem(sd, lay, 500.0, 2500.0, 0.001, 25.0, 2.0)
    ks = list(range(20, 30))
    sub = lambda a: np.asarray(a)[ks]
    tiles = [0, 500, 1000]
    st = {}
    O.optical_depth(prob.nu, sd.lines, sd.tips, sd.molar_masses, sub(lay.pressure_atm),
                    sub(lay.temperature_k), {g: sub(v) for g, v in lay.vmr.items()},
                    {g: sub(v) for g, v in lay.column_cm2.items()}, 25.0, 2.0, tiles=tiles, stats=st)
    assert prob.count_evaluations(layers=ks, tiles=tiles) == st["evals"] > 0


def _check_kernel_counts(prob, precision):
    import torch
    plain = prob.run(precision, mus=(1.0,), want_tau=True)["optical_depth"].double().cpu().numpy()
    kw = prob.kernel_work(precision)
    counted = prob.run(precision, mus=(1.0,), want_tau=True)["optical_depth"].double().cpu().numpy()
    # a counting run leaves the counts of the next plain run at zero
    from pylblrtm_amd import _capi
    assert all(v == 0 for v in _capi.last_work_count().values())
    hw = prob.count_series_work(precision)
    say(f"  kernels {kw}")
    say(f"  host    {hw}")
    assert kw["rescans"] == 0
    # the near field evaluates the lines in each group's own near range, and
    # those the group series left to it (its form fails at some point of the
    # group: one-node and middle tier mixed, four-node FP64); the group level
    # sees every other line of the wavefront's near range once per group
    assert kw["near_evals"] == hw["near_evals"] + kw["group_deferred_evals"]
    assert kw["group_series"] + kw["group_deferred"] == hw["group_pairs"]
    assert (kw["group_deferred"] == 0) == (kw["group_deferred_evals"] == 0)
    if precision == "fp64":
        assert kw["group_series_middle"] == 0
    assert kw["slice_series"] + kw["slice_exceptions"] == hw["slice_pairs"]
    # the wavefront level sees the records between the two near ranges, plus
    # every record the slice level passed on, once per live wavefront
    assert kw["wave_series"] + kw["wave_exceptions"] == hw["wave_pairs"] + kw["wave_from_slice_list"]
    assert kw["wave_from_slice_list"] >= kw["slice_exceptions"]
    assert (kw["slice_exceptions"] == 0) == (kw["wave_from_slice_list"] == 0)
    # the counting variant computes the same optical depths
    tol = 1e-12 if precision == "fp64" else 1e-5
    pos = plain > 0
    assert np.max(np.abs(counted[pos] - plain[pos]) / plain[pos]) <= tol
    torch.cuda.synchronize()
    return kw


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_kernel_counts_equal_host_prediction_co2_window(precision):
    say(f"case: kernel work counts, CO2 window (5,000 lines, 20 layers, 700-750 @ 0.001) {precision}")
    sd, prof = synth.co2_window()
    prob = OpticalDepthProblem(sd, prof.build_layers(), 700.0, 750.0, 0.001, 25.0, 2.0)
    kw = _check_kernel_counts(prob, precision)
    assert kw["near_evals"] > 0 and kw["slice_series"] > 0 and kw["wave_series"] > 0
    if precision == "mixed":
        assert kw["group_series"] > 0


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_kernel_counts_equal_host_prediction_five_gas(precision):
    say(f"case: kernel work counts, five gases (2,000 lines/gas, 50 layers, 2,000,001 points) {precision}")
    sd, prof = synth.five_gas(lines_per_gas=2000)
    prob = OpticalDepthProblem(sd, prof.build_layers(), 500.0, 2500.0, 0.001, 25.0, 2.0)
    _check_kernel_counts(prob, precision)


def test_kernel_counts_on_origin_dependent_tiles():
    say("case: kernel work counts, 1999-point tiles at 2400 cm-1 and step 5e-4 (3999-point tiles)")
    sd, prof = synth.co2_window(n_lines=3000, nu_lo=2390.0, nu_hi=2420.0)
    lay = prof.build_layers()
    for lo, hi, step in ((2400.0, 2410.0, 0.001), (1200.0, 1205.0, 0.0005)):
        sdx = sd
        if lo < 2000:
            lines = dict(sd.lines)
            lines["nu0"] = np.asarray(sd.lines["nu0"]) - 1195.0
            sdx = SpectralData(lines=lines, tips=sd.tips, molar_masses=sd.molar_masses)
        prob = OpticalDepthProblem(sdx, lay, lo, hi, step, 25.0, 2.0)
        for precision in ("mixed", "fp64"):
            _check_kernel_counts(prob, precision)


def test_kernel_counts_with_far_field_exceptions():
    """Broad H2O-like lines (g_air, g_self 0.4-0.7 cm-1/atm) near a 0.6 cm-1
    window at 1e-4: in the lower layers many records outside a slice's (or
    wavefront's) near range are too near for its series form to hold at all
    its points; they pass to the wavefront level or are evaluated point by
    point there, and the identities hold with those exception counts (300
    lines keep each slice's list of them below its 1,024 entries)."""
    say("case: kernel work counts with far-field exceptions (broad lines at 1e-4)")
    rng = np.random.default_rng(17)
    lines = synth.random_lines(rng, 300, 2497.5, 2502.5, 1)
    lines["g_air"] = rng.uniform(0.4, 0.7, 300)
    lines["g_self"] = rng.uniform(0.4, 0.7, 300)
    sd = SpectralData(lines=lines, tips={(1, 1): synth.tips_table(174.6)},
                      molar_masses={(1, 1): 18.01})
    prof = AtmosphericProfile(np.array([10.0, 100.0, 300.0, 600.0, 1000.0]),
                              np.array([215.0, 225.0, 245.0, 270.0, 290.0]),
                              {"H2O": np.array([2e-6, 8e-6, 4e-4, 3.6e-3, 1.2e-2])})
    prob = OpticalDepthProblem(sd, prof.build_layers(), 2499.7, 2500.3, 1e-4, 25.0, 2.0)
    for precision in ("mixed", "fp64"):
        kw = _check_kernel_counts(prob, precision)
        assert kw["slice_exceptions"] > 0 and kw["wave_exceptions"] > 0
        assert kw["exception_evals"] >= kw["wave_exceptions"]
        if precision == "mixed":   # broad lines: the group series cannot take all
            assert kw["group_deferred"] > 0

"""Pins of the CPU oracle (oracle/lbl_oracle.py) by closed forms, the
reference's own test values, published known-answer values and brute force.
CPU only."""

import json
import math
import os

import numpy as np
import pytest
from scipy.special import erfcx, wofz

from oracle import lbl_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ── grid / tile_samples (forward.py:141-142,175-176) ─────────────────────────

@pytest.mark.parametrize("nu_min,step,ts", [
    (500.0, 0.001, 2000), (700.0, 0.001, 2000), (1000.0, 0.001, 2000), (2400.0, 0.001, 1999),
    (1200.0, 0.0005, 3999), (1450.0, 0.0005, 3999), (2000.0, 0.0005, 3999),
    (2400.0, 0.0005, 3999), (2400.0, 1e-4, 19999),
])
def test_tile_samples_depends_on_grid_origin(nu_min, step, ts):
    nu = nu_min + np.arange(3, dtype=np.float64) * step
    assert O.tile_samples(nu, 2.0) == ts


def test_tile_samples_single_point_uses_default_step():
    assert O.tile_samples(np.array([700.0]), 2.0) == 2000


def test_grid_closed_form():
    nu = O.grid(500.0, 2500.0, 0.001)
    assert nu.size == 2_000_001
    assert nu[0] == 500.0 and nu[1234] == 500.0 + 1234.0 * 0.001
    nu = O.grid(700.0, 700.0105, 0.001)   # non-integral ratio: last point passes nu_max
    assert nu.size == 12 and nu[-1] > 700.0105


# ── layers (atmosphere.py:110-141) ────────────────────────────────────────────

def test_build_layers_hand_case():
    p = np.array([500.0, 1000.0])
    t = np.array([250.0, 290.0])
    pa, tl, vl, col = O.build_layers(p, t, {"X": np.array([0.01, 0.03])})
    assert pa[0] == pytest.approx(math.sqrt(500.0 * 1000.0) / 1013.25, rel=1e-15)
    assert tl[0] == 270.0 and vl["X"][0] == pytest.approx(0.02, rel=1e-15)
    per_pa = 6.02214086e23 / (9.80665 * 0.0289644)
    assert col["X"][0] == pytest.approx(0.02 * per_pa * 50000.0 / 1e4, rel=1e-14)


def test_full_column_constant():
    _, _, _, col = O.build_layers(np.array([1e-9, 1013.25]), np.array([250.0, 250.0]),
                                  {"air": np.ones(2)})
    assert col["air"][0] == pytest.approx(2.14824e25, rel=2e-6)


# ── partition sums (spectral_data.py:43-52; test_spectral_data.py:72-75) ─────

def test_partition_sum_reference_value():
    T = np.arange(150.0, 401.0)
    assert abs(O.partition_sum(200.0, T, 0.5 * T) - 100.0) < 1e-6


def test_partition_sum_clamps_like_interp():
    T = np.arange(70.0, 401.0)
    Q = 286.1 * (T / 296.0) ** 1.5
    assert O.partition_sum(10.0, T, Q) == Q[0]
    assert O.partition_sum(1000.0, T, Q) == Q[-1]
    assert O.partition_sum(296.0, T, Q) == pytest.approx(286.1, rel=1e-15)


# ── per-(layer, line) scaling (forward.py:257-292) ────────────────────────────

def _lines(n=7, seed=0):
    rng = np.random.default_rng(seed)
    return {"nu0": rng.uniform(700, 750, n), "Sref": 10 ** rng.uniform(-27, -19, n),
            "g_air": rng.uniform(0.05, 0.1, n), "g_self": rng.uniform(0.07, 0.12, n),
            "Elow": rng.uniform(0, 3000, n), "n_air": rng.uniform(0.6, 0.8, n),
            "shift": rng.normal(-0.002, 0.001, n)}


def test_line_params_reference_temperature_special_case():
    L = _lines()
    idx = np.arange(7)
    P, col = 0.7, 3.0e21
    c, S, gl, ad = O.line_params(L, idx, 296.0, P, 0.0, col, 1.0, 43.99)
    np.testing.assert_array_equal(c, L["nu0"] + L["shift"] * P)
    R = c * np.tanh(0.5 * 1.4387752 * c / 296.0)
    np.testing.assert_allclose(S, L["Sref"] * col / R, rtol=1e-14)
    np.testing.assert_allclose(gl, L["g_air"] * P, rtol=1e-15)
    np.testing.assert_allclose(ad, L["nu0"] * 3.5810e-7 * math.sqrt(296.0 / 43.99), rtol=1e-15)


def test_line_params_linear_in_sref_and_column():
    L = _lines()
    idx = np.arange(7)
    a = O.line_params(L, idx, 231.0, 0.3, 1e-3, 1e21, 1.3, 18.01)[1]
    L2 = dict(L, Sref=3.0 * L["Sref"])
    b = O.line_params(L2, idx, 231.0, 0.3, 1e-3, 2e21, 1.3, 18.01)[1]
    np.testing.assert_allclose(b, 6.0 * a, rtol=1e-14)


def test_self_pressure_enters_width_and_shift():
    L = _lines()
    idx = np.arange(7)
    c, _, gl, _ = O.line_params(L, idx, 296.0, 1.0, 0.25, 1e21, 1.0, 44.0)
    np.testing.assert_allclose(c, L["nu0"] + L["shift"] * 0.75, rtol=1e-15)
    np.testing.assert_allclose(gl, L["g_air"] * 0.75 + L["g_self"] * 0.25, rtol=1e-15)


# ── Voigt / Faddeeva (forward.py:51-62) ───────────────────────────────────────

def test_wofz_known_answers():
    kat = json.load(open(os.path.join(HERE, "golden", "wofz_kat.json")))
    z = np.array([complex(*v) for v in kat["z"]])
    w = np.array([complex(*v) for v in kat["w"]])
    np.testing.assert_allclose(wofz(z), w, rtol=1e-13)


def test_voigt_limits():
    x = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(wofz(x + 0j).real, np.exp(-x * x), rtol=1e-13, atol=1e-300)
    y = np.geomspace(1e-4, 1e2, 50)
    np.testing.assert_allclose(wofz(1j * y).real, erfcx(y), rtol=1e-13)
    xf = np.geomspace(3e3, 1e5, 20)          # next term is 1.5/x^2 relative
    np.testing.assert_allclose(wofz(xf + 0.5j).real, 0.5 / (math.sqrt(math.pi) * (xf * xf + 0.25)),
                               rtol=1e-6)


def test_voigt_area_is_one_minus_wing_mass():
    nu = np.linspace(-40.0, 40.0, 800_001)
    phi = O.voigt(nu, np.array([0.0]), np.array([0.01]), np.array([0.02]))[:, 0]
    area = np.trapezoid(phi, nu)
    wing = 2.0 * 0.02 / math.pi / 40.0          # Lorentz mass beyond +-40 (1st order)
    assert area == pytest.approx(1.0 - wing, rel=2e-5)


# ── tile inclusion (forward.py:295-304) ───────────────────────────────────────

def test_one_point_tiles_equal_per_point_cutoff():
    rng = np.random.default_rng(5)
    n = 40
    L = _lines(n, seed=5)
    L["nu0"] = rng.uniform(699.6, 701.4, n)
    L["mol_id"] = np.full(n, 2, np.int32)
    L["iso_id"] = np.full(n, 1, np.int32)
    T = np.arange(70.0, 401.0)
    tips = {(2, 1): (T, 286.1 * (T / 296) ** 1.5)}
    mm = {(2, 1): 43.99}
    pa, tl, vl, col = O.build_layers(np.array([100.0, 500.0]), np.array([230.0, 260.0]),
                                     {"CO2": np.full(2, 4e-4)})
    nu = O.grid(700.0, 701.0, 0.01)
    cut = 0.3
    od = O.optical_depth(nu, L, tips, mm, pa, tl, vl, col, cut, 0.001)
    # brute force, point by point
    g = O.prepare_groups(nu, L, tips, mm, pa, tl, ["CO2"], cut)[0]
    c, S, gl, ad = O.line_params(L, g["idx"], tl[0], pa[0], vl["CO2"][0], col["CO2"][0],
                                 g["q_ratio"][0], 43.99)
    ref = np.zeros_like(nu)
    for i, v in enumerate(nu):
        m = (c >= v - cut) & (c <= v + cut)
        if m.any():
            ref[i] = (O.voigt(np.array([v]), c[m], ad[m], gl[m]) @ S[m])[0]
    ref *= O.radiation_term(nu, tl[0])
    np.testing.assert_allclose(od[0], ref, rtol=1e-13, atol=0)


def test_inclusion_edges_are_inclusive():
    # a line exactly at tile[0] - cut is in, one ulp lower is out
    nu = O.grid(700.0, 701.0, 0.5)       # tiles of one point with tile_width 0.1
    cut = 1.0
    L = {"nu0": np.array([699.0, np.nextafter(699.0, 0.0)]), "Sref": np.array([1e-20, 1e-20]),
         "g_air": np.array([0.07, 0.07]), "g_self": np.array([0.1, 0.1]), "Elow": np.zeros(2),
         "n_air": np.full(2, 0.7), "shift": np.zeros(2), "mol_id": np.array([2, 2], np.int32),
         "iso_id": np.array([1, 2], np.int32)}
    T = np.arange(70.0, 401.0)
    tips = {(2, 1): (T, T), (2, 2): (T, T)}
    mm = {(2, 1): 44.0, (2, 2): 44.0}
    pa, tl, vl, col = O.build_layers(np.array([100.0, 500.0]), np.array([296.0, 296.0]),
                                     {"CO2": np.full(2, 4e-4)})
    st = {}
    O.optical_depth(nu, L, tips, mm, pa, tl, vl, col, cut, 0.1, stats=st)
    # first tile (700.0) sees line 1 only; the other tiles see neither
    assert st["evals"] == 1


def test_evaluation_count_brute_force():
    """stats['evals'] (the work unit of the headline rate) against a plain
    loop over (layer, group, tile, line) with the reference's tile-level
    rule (forward.py:295-304): a line counts for every point of a tile when
    its shifted centre lies in [tile first - cut, tile last + cut]; groups
    with a zero column are skipped (forward.py:253-255).  Tiles of 7 points
    over a 25-point grid (a short last tile), two gases, one of them absent
    from a layer, centres on and one ulp outside a tile edge."""
    rng = np.random.default_rng(11)
    n = 30
    L = _lines(n, seed=11)
    L["nu0"] = rng.uniform(699.5, 700.8, n)
    nu = O.grid(700.0, 700.24, 0.01)                # 25 points
    cut = 0.2
    # a CO2 line whose shifted centre is exactly tile 0's first point - cut
    # (it is in), and one an ulp below it (it is out) in the top layer
    L["shift"][:2] = 0.0
    L["nu0"][0] = nu[0] - cut
    L["nu0"][1] = np.nextafter(nu[0] - cut, 0.0)
    L["mol_id"] = np.where(np.arange(n) % 3 == 0, 1, 2).astype(np.int32)
    L["mol_id"][:2] = 2
    L["iso_id"] = np.ones(n, np.int32)
    T = np.arange(70.0, 401.0)
    tips = {(1, 1): (T, T), (2, 1): (T, 286.1 * (T / 296) ** 1.5)}
    mm = {(1, 1): 18.01, (2, 1): 43.99}
    vmr = {"CO2": np.full(4, 4e-4), "H2O": np.array([0.0, 0.0, 1e-3, 2e-3])}
    pa, tl, vl, col = O.build_layers(np.array([10.0, 100.0, 500.0, 1000.0]),
                                     np.array([220.0, 240.0, 270.0, 290.0]), vmr)
    assert col["H2O"][0] == 0.0 and col["H2O"][1] > 0.0   # layer 0 skips H2O
    st = {}
    O.optical_depth(nu, L, tips, mm, pa, tl, vl, col, cut, 0.07, stats=st)
    ts = 7
    expect = 0
    for k in range(len(pa)):
        for mol, gas in ((1, "H2O"), (2, "CO2")):
            if col[gas][k] == 0.0:
                continue
            for t0 in range(0, nu.size, ts):
                t1 = min(t0 + ts, nu.size)
                for i in range(n):
                    if L["mol_id"][i] != mol:
                        continue
                    p_self = vl[gas][k] * pa[k]
                    c = L["nu0"][i] + L["shift"][i] * (pa[k] - p_self)
                    if nu[t0] - cut <= c <= nu[t1 - 1] + cut:
                        expect += t1 - t0
    assert O.tile_samples(nu, 0.07) == ts
    assert st["evals"] == expect
    # the edge line is counted in every layer for tile 0, its neighbour never
    assert expect > 0


# ── integrated path vs the reference's own outputs ────────────────────────────

@pytest.mark.parametrize("name", ["co2_mini", "three_gas_edge", "tile1"])
def test_oracle_matches_reference_golden(name):
    import warnings
    from golden_io import unpack_case
    z = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    lines, tips, mm, p, t, vmr, meta = unpack_case(z)
    lo, hi, step = meta["grid"]
    nu = O.grid(lo, hi, step)
    np.testing.assert_array_equal(nu, z["out_nu"])
    pa, tl, vl, col = O.build_layers(p, t, vmr)
    np.testing.assert_array_equal(pa, z["out_lay_p_atm"])
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        od = O.optical_depth(nu, lines, tips, mm, pa, tl, vl, col, meta["cutoff"], meta["tile"])
    assert sorted({str(x.message) for x in w}) == meta["warnings"]
    ref = z["out_od_layers"]
    np.testing.assert_allclose(od, ref, rtol=1e-13, atol=0)
    for j, mu in enumerate(meta["mus"]):
        np.testing.assert_allclose(O.transmittance(od, mu), z[f"out_T_mu{j}"], rtol=1e-12, atol=1e-300)


def test_oracle_error_message_matches_reference():
    from golden_io import unpack_case
    z = np.load(os.path.join(HERE, "golden", "nan_strength.npz"))
    lines, tips, mm, p, t, vmr, meta = unpack_case(z)
    lo, hi, step = meta["grid"]
    nu = O.grid(lo, hi, step)
    pa, tl, vl, col = O.build_layers(p, t, vmr)
    with pytest.raises(FloatingPointError) as e:
        O.optical_depth(nu, lines, tips, mm, pa, tl, vl, col, meta["cutoff"], meta["tile"])
    assert meta["error"] == ["FloatingPointError", str(e.value)]


def test_planted_mistake_is_caught():
    """A dropped stimulated-emission factor in the oracle must break parity
    with the reference's outputs (the pins above are not vacuous)."""
    from golden_io import unpack_case
    z = np.load(os.path.join(HERE, "golden", "co2_mini.npz"))
    lines, tips, mm, p, t, vmr, meta = unpack_case(z)
    lo, hi, step = meta["grid"]
    nu = O.grid(lo, hi, step)
    pa, tl, vl, col = O.build_layers(p, t, vmr)
    orig = O.line_params

    def broken(*a, **k):
        c, S, gl, ad = orig(*a, **k)
        return c, S * 1.000001, gl, ad

    O.line_params = broken
    try:
        od = O.optical_depth(nu, lines, tips, mm, pa, tl, vl, col, meta["cutoff"], meta["tile"])
    finally:
        O.line_params = orig
    assert np.max(np.abs(od - z["out_od_layers"]) / z["out_od_layers"]) > 1e-7

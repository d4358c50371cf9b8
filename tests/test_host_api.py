"""Host-side API of pylblrtm_amd mirrors the reference (CPU only): data
containers, validation, config resolution, prelude, writers."""

import os
import warnings

import numpy as np
import pytest

from pylblrtm_amd import AtmosphericProfile, ForwardResult, SpectralData, load_config
from pylblrtm_amd import compute_transmittance, run_forward
from pylblrtm_amd.config import _write
from pylblrtm_amd.device import make_grid, prepare_groups, tile_samples_for

HERE = os.path.dirname(os.path.abspath(__file__))


def _make_lines(n, mol_id=1, iso_id=1):
    return {"nu0": np.linspace(700.0, 1400.0, n), "Sref": np.ones(n) * 1e-23,
            "g_air": np.ones(n) * 0.07, "g_self": np.ones(n) * 0.35, "Elow": np.ones(n) * 100.0,
            "n_air": np.ones(n) * 0.67, "shift": np.zeros(n),
            "mol_id": np.full(n, mol_id, np.int32), "iso_id": np.full(n, iso_id, np.int32)}


@pytest.fixture
def sd():
    T = np.arange(150.0, 401.0)
    return SpectralData(lines=_make_lines(20), tips={(1, 1): (T, 0.5 * T)},
                        molar_masses={(1, 1): 18.010565}, molecule="H2O")


# SpectralData (same checks as the reference's tests/test_spectral_data.py)

def test_spectral_data_properties(sd):
    assert sd.n_lines == 20 and len(sd.nu0) == 20 and len(sd.Sref) == 20


def test_Q_interpolates(sd):
    assert abs(sd.Q(200.0, 1, 1) - 100.0) < 1e-6
    assert isinstance(sd.Q(200.0, np.int32(1), np.int32(1)), float)


def test_Q_missing_key(sd):
    with pytest.raises(KeyError, match="mol_id=1, iso_id=99"):
        sd.Q(200.0, mol_id=1, iso_id=99)


def test_repr(sd):
    assert "H2O" in repr(sd) and "(1,1)" in repr(sd)
    sd2 = SpectralData(lines=_make_lines(5), tips=sd.tips, molar_masses={(1, 1): 18.0})
    assert "multi-molecule" in repr(sd2)


# AtmosphericProfile (atmosphere.py:69-100)

@pytest.mark.parametrize("p,t,vmr,msg", [
    ([1.0], [200.0], {}, "at least 2 levels"),
    ([1.0, -2.0], [200.0, 210.0], {}, "finite and positive"),
    ([2.0, 1.0], [200.0, 210.0], {}, "strictly increasing"),
    ([1.0, 2.0], [200.0], {}, "same length"),
    ([1.0, 2.0], [200.0, 0.0], {}, "temperature_k must be finite"),
    ([1.0, 2.0], [200.0, 210.0], {"CO2": [1e-4]}, "must have the same length"),
    ([1.0, 2.0], [200.0, 210.0], {"CO2": [1e-4, 1.5]}, "finite and in"),
])
def test_profile_validation(p, t, vmr, msg):
    with pytest.raises(ValueError, match=msg):
        AtmosphericProfile(p, t, vmr)


def test_build_layers_matches_reference_golden():
    from golden_io import unpack_case
    z = np.load(os.path.join(HERE, "golden", "three_gas_edge.npz"))
    lines, tips, mm, p, t, vmr, meta = unpack_case(z)
    lay = AtmosphericProfile(p, t, vmr).build_layers()
    np.testing.assert_array_equal(lay.pressure_atm, z["out_lay_p_atm"])
    np.testing.assert_array_equal(lay.temperature_k, z["out_lay_t_k"])
    for g in vmr:
        np.testing.assert_array_equal(lay.vmr[g], z[f"out_lay_vmr_{g}"])
        np.testing.assert_array_equal(lay.column_cm2[g], z[f"out_lay_col_{g}"])


# compute_transmittance validation (forward.py:129-138) happens before any device use

def test_compute_transmittance_validation(sd):
    prof = AtmosphericProfile([1.0, 10.0], [220.0, 240.0], {"H2O": [1e-5, 1e-4]})
    with pytest.raises(TypeError):
        compute_transmittance(object(), prof, 700, 710)
    with pytest.raises(TypeError):
        compute_transmittance(sd, object(), 700, 710)
    with pytest.raises(ValueError, match="strictly less"):
        compute_transmittance(sd, prof, 710, 700)
    with pytest.raises(ValueError, match="step_cm1"):
        compute_transmittance(sd, prof, 700, 710, step_cm1=0)
    with pytest.raises(ValueError, match="mu_view"):
        compute_transmittance(sd, prof, 700, 710, mu_view=0.0)


# config (config.py:23-124)

def test_load_config_defaults_and_merge():
    cfg = load_config({"spectral": {"nu_min_cm1": 710}})
    assert cfg["spectral"] == {"nu_min_cm1": 710, "nu_max_cm1": 750.0, "step_cm1": 0.001}
    assert cfg["calculation"]["tile_width_cm1"] == 2.0
    assert cfg["calculation"]["precision"] == "fp64"
    assert cfg["output"]["save"]["format"] == "npz"


def test_load_config_yaml(tmp_path):
    p = tmp_path / "c.yaml"
    p.write_text("spectral:\n  nu_min_cm1: 720.0\ncalculation:\n  mu_view: 0.5\n")
    cfg = load_config(str(p))
    assert cfg["spectral"]["nu_min_cm1"] == 720.0 and cfg["calculation"]["mu_view"] == 0.5
    with pytest.raises(FileNotFoundError):
        load_config(str(tmp_path / "missing.yaml"))


@pytest.mark.parametrize("override,msg", [
    ({"spectral": {"nu_min_cm1": 800}}, "strictly less"),
    ({"spectral": {"step_cm1": 0}}, "step_cm1"),
    ({"calculation": {"line_cutoff_cm1": 0}}, "line_cutoff"),
    ({"calculation": {"mu_view": 1.5}}, "mu_view"),
    ({"output": {"transmittance": False}}, "At least one output"),
    ({"output": {"save": {"format": "xlsx"}}}, "npz"),
    ({"calculation": {"precision": "fp16"}}, "precision"),
])
def test_config_validation(override, msg):
    with pytest.raises(ValueError, match=msg):
        load_config(override)


def test_run_forward_type_checks(sd):
    prof = AtmosphericProfile([1.0, 10.0], [220.0, 240.0], {"H2O": [1e-5, 1e-4]})
    with pytest.raises(TypeError):
        run_forward(object(), prof, {})
    with pytest.raises(TypeError):
        run_forward(sd, object(), {})


# writers (config.py:284-319) against the headers the reference wrote

def test_writer_headers_match_reference(tmp_path, capsys):
    import json
    z = np.load(os.path.join(HERE, "golden", "co2_mini.npz"))
    meta = json.loads(str(z["meta_json"]))
    nu = z["out_nu"]
    res = ForwardResult(wavenumber_cm1=nu, config=load_config({}),
                        transmittance=z["out_rf_transmittance"],
                        optical_depth=z["out_rf_optical_depth"],
                        optical_depth_layers=z["out_rf_optical_depth_layers"])
    for fmt in ("csv", "txt"):
        path = _write(res, str(tmp_path), fmt, "gold")
        assert path.read_text().splitlines()[0] == meta["run_forward"][fmt]
        back = np.loadtxt(path, delimiter="," if fmt == "csv" else "\t", skiprows=1)
        np.testing.assert_array_equal(back[:, 0], nu)
        np.testing.assert_array_equal(back[:, 1], z["out_rf_transmittance"])
    path = _write(res, str(tmp_path), "npz", "gold")
    with np.load(path) as f:
        assert sorted(f.files) == sorted(["wavenumber_cm1", "transmittance", "optical_depth",
                                          "optical_depth_layers"])
        np.testing.assert_array_equal(f["optical_depth_layers"], z["out_rf_optical_depth_layers"])
    out = capsys.readouterr().out
    assert out.splitlines()[0].replace(str(tmp_path), "<DIR>") == \
        meta["run_forward"]["stdout_csv"].strip()
    assert repr(res) == meta["repr"]


# prelude (forward.py:171-245)

def test_tile_samples_and_grid_like_reference():
    from golden_io import unpack_case
    for name in ("co2_mini", "three_gas_edge", "tile1"):
        z = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
        meta = unpack_case(z)[-1]
        lo, hi, step = meta["grid"]
        nu = make_grid(lo, hi, step)
        np.testing.assert_array_equal(nu, z["out_nu"])
    assert tile_samples_for(make_grid(2400.0, 2401.0, 0.001), 0.2) == 199


def test_prelude_warnings_match_reference():
    from golden_io import unpack_case
    z = np.load(os.path.join(HERE, "golden", "three_gas_edge.npz"))
    lines, tips, mm, p, t, vmr, meta = unpack_case(z)
    sdx = SpectralData(lines=lines, tips=tips, molar_masses=mm)
    lay = AtmosphericProfile(p, t, vmr).build_layers()
    lo, hi, step = meta["grid"]
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        groups = prepare_groups(make_grid(lo, hi, step), sdx, lay, meta["cutoff"])
    assert sorted({str(x.message) for x in w}) == meta["warnings"]
    assert [(g.mol_id, g.iso_id) for g in groups] == [(1, 1), (2, 1), (2, 2), (3, 1)]
    assert [g.molar_mass for g in groups] == [18.01, 43.99, 28.97, 47.98]


def test_prelude_rejects_tables_the_reference_rejects(tmp_path):
    """A TIPS table whose T and Q differ in length (what txt_to_tips returns
    for a row whose Q does not convert, hitran_manager.py:955-956) or that is
    empty fails where the reference first interpolates it, q_ref = Q(296)
    (forward.py:236): numpy's own ValueError, before any device work."""
    from pylblrtm_amd import synth
    from pylblrtm_amd.hitran_manager import txt_to_tips
    sd, prof = synth.co2_window(n_lines=50, n_levels=4)
    lay = prof.build_layers()
    p = tmp_path / "tips.txt"
    p.write_text("200 100\n296 x\n300 150\n", encoding="utf-8")
    t, q = txt_to_tips(p)
    assert (t.size, q.size) == (3, 2)
    for bad in ((t, q), (np.zeros(0), np.zeros(0))):
        sdx = SpectralData(lines=sd.lines, tips={(2, 1): bad}, molar_masses=sd.molar_masses)
        with pytest.raises(ValueError) as exc:
            prepare_groups(make_grid(700.0, 701.0, 0.001), sdx, lay, 25.0)
        with pytest.raises(ValueError) as ref_exc:
            np.interp(296.0, *bad)
        assert str(exc.value) == str(ref_exc.value)

"""The C-ABI library loads and exports every function include/*.h declares;
argument validation happens before any device call (CPU only)."""

import ctypes
import json
import re
from pathlib import Path

import pytest

from pylblrtm_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lbl_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    names = declared_functions()
    assert "lbl_optical_depth" in names and len(names) >= 4
    for n in names:
        assert hasattr(lib, n), n
    assert set(_capi.EXPORTED) <= set(names)


def test_last_error_is_empty_string_initially():
    assert isinstance(_capi.last_error(), str)


def test_invalid_arguments_rejected_without_device():
    lib = _capi.load()
    grid = _capi.LblGrid(700.0, 0.001, 100, 10, 25.0, 0.0)
    lines = _capi.LblLines(0, None, None, None, None, None, None, None, None)
    groups = _capi.LblGroups(0, None, None, None, None, None)
    layers = _capi.LblLayers(0, 0, None, None, None, None)
    mu = (ctypes.c_double * 1)(1.0)
    out = _capi.LblOutputs(1, mu, None, None, None)
    st = lib.lbl_optical_depth(ctypes.byref(lines), ctypes.byref(groups), ctypes.byref(layers),
                               ctypes.byref(grid), 0, ctypes.byref(out), None, 0, None, None, None)
    assert st == _capi.LBL_EINVAL and "transmittance" in _capi.last_error()
    st = lib.lbl_optical_depth(ctypes.byref(lines), ctypes.byref(groups), ctypes.byref(layers),
                               ctypes.byref(grid), 7, ctypes.byref(out), None, 0, None, None, None)
    assert st == _capi.LBL_EINVAL and "precision" in _capi.last_error()
    with pytest.raises(ValueError):
        _capi.raise_for(st, "x")


def test_workspace_size_is_host_arithmetic():
    lib = _capi.load()
    grid = _capi.LblGrid(500.0, 0.001, 2_000_001, 2000, 25.0, 0.01)
    lines = _capi.LblLines(500_000, None, None, None, None, None, None, None, None)
    layers = _capi.LblLayers(50, 5, None, None, None, None)
    n = ctypes.c_size_t(0)
    assert lib.lbl_workspace_bytes(ctypes.byref(lines), ctypes.byref(layers), ctypes.byref(grid),
                                   1, ctypes.byref(n)) == 0
    assert n.value >= 25_000_000 * 48          # one 48-byte record per (layer, line)
    m = ctypes.c_size_t(0)
    lib.lbl_workspace_bytes(ctypes.byref(lines), ctypes.byref(layers), ctypes.byref(grid), 0,
                            ctypes.byref(m))
    assert m.value >= 25_000_000 * 32          # FP64: four 8-byte arrays


def test_noalloc_entry_point_rejects_oversize_before_device_work():
    """lbl_optical_depth_noalloc checks the problem's scratch against its
    fixed array on the host: the five-gas workload (~1.7 GB) is refused with
    LBL_EINVAL, and no device call is made (none is possible here)."""
    lib = _capi.load()
    fake = ctypes.c_void_p(256)   # never dereferenced: the size check comes first
    lines = _capi.LblLines(500_000, fake, fake, fake, fake, fake, fake, fake, fake)
    groups = _capi.LblGroups(5, fake, fake, fake, fake, fake)
    layers = _capi.LblLayers(50, 5, fake, fake, fake, fake)
    grid = _capi.LblGrid(500.0, 0.001, 2_000_001, 2000, 25.0, 0.01)
    mu = (ctypes.c_double * 1)(1.0)
    out = _capi.LblOutputs(1, mu, fake, None, None)
    for prec in (0, 1):
        st = lib.lbl_optical_depth_noalloc(ctypes.byref(lines), ctypes.byref(groups),
                                           ctypes.byref(layers), ctypes.byref(grid), prec,
                                           ctypes.byref(out), None, None, None)
        assert st == _capi.LBL_EINVAL
        assert "scratch" in _capi.last_error() and "lbl_optical_depth" in _capi.last_error()
    st = lib.lbl_optical_depth_noalloc(ctypes.byref(lines), ctypes.byref(groups),
                                       ctypes.byref(layers), ctypes.byref(grid), 2,
                                       ctypes.byref(out), None, None, None)
    assert st == _capi.LBL_EINVAL and "precision" in _capi.last_error()
    # the reference's own CPU case fits the array
    n = ctypes.c_size_t(0)
    small_lines = _capi.LblLines(5000, fake, fake, fake, fake, fake, fake, fake, fake)
    small_layers = _capi.LblLayers(20, 1, fake, fake, fake, fake)
    small_grid = _capi.LblGrid(700.0, 0.001, 50_001, 2000, 25.0, 0.01)
    for prec in (0, 1):
        assert lib.lbl_workspace_bytes(ctypes.byref(small_lines), ctypes.byref(small_layers),
                                       ctypes.byref(small_grid), prec, ctypes.byref(n)) == 0
        assert n.value <= 64 << 20


def test_grid_from_nu_reproduces_the_seams_grid():
    """lbl_grid_from_nu (host code) on the arrays the seam receives: the
    golden fixtures' grids (written by the reference) and the workload grids.
    Its points are the caller's bit for bit, its size is nu.size and its tile
    size follows forward.py:175-176 (the tile_samples table of the oracle's
    tests: 1999 at 2400 cm-1, 3999 at step 5e-4 from 1200 cm-1)."""
    import numpy as np
    from pylblrtm_amd.device import make_grid
    arrays = []
    for name in ("co2_mini", "three_gas_edge", "tile1"):
        z = np.load(ROOT / "tests" / "golden" / f"{name}.npz")
        arrays.append((z["out_nu"], json.loads(str(z["meta_json"]))["tile"]))
    for lo, hi, st in [(700.0, 750.0, 0.001), (500.0, 2500.0, 0.001), (500.0, 2500.0, 5e-4),
                       (1000.0, 1100.0, 1e-4), (2400.0, 2401.0, 0.001), (1200.0, 1201.0, 5e-4),
                       (700.0, 700.0, 0.001), (0.1, 1.0, 0.003)]:
        arrays.append((make_grid(lo, hi, st), 2.0))
    for nu, tile in arrays:
        g = _capi.grid_from_nu(nu, tile, 25.0)
        assert g.n_nu == nu.size and g.nu_min == nu[0] and g.cutoff == 25.0
        rebuilt = g.nu_min + np.arange(g.n_nu, dtype=np.float64) * g.step
        np.testing.assert_array_equal(rebuilt, nu)
        ref_step = float(nu[1] - nu[0]) if nu.size > 1 else 0.001
        assert g.tile_samples == max(1, int(np.floor(tile / ref_step + 1e-8)))
    assert _capi.grid_from_nu(make_grid(2400.0, 2401.0, 0.001), 2.0, 25.0).tile_samples == 1999
    assert _capi.grid_from_nu(make_grid(1200.0, 1201.0, 5e-4), 2.0, 25.0).tile_samples == 3999
    # forward.py:141-142's own count: the seam's nu.size, not a re-derived one
    nu = make_grid(700.0, 750.0, 0.001)
    assert nu.size == 50_001 and _capi.grid_from_nu(nu, 2.0, 25.0).n_nu == 50_001


def test_grid_from_nu_rejects_what_the_kernels_cannot_grid():
    import numpy as np
    for bad in (np.array([1.0, 2.0, 4.0]), np.array([3.0, 2.0, 1.0]), np.array([1.0, np.nan]),
                np.zeros(0), 700.0 + np.sort(np.random.default_rng(0).uniform(0, 1, 50))):
        with pytest.raises(ValueError):
            _capi.grid_from_nu(bad, 2.0, 25.0)
    with pytest.raises(ValueError):
        _capi.grid_from_nu(np.array([700.0, 700.001]), 0.0, 25.0)

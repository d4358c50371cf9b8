"""The C-ABI library loads and exports every function include/*.h declares;
argument validation happens before any device call (CPU only)."""

import ctypes
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

"""HITRAN file formats on the host: the HITRAN-2004 ``.par`` line list and
two-column TIPS tables, mirroring pylblrtm's hitran_manager
(``par_to_dict`` :816-872, ``txt_to_tips`` :917-963, the column layout of
:663-686, ``HITRANDataError`` :134, ``get_molar_mass`` :512-552).

``par_to_dict`` runs the line loop in C++ (``lbl_par_read`` in
libpylblrtm_amd, host code: no GPU needed) with the reference's exact
semantics: Python's universal newlines, lines under 155 characters
skipped, fields converted with Python's ``int``/``float`` rules in file
order and appended one by one (a malformed line keeps the fields before the
bad one, as in the reference), one warning per malformed line.  A file
holding non-ASCII bytes is decoded here instead, exactly as the reference
opens it (UTF-8, errors replaced), since Python's number parsing then also
accepts non-ASCII digits and spaces.

Downloading from HITRANonline (HAPI, credentials, the cache) is not part of
this package; nor is HAPI's isotopologue table, so ``get_molar_mass``
raises and molar masses must be passed explicitly.
"""

from __future__ import annotations

import ctypes
import logging
from pathlib import Path
from typing import Dict, Tuple

import numpy as np

from . import _capi

logger = logging.getLogger(__name__)

PAR_MIN_LEN = 155
# (name, first column, end column) in the order the reference converts them
PAR_FIELDS = (("mol_id", 0, 2), ("iso_id", 2, 3), ("nu0", 3, 15), ("Sref", 15, 25),
              ("g_air", 35, 40), ("g_self", 40, 45), ("Elow", 45, 55), ("n_air", 55, 59),
              ("shift", 59, 67))
_INT_FIELDS = 2


class HITRANDataError(ValueError):
    """HITRAN returned empty or unexpected data (hitran_manager.py:134)."""


def get_molar_mass(molecule: str, iso_id: int = 1) -> float:
    """The reference reads HAPI's isotopologue table (hitran_manager.py:512-552),
    which this package does not ship; pass ``molar_masses`` to ``from_files``."""
    raise HITRANDataError(
        f"No isotopologue mass table for {molecule} iso_id={iso_id}: HAPI's catalogue is not "
        "part of pylblrtm_amd. Pass molar_masses={(mol_id, iso_id): g/mol} explicitly.")


def _convert(raw: str, lists) -> None:
    """The reference's per-line conversion: int/float of each field in order,
    appended as soon as it converts."""
    for k, (_name, a, b) in enumerate(PAR_FIELDS):
        lists[k].append(int(raw[a:b]) if k < _INT_FIELDS else float(raw[a:b]))


def _finish(lists, par_path: Path) -> Dict[str, np.ndarray]:
    if not lists[2]:
        raise HITRANDataError(f"No valid lines read from {par_path}.")
    out = {}
    for k, (name, _a, _b) in enumerate(PAR_FIELDS):
        if name in ("mol_id", "iso_id"):
            continue
        out[name] = np.asarray(lists[k], dtype=np.float64)
    out["mol_id"] = np.asarray(lists[0], dtype=np.int32)
    out["iso_id"] = np.asarray(lists[1], dtype=np.int32)
    return out


def _par_to_dict_text(par_path: Path) -> Dict[str, np.ndarray]:
    lists = [[] for _ in PAR_FIELDS]
    with par_path.open(encoding="utf-8", errors="replace") as f:
        for line_num, raw in enumerate(f, 1):
            if len(raw.rstrip("\n")) < PAR_MIN_LEN:
                continue
            try:
                _convert(raw, lists)
            except ValueError as exc:
                logger.warning("Malformed line %d in %s: %s", line_num, par_path.name, exc)
    return _finish(lists, par_path)


def par_to_dict(par_path) -> Dict[str, np.ndarray]:
    """Read a HITRAN-2004 .par file into nu0, Sref, g_air, g_self, Elow,
    n_air, shift (float64) and mol_id, iso_id (int32) arrays."""
    par_path = Path(par_path)
    if not par_path.exists():
        raise FileNotFoundError(f".par file not found: {par_path}")
    lib = _capi.load()
    path = str(par_path).encode()
    n = ctypes.c_int64(0)
    _capi.raise_for(lib.lbl_par_read(path, 0, ctypes.byref(n), None, None, None), "lbl_par_read")
    n_lines = int(n.value)
    status = np.empty(max(n_lines, 1), dtype=np.int8)
    ids = np.empty((max(n_lines, 1), 2), dtype=np.int32)
    vals = np.empty((max(n_lines, 1), 7), dtype=np.float64)
    if n_lines:
        _capi.raise_for(lib.lbl_par_read(path, n_lines, ctypes.byref(n),
                                         status.ctypes.data_as(ctypes.c_void_p),
                                         ids.ctypes.data_as(ctypes.c_void_p),
                                         vals.ctypes.data_as(ctypes.c_void_p)), "lbl_par_read")
    status = status[:n_lines]
    if (status == -2).any():
        return _par_to_dict_text(par_path)
    bad = np.nonzero((status >= 0) & (status < len(PAR_FIELDS)))[0]
    if bad.size:
        _warn_malformed(par_path, set(int(i) for i in bad))
    out = {}
    for k, (name, _a, _b) in enumerate(PAR_FIELDS):
        rows = status > k
        if k < _INT_FIELDS:
            out[name] = ids[:n_lines][rows, k].astype(np.int32)
        else:
            out[name] = vals[:n_lines][rows, k - _INT_FIELDS].astype(np.float64)
    if out["nu0"].size == 0:
        raise HITRANDataError(f"No valid lines read from {par_path}.")
    return {k: out[k] for k in ("nu0", "Sref", "g_air", "g_self", "Elow", "n_air", "shift",
                                "mol_id", "iso_id")}


def _warn_malformed(par_path: Path, bad_lines) -> None:
    """The reference's warning text for each malformed line (0-based indices),
    with the message of the conversion that failed."""
    with par_path.open(encoding="utf-8", errors="replace") as f:
        for idx, raw in enumerate(f):
            if idx not in bad_lines:
                continue
            try:
                _convert(raw, [[] for _ in PAR_FIELDS])
            except ValueError as exc:
                logger.warning("Malformed line %d in %s: %s", idx + 1, par_path.name, exc)


def txt_to_tips(txt_path) -> Tuple[np.ndarray, np.ndarray]:
    """Two-column TIPS table (T [K], Q(T)); blank lines, '#' comments and rows
    with fewer than two fields are skipped (hitran_manager.py:917-963).

    As in the reference loop (hitran_manager.py:950-958), T is appended
    before Q is converted: a row whose first field converts and whose second
    does not keeps its T, so the two arrays come back with different lengths
    (and np.interp, hence ``SpectralData.Q``, rejects the table)."""
    txt_path = Path(txt_path)
    if not txt_path.exists():
        raise FileNotFoundError(f"TIPS file not found: {txt_path}")
    t_list, q_list = [], []
    with txt_path.open(encoding="utf-8", errors="replace") as f:
        for raw in f:
            raw = raw.strip()
            if not raw or raw.startswith("#"):
                continue
            parts = raw.split()
            if len(parts) < 2:
                continue
            try:
                t_list.append(float(parts[0]))
                q_list.append(float(parts[1]))
            except ValueError:
                continue
    if not t_list:
        raise HITRANDataError(f"No data read from {txt_path}.")
    return np.asarray(t_list, dtype=np.float64), np.asarray(q_list, dtype=np.float64)


# ------------------------------------------------------------------ writing --
def _f5_4(v: float) -> str:
    """Fortran F5.4 as HITRAN writes g_air: '.xxxx' below 1, else cut to 5."""
    s = f"{v:.4f}"
    return s[1:] if s.startswith("0.") else s[:5]


def _f4_2(v: float) -> str:
    """F4.2 (n_air): '-.xx' for -1 < v < 0, else cut to 4."""
    s = f"{v:.2f}"
    return "-" + s[2:] if s.startswith("-0.") else s[:4]


def _f8_6(v: float) -> str:
    """F8.6 (pressure shift): '-.xxxxxx' for -1 < v < 0, else cut to 8."""
    s = f"{v:.6f}"
    return "-" + s[2:] if s.startswith("-0.") else s[:8]


def dict_to_par(lines: Dict[str, np.ndarray], par_path) -> Path:
    """Write the fields a line dict holds as HITRAN-2004 .par records (160
    characters: the column layout of hitran_manager.py:663-686 with the
    number formats hapi_to_par uses, :755-808; Einstein A, quanta,
    uncertainty/reference codes and statistical weights, which the forward
    model does not use, are written blank or zero)."""
    par_path = Path(par_path)
    n = len(lines["nu0"])
    blank = " " * 15
    rows = []
    for i in range(n):
        rows.append(
            f"{int(lines['mol_id'][i]):2d}{int(lines['iso_id'][i]):1d}"
            f"{float(lines['nu0'][i]):12.6f}"
            f"{float(lines['Sref'][i]):10.3E}"
            f"{0.0:10.3E}"
            f"{_f5_4(float(lines['g_air'][i]))}"
            f"{float(lines['g_self'][i]):5.3f}"
            f"{float(lines['Elow'][i]):10.4f}"
            f"{_f4_2(float(lines['n_air'][i]))}"
            f"{_f8_6(float(lines['shift'][i]))}"
            f"{blank}{blank}{blank}{blank}"
            f"{' ' * 6}{' ' * 12}{' '}"
            f"{0.0:7.1f}{0.0:7.1f}")
    par_path.write_text("\n".join(rows) + "\n", encoding="utf-8")
    return par_path

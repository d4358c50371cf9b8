"""Oracle (test infrastructure only, never imported by the package): a plain
Python restatement of the reference's HITRAN file readers.

* ``par_to_dict``  -- pylblrtm/hitran_manager.py:816-872: universal-newline
  text lines, lines shorter than 155 characters skipped, fields
  mol_id [0:2], iso_id [2:3], nu0 [3:15], Sref [15:25], g_air [35:40],
  g_self [40:45], Elow [45:55], n_air [55:59], shift [59:67] converted with
  int()/float() and appended one at a time (a line that fails part-way keeps
  its earlier fields), one warning per malformed line, HITRANDataError-like
  ValueError when no nu0 was read.
* ``txt_to_tips``  -- :917-963.

Pinned by hand-built records with known values, the edge cases of the
conversion rules and the column layout of :663-686 (tests/test_par_reader.py).
"""

from __future__ import annotations

import logging
from pathlib import Path

import numpy as np

log = logging.getLogger("oracle.par")

FIELDS = ((0, 2), (2, 3), (3, 15), (15, 25), (35, 40), (40, 45), (45, 55), (55, 59), (59, 67))
NAMES = ("mol_id", "iso_id", "nu0", "Sref", "g_air", "g_self", "Elow", "n_air", "shift")


class NoLines(ValueError):
    pass


def par_to_dict(par_path):
    """Returns (dict of arrays, list of (line number, message) for malformed lines)."""
    par_path = Path(par_path)
    if not par_path.exists():
        raise FileNotFoundError(f".par file not found: {par_path}")
    cols = [[] for _ in NAMES]
    bad = []
    with par_path.open(encoding="utf-8", errors="replace") as f:
        for num, raw in enumerate(f, 1):
            if len(raw.rstrip("\n")) < 155:
                continue
            try:
                for k, (a, b) in enumerate(FIELDS):
                    cols[k].append(int(raw[a:b]) if k < 2 else float(raw[a:b]))
            except ValueError as exc:
                bad.append((num, str(exc)))
    if not cols[2]:
        raise NoLines(f"No valid lines read from {par_path}.")
    out = {name: np.asarray(c, dtype=np.int32 if k < 2 else np.float64)
           for k, (name, c) in enumerate(zip(NAMES, cols))}
    return out, bad


def txt_to_tips(txt_path):
    txt_path = Path(txt_path)
    if not txt_path.exists():
        raise FileNotFoundError(f"TIPS file not found: {txt_path}")
    ts, qs = [], []
    with txt_path.open(encoding="utf-8", errors="replace") as f:
        for raw in f:
            raw = raw.strip()
            if not raw or raw.startswith("#"):
                continue
            parts = raw.split()
            if len(parts) < 2:
                continue
            try:   # T first, then Q (hitran_manager.py:955-956): a bad Q leaves T appended
                ts.append(float(parts[0]))
                qs.append(float(parts[1]))
            except ValueError:
                continue
    if not ts:
        raise NoLines(f"No data read from {txt_path}.")
    return np.asarray(ts, dtype=np.float64), np.asarray(qs, dtype=np.float64)

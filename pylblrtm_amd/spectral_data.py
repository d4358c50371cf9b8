"""Line-list container handed to the forward model.

Mirrors ``SpectralData`` of the reference (pylblrtm/spectral_data.py:14-78):
the same field names, the same structure-of-arrays ``lines`` dict, the same
``Q(T, mol_id, iso_id)`` partition-sum lookup (linear interpolation clamped at
the table ends, i.e. ``numpy.interp``) and the same ``KeyError`` for a missing
table.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional, Tuple

import numpy as np

LINE_FIELDS = ("nu0", "Sref", "g_air", "g_self", "Elow", "n_air", "shift")
ID_FIELDS = ("mol_id", "iso_id")


@dataclass
class SpectralData:
    """Line parameters, partition-sum tables and molar masses.

    ``lines`` holds float64 arrays under the keys of ``LINE_FIELDS`` and
    integer ``mol_id`` / ``iso_id`` arrays (HITRAN numbering), one entry per
    line.  ``tips[(mol_id, iso_id)]`` is a ``(T, Q)`` pair of 1-D arrays and
    ``molar_masses[(mol_id, iso_id)]`` is in g/mol.
    """

    lines: Dict[str, np.ndarray]
    tips: Dict[tuple, Tuple[np.ndarray, np.ndarray]]
    molar_masses: Dict[tuple, float]
    molecule: Optional[str] = None

    def Q(self, T: float, mol_id: int, iso_id: int) -> float:
        """Partition sum at temperature ``T`` (reference spectral_data.py:43-52).

        Linear interpolation in the tabulated (T, Q) pairs, clamped to the end
        values outside the table, exactly as ``numpy.interp``.
        """
        key = (int(mol_id), int(iso_id))
        table = self.tips.get(key)
        if table is None:
            raise KeyError(
                f"No TIPS data for (mol_id={mol_id}, iso_id={iso_id}). "
                f"Available: {sorted(self.tips.keys())}"
            )
        t_nodes, q_nodes = table
        return float(np.interp(T, t_nodes, q_nodes))

    @property
    def nu0(self) -> np.ndarray:
        return self.lines["nu0"]

    @property
    def Sref(self) -> np.ndarray:
        return self.lines["Sref"]

    @property
    def n_lines(self) -> int:
        return len(self.lines["nu0"])

    def __repr__(self) -> str:
        keys = ", ".join(f"({m},{i})" for m, i in sorted(self.tips.keys()))
        nu = self.lines["nu0"]
        return (
            f"SpectralData({self.molecule or 'multi-molecule'}, "
            f"(mol_id,iso_id)=[{keys}], {self.n_lines} lines, "
            f"nu=[{nu.min():.1f}, {nu.max():.1f}] cm-1)"
        )

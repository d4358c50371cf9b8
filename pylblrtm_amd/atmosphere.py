"""Atmospheric levels and the layer averages the forward model consumes.

Mirrors ``AtmosphericProfile`` / ``_Layers`` / ``build_layers`` of the
reference (pylblrtm/atmosphere.py:21-141): levels ordered top of atmosphere
first (pressure strictly increasing), the same validation and messages, and
the same layer averages:

* pressure: geometric mean of the bounding levels, converted to atm;
* temperature and mixing ratio: arithmetic mean of the bounding levels;
* column: hydrostatic, ``vmr * N_A / (g * M_air) * dp[Pa] / 1e4`` molec/cm^2.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

# Constants of atmosphere.py:15-18 (used verbatim so columns agree bit for bit).
AVOGADRO = 6.02214086e23     # 1/mol
GRAVITY = 9.80665            # m/s^2
M_DRY_AIR = 0.0289644        # kg/mol
P_STD_HPA = 1013.25          # hPa per atm


@dataclass
class _Layers:
    """Per-layer arrays, index 0 = topmost layer (atmosphere.py:21-35)."""

    pressure_atm: np.ndarray
    temperature_k: np.ndarray
    vmr: Dict[str, np.ndarray]
    column_cm2: Dict[str, np.ndarray]
    # hydrostatic air column per layer [molec cm^-2] (column_cm2 = vmr * air_cm2);
    # not a field of the reference's _Layers: the gradient at vmr = 0 needs it,
    # since a zero mixing ratio leaves a zero column that does not carry it
    air_cm2: Optional[np.ndarray] = None

    @property
    def n_layers(self) -> int:
        return len(self.pressure_atm)


def _as_levels(values, name: str) -> np.ndarray:
    return np.asarray(values, dtype=np.float64)


class AtmosphericProfile:
    """Pressure [hPa], temperature [K] and mixing ratios at each level.

    Validation follows atmosphere.py:69-100 (same exception types and text).
    """

    def __init__(self, pressure_hpa, temperature_k, vmr: Dict[str, object]) -> None:
        p = _as_levels(pressure_hpa, "pressure_hpa")
        t = _as_levels(temperature_k, "temperature_k")

        if p.ndim != 1 or p.size < 2:
            raise ValueError("pressure_hpa must be a 1-D array with at least 2 levels")
        if not np.isfinite(p).all() or (p <= 0).any():
            raise ValueError("pressure_hpa must be finite and positive")
        if (np.diff(p) <= 0).any():
            raise ValueError(
                "pressure_hpa must be strictly increasing from TOA "
                "(smallest pressure = highest altitude = index 0)"
            )
        if t.shape != p.shape:
            raise ValueError("temperature_k must have the same length as pressure_hpa")
        if not np.isfinite(t).all() or (t <= 0).any():
            raise ValueError("temperature_k must be finite and positive")

        checked: Dict[str, np.ndarray] = {}
        for gas, values in vmr.items():
            v = _as_levels(values, gas)
            if v.shape != p.shape:
                raise ValueError(
                    f"vmr['{gas}'] must have the same length as pressure_hpa "
                    f"({p.size}), got {v.size}"
                )
            if not np.isfinite(v).all() or (v < 0).any() or (v > 1).any():
                raise ValueError(f"vmr['{gas}'] must be finite and in [0, 1]")
            checked[gas] = v

        self.pressure_hpa = p
        self.temperature_k = t
        self.vmr = checked

    @property
    def n_levels(self) -> int:
        return self.pressure_hpa.size

    @property
    def n_layers(self) -> int:
        return self.n_levels - 1

    def build_layers(self) -> _Layers:
        """Layer averages (atmosphere.py:110-141)."""
        p = self.pressure_hpa
        t = self.temperature_k
        p_layer = np.sqrt(p[:-1] * p[1:]) / P_STD_HPA
        t_layer = 0.5 * (t[:-1] + t[1:])
        vmr_layer = {g: 0.5 * (v[:-1] + v[1:]) for g, v in self.vmr.items()}
        # molecules per m^2 per Pa, then per cm^2 (1e4 cm^2 per m^2)
        per_pa = AVOGADRO / (GRAVITY * M_DRY_AIR)
        air_cm2 = per_pa * (np.abs(np.diff(p)) * 100.0) / 1e4
        col = {g: v * air_cm2 for g, v in vmr_layer.items()}
        return _Layers(p_layer, t_layer, vmr_layer, col, air_cm2)

    def __repr__(self) -> str:
        return (
            f"AtmosphericProfile({self.n_levels} levels, "
            f"P=[{self.pressure_hpa[0]:.4g}–{self.pressure_hpa[-1]:.4g}] hPa, "
            f"T=[{self.temperature_k.min():.1f}–{self.temperature_k.max():.1f}] K, "
            f"gases={list(self.vmr.keys())})"
        )

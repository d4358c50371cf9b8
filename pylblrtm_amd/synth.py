"""Seeded synthetic inputs for the benchmark and parity workloads.

Every generator uses numpy's PCG64 with a fixed seed, so the oracle and the
HIP path always consume identical arrays.  Workloads:

* ``co2_window``   — 1 CO2-like isotopologue (43.99 g/mol), 5,000 lines in
  700-750 cm^-1, 21 levels (20 layers), grid 700-750 cm^-1 at 0.001.
* ``five_gas``     — H2O, CO2, O3, N2O, CH4 (HITRAN ids 1, 2, 3, 4, 6;
  18.01, 43.99, 47.98, 44.01, 16.04 g/mol), 100,000 lines each in
  500-2500 cm^-1, 51 levels (50 layers), grid 500-2500 cm^-1 at 0.001.
* ``perturbed``    — 64 profile perturbations of the five-gas atmosphere
  (layer T + N(0, 2 K), vmr x lognormal(0, 0.1)).

Choices the workload descriptions leave open (fixed here, once):

* altitude of a pressure level: hypsometric relation for T(z) = 288 - 6.5 z
  (z in km) down to 216.65 K, isothermal above;
* mixing-ratio profiles of the five-gas atmosphere: H2O 7.745e-3 exp(-z/2),
  CO2 4.15e-4, O3 7e-6 exp(-((z-25)/5)^2 / 2), N2O 3.2e-7 exp(-z/10),
  CH4 1.8e-6 exp(-z/15) (surface values of the AFGL mid-latitude summer
  profile, O3 a Gaussian of sigma 5 km peaking at 25 km);
* partition sums: Q(T) = 286.1 (T/296)^1.5 tabulated at 70..400 K every 1 K
  for every isotopologue;
* the perturbation window: 1000-1100 cm^-1 at 1e-4 (1,000,001 points).
"""

from __future__ import annotations

import math
from typing import Dict, List, Optional, Tuple

import numpy as np

from .atmosphere import AtmosphericProfile
from .spectral_data import SpectralData

R_GAS = 8.31446
G0 = 9.80665
M_AIR = 0.0289644
T_SURF = 288.0
LAPSE = 6.5          # K/km
T_TROP = 216.65

SEEDS = {"co2_window": 20260101, "five_gas": 20260102, "perturbed": 20260104}

FIVE_GASES = (  # name, HITRAN id, molar mass [g/mol]
    ("H2O", 1, 18.01), ("CO2", 2, 43.99), ("O3", 3, 47.98), ("N2O", 4, 44.01), ("CH4", 6, 16.04),
)


def tips_table(q296: float = 286.1) -> Tuple[np.ndarray, np.ndarray]:
    t = np.arange(70.0, 401.0, 1.0)
    return t, q296 * (t / 296.0) ** 1.5


def altitude_km(p_hpa: np.ndarray, p_surf: float = 1013.25) -> np.ndarray:
    """Hypsometric altitude for the 6.5 K/km troposphere + isothermal layer."""
    expo = R_GAS * (LAPSE / 1000.0) / (G0 * M_AIR)
    z_trop = (T_SURF - T_TROP) / LAPSE
    p_trop = p_surf * (T_TROP / T_SURF) ** (1.0 / expo)
    p = np.asarray(p_hpa, np.float64)
    z_low = (T_SURF / LAPSE) * (1.0 - (p / p_surf) ** expo)
    z_high = z_trop + (R_GAS * T_TROP / (G0 * M_AIR)) * np.log(p_trop / p) / 1000.0
    return np.where(p >= p_trop, z_low, z_high)


def levels(n_levels: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Pressure (TOA first, increasing), temperature and altitude per level."""
    p = np.logspace(np.log10(1013.25), np.log10(1.0), n_levels)[::-1].copy()
    z = altitude_km(p)
    t = np.maximum(T_SURF - LAPSE * z, T_TROP)
    return p, t, z


def random_lines(rng: np.random.Generator, n: int, nu_lo: float, nu_hi: float,
                 mol_id: int, iso_id: int = 1) -> Dict[str, np.ndarray]:
    return {
        "nu0": rng.uniform(nu_lo, nu_hi, n),
        "Sref": 10.0 ** rng.uniform(-27.0, -19.0, n),
        "g_air": rng.uniform(0.05, 0.10, n),
        "g_self": rng.uniform(0.07, 0.12, n),
        "n_air": rng.uniform(0.6, 0.8, n),
        "shift": rng.normal(-0.002, 0.001, n),
        "Elow": rng.uniform(0.0, 3000.0, n),
        "mol_id": np.full(n, mol_id, np.int32),
        "iso_id": np.full(n, iso_id, np.int32),
    }


def _concat(parts: List[Dict[str, np.ndarray]]) -> Dict[str, np.ndarray]:
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def co2_window(n_lines: int = 5000, n_levels: int = 21, nu_lo: float = 700.0,
               nu_hi: float = 750.0, seed: Optional[int] = None):
    """Single CO2-like isotopologue workload -> (SpectralData, AtmosphericProfile)."""
    rng = np.random.default_rng(SEEDS["co2_window"] if seed is None else seed)
    lines = random_lines(rng, n_lines, nu_lo, nu_hi, 2, 1)
    sd = SpectralData(lines=lines, tips={(2, 1): tips_table()},
                      molar_masses={(2, 1): 43.99}, molecule="CO2")
    p, t, _ = levels(n_levels)
    prof = AtmosphericProfile(p, t, {"CO2": np.full(n_levels, 4.15e-4)})
    return sd, prof


def five_gas_vmr(z_km: np.ndarray) -> Dict[str, np.ndarray]:
    return {
        "H2O": 7.745e-3 * np.exp(-z_km / 2.0),
        "CO2": np.full_like(z_km, 4.15e-4),
        "O3": 7e-6 * np.exp(-0.5 * ((z_km - 25.0) / 5.0) ** 2),
        "N2O": 3.2e-7 * np.exp(-z_km / 10.0),
        "CH4": 1.8e-6 * np.exp(-z_km / 15.0),
    }


def five_gas(lines_per_gas: int = 100_000, n_levels: int = 51, nu_lo: float = 500.0,
             nu_hi: float = 2500.0, seed: Optional[int] = None):
    """Five-gas workload -> (SpectralData, AtmosphericProfile)."""
    rng = np.random.default_rng(SEEDS["five_gas"] if seed is None else seed)
    parts, tips, mm = [], {}, {}
    for _name, mol, mass in FIVE_GASES:
        parts.append(random_lines(rng, lines_per_gas, nu_lo, nu_hi, mol, 1))
        tips[(mol, 1)] = tips_table()
        mm[(mol, 1)] = mass
    sd = SpectralData(lines=_concat(parts), tips=tips, molar_masses=mm, molecule=None)
    p, t, z = levels(n_levels)
    prof = AtmosphericProfile(p, t, five_gas_vmr(z))
    return sd, prof


def perturbations(layers, n_profiles: int = 64, seed: Optional[int] = None):
    """Seeded (T_k, vmr_{k,g}) perturbations of a layer set.

    Returns (T[n_prof, n_layers], vmr{gas: [n_prof, n_layers]},
    col{gas: [n_prof, n_layers]}) with col = vmr x the layer's air column, so
    each perturbed layer set stays self-consistent.
    """
    rng = np.random.default_rng(SEEDS["perturbed"] if seed is None else seed)
    nl = layers.n_layers
    T = layers.temperature_k[None, :] + rng.normal(0.0, 2.0, (n_profiles, nl))
    vmr, col = {}, {}
    for gas in layers.vmr:
        v0 = layers.vmr[gas]
        air = np.where(v0 > 0, layers.column_cm2[gas] / np.where(v0 > 0, v0, 1.0), 0.0)
        f = rng.lognormal(0.0, 0.1, (n_profiles, nl))
        vmr[gas] = np.minimum(v0[None, :] * f, 1.0)
        col[gas] = vmr[gas] * air[None, :]
    return T, vmr, col

"""FP64 CPU restatement of PyLBLRTM's line-by-line optical-depth path.

TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it.

It restates, function by function, the reference's algorithm (file:line
citations are into the reference tree) and evaluates the Faddeeva function
with ``scipy.special.wofz``, the routine the reference itself calls
(pylblrtm/forward.py:22,62).  Summation order inside a tile differs from the
reference's 256-line BLAS blocks; every summand is positive, so this changes
results only at the 1e-15 level.

Parity pins (see tests/test_oracle.py):
  * grid / tile_samples: closed-form table incl. the origin-dependent cases;
  * layer builder: hand closed form and the 1013.25 hPa column constant;
  * Q(T): the reference's own test (Q = 0.5 T -> Q(200) = 100) and np.interp;
  * line scaling: T = 296 / vmr -> 0 special case, linearity in Sref and col;
  * tile inclusion: brute force per point for tile_samples = 1, edge cases;
  * Voigt: scipy's published wofz values, mpmath w(z), K(x,0) = exp(-x^2),
    K(0,y) = erfcx(y), unit area;
  * integrated path: golden vectors written by running the reference
    (tests/golden/make_golden.py).
"""

from __future__ import annotations

import math
import warnings
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
from scipy.special import wofz

# forward.py:32-36
T_REF = 296.0
C2 = 1.4387752
DOPPLER = 3.5810e-7
SQRT_LN2 = math.sqrt(math.log(2.0))
VOIGT_K = SQRT_LN2 / math.sqrt(math.pi)

# atmosphere.py:15-18
AVOGADRO = 6.02214086e23
GRAVITY = 9.80665
M_DRY_AIR = 0.0289644
P_STD_HPA = 1013.25

MOL_ID_TO_NAME = {
    1: "H2O", 2: "CO2", 3: "O3", 4: "N2O", 5: "CO", 6: "CH4", 7: "O2", 8: "NO",
    9: "SO2", 10: "NO2", 11: "NH3", 12: "HNO3", 13: "OH", 14: "HF", 15: "HCl",
    16: "HBr", 17: "HI", 18: "ClO", 19: "OCS", 20: "H2CO", 21: "HOCl", 22: "N2",
    23: "HCN", 24: "CH3Cl", 25: "H2O2", 26: "C2H2", 27: "C2H6", 28: "PH3",
    29: "COF2", 30: "SF6", 31: "H2S", 32: "HCOOH", 33: "HO2", 34: "O",
    35: "ClONO2", 36: "NO+", 37: "HOBr", 38: "C2H4", 39: "CH3OH", 40: "CH3Br",
    41: "CH3CN", 42: "CF4", 43: "C4H2", 44: "HC3N", 45: "H2", 46: "CS",
    47: "SO3", 48: "C2N2", 49: "COCl2", 51: "SO", 52: "CH3F", 53: "GeH4",
    54: "CS2", 55: "CH3I", 56: "NF3",
}  # hitran_manager.py:89-101 (HITRAN numbering)


# ── grid and tiles ────────────────────────────────────────────────────────────

def grid(nu_min: float, nu_max: float, step: float) -> np.ndarray:
    """forward.py:141-142: n = ceil((max-min)/step)+1, nu_i = min + i*step."""
    n = int(np.ceil((nu_max - nu_min) / step)) + 1
    return nu_min + np.arange(n, dtype=np.float64) * step


def tile_samples(nu: np.ndarray, tile_width: float) -> int:
    """forward.py:175-176 (step taken from the first two grid points)."""
    step = float(nu[1] - nu[0]) if nu.size > 1 else 0.001
    return max(1, int(np.floor(tile_width / step + 1e-8)))


def tile_bounds(nu: np.ndarray, ts: int) -> List[Tuple[int, int]]:
    return [(t0, min(t0 + ts, nu.size)) for t0 in range(0, nu.size, ts)]


# ── layers and partition sums ─────────────────────────────────────────────────

def build_layers(p_hpa: np.ndarray, t_k: np.ndarray, vmr: Dict[str, np.ndarray]):
    """atmosphere.py:110-141 -> (P_atm, T, vmr_layer, col_layer)."""
    p = np.asarray(p_hpa, np.float64)
    t = np.asarray(t_k, np.float64)
    p_lay = np.sqrt(p[:-1] * p[1:]) / P_STD_HPA
    t_lay = 0.5 * (t[:-1] + t[1:])
    v_lay = {g: 0.5 * (np.asarray(v)[:-1] + np.asarray(v)[1:]) for g, v in vmr.items()}
    air = (AVOGADRO / (GRAVITY * M_DRY_AIR)) * (np.abs(np.diff(p)) * 100.0) / 1e4
    col = {g: v * air for g, v in v_lay.items()}
    return p_lay, t_lay, v_lay, col


def partition_sum(T: float, t_nodes: np.ndarray, q_nodes: np.ndarray) -> float:
    """spectral_data.py:43-52: np.interp, clamped at the table ends."""
    return float(np.interp(T, t_nodes, q_nodes))


# ── prelude: groups, band mask, q ratios ──────────────────────────────────────

def prepare_groups(nu, lines, tips, molar_masses, p_atm, t_lay, gas_names: Iterable[str],
                   cutoff: float, stacklevel: int = 3):
    """forward.py:185-245.  Returns a list of dicts, in sorted (mol, iso) order."""
    nu0 = lines["nu0"]
    delta = lines["shift"]
    mol = lines["mol_id"]
    iso = lines["iso_id"]
    gas_names = set(gas_names)
    p_max = float(np.max(p_atm)) if len(p_atm) > 0 else 1.0
    bound = np.abs(delta) * p_max
    out = []
    for m, i in sorted(set(zip(mol.tolist(), iso.tolist()))):
        name = MOL_ID_TO_NAME.get(int(m))
        if name is None:
            warnings.warn(f"mol_id={m} is not in the HITRAN catalogue; lines skipped.",
                          RuntimeWarning, stacklevel=stacklevel)
            continue
        if name not in gas_names:
            warnings.warn(
                f"Gas '{name}' (mol_id={m}, iso_id={i}) is in the spectral data but not "
                f"in the atmospheric profile vmr; its lines are skipped.",
                RuntimeWarning, stacklevel=stacklevel)
            continue
        mask = ((mol == m) & (iso == i)
                & (nu0 >= nu[0] - cutoff - bound) & (nu0 <= nu[-1] + cutoff + bound))
        if not mask.any():
            continue
        mm = molar_masses.get((int(m), int(i)))
        if mm is None:
            warnings.warn(f"No molar mass found for (mol_id={m}, iso_id={i}); "
                          f"using dry-air mass as fallback.", RuntimeWarning, stacklevel=stacklevel)
            mm = 28.97
        key = (int(m), int(i))
        if key not in tips:
            raise KeyError(f"No TIPS data for (mol_id={m}, iso_id={i}). "
                           f"Available: {sorted(tips.keys())}")
        tn, qn = tips[key]
        q_ref = partition_sum(T_REF, tn, qn)
        q_ratio = np.array([q_ref / partition_sum(float(T), tn, qn) for T in t_lay], np.float64)
        out.append(dict(mol=int(m), iso=int(i), gas=name, idx=np.flatnonzero(mask),
                        molar_mass=float(mm), q_ratio=q_ratio))
    return out


# ── per-(layer, line) scaling ─────────────────────────────────────────────────

def radiation_term(nu, T):
    """forward.py:41-48: R(nu, T) = nu * tanh(c2 nu / 2T)."""
    return nu * np.tanh(0.5 * C2 * nu / T)


def line_params(lines, idx, T, P, vmr_gas, col, q_ratio_k, molar_mass):
    """forward.py:257-292 for one (layer, group).  Returns shifted centre,
    strength (column and 1/R at the shifted centre folded in), Lorentz HWHM
    and Doppler HWHM, all float64 arrays over ``idx``."""
    p_self = vmr_gas * P
    p_air = P - p_self
    c = lines["nu0"][idx]
    shifted = c + lines["shift"][idx] * p_air
    strength = (lines["Sref"][idx] * q_ratio_k
                * np.exp(-C2 * lines["Elow"][idx] * (1.0 / T - 1.0 / T_REF))
                * (-np.expm1(-C2 * c / T)) / (-np.expm1(-C2 * c / T_REF))
                * col / radiation_term(shifted, T))
    lorentz = (T_REF / T) ** lines["n_air"][idx] * (lines["g_air"][idx] * p_air
                                                    + lines["g_self"][idx] * p_self)
    doppler = c * DOPPLER * math.sqrt(T / molar_mass)
    return shifted, strength, lorentz, doppler


def voigt(nu, centre, doppler, lorentz):
    """forward.py:51-62: area-normalised Voigt, [n_nu, n_lines]."""
    z = SQRT_LN2 * ((nu[:, None] - centre) + 1j * lorentz) / doppler
    return wofz(z).real * (VOIGT_K / doppler)


# ── the optical-depth stack ───────────────────────────────────────────────────

class OracleError(FloatingPointError):
    pass


def optical_depth(nu, lines, tips, molar_masses, p_atm, t_lay, vmr_lay, col_lay,
                  cutoff: float, tile_width: float, tiles: Optional[Sequence[int]] = None,
                  block: int = 256, stats: Optional[dict] = None):
    """forward.py:157-317.  od[n_layers, n_nu] in float64.

    ``tiles`` restricts the evaluation to a subset of reference tiles (tiles
    are independent, so the subset is exact where computed; elsewhere od is
    NaN).  ``stats['evals']`` accumulates the line-gridpoint evaluation count.
    """
    n_layers = len(p_atm)
    n_nu = nu.size
    od = np.zeros((n_layers, n_nu), np.float64)
    ts = tile_samples(nu, tile_width)
    bounds = tile_bounds(nu, ts)
    chosen = range(len(bounds)) if tiles is None else sorted(set(int(t) for t in tiles))
    if tiles is not None:
        keep = np.zeros(n_nu, bool)
        for t in chosen:
            keep[bounds[t][0]:bounds[t][1]] = True
        od[:, ~keep] = np.nan
    groups = prepare_groups(nu, lines, tips, molar_masses, p_atm, t_lay, col_lay.keys(), cutoff,
                            stacklevel=4)
    evals = 0
    if groups:
        for k in range(n_layers):
            T = float(t_lay[k])
            P = float(p_atm[k])
            for g in groups:
                col = float(col_lay[g["gas"]][k])
                if col == 0.0:
                    continue
                shifted, strength, lorentz, doppler = line_params(
                    lines, g["idx"], T, P, float(vmr_lay[g["gas"]][k]), col,
                    g["q_ratio"][k], g["molar_mass"])
                if not np.isfinite(strength).all() or np.any(strength < 0):
                    raise OracleError(
                        f"Layer {k}, gas '{g['gas']}': non-finite or negative line strength")
                order = np.argsort(shifted, kind="stable")
                sc = shifted[order]
                for t in chosen:
                    t0, t1 = bounds[t]
                    lo = np.searchsorted(sc, nu[t0] - cutoff, side="left")
                    hi = np.searchsorted(sc, nu[t1 - 1] + cutoff, side="right")
                    if hi <= lo:
                        continue
                    sel = order[lo:hi]
                    evals += (hi - lo) * (t1 - t0)
                    tile = nu[t0:t1]
                    for b0 in range(0, sel.size, block):
                        b = sel[b0:b0 + block]
                        od[k, t0:t1] += voigt(tile, shifted[b], doppler[b], lorentz[b]) @ strength[b]
            od[k] *= radiation_term(nu, T)
            if not np.isfinite(od[k][~np.isnan(od[k])] if tiles is not None else od[k]).all():
                raise OracleError(f"Layer {k}: non-finite optical depth after radiation term")
    if stats is not None:
        stats["evals"] = stats.get("evals", 0) + evals
    return od


def transmittance(od: np.ndarray, mu: float) -> np.ndarray:
    """forward.py:151: T = exp(-sum_k od_k / mu)."""
    return np.exp(-od.sum(axis=0) / mu)


def compute_transmittance(sd, profile, nu_min, nu_max, *, step_cm1=0.001, line_cutoff_cm1=25.0,
                          tile_width_cm1=2.0, mu_view=1.0, tiles=None):
    """forward.py:67-152 on plain containers (objects with the reference's
    SpectralData / AtmosphericProfile attributes)."""
    nu = grid(nu_min, nu_max, step_cm1)
    p_atm, t_lay, v_lay, col = build_layers(profile.pressure_hpa, profile.temperature_k, profile.vmr)
    od = optical_depth(nu, sd.lines, sd.tips, sd.molar_masses, p_atm, t_lay, v_lay, col,
                       line_cutoff_cm1, tile_width_cm1, tiles=tiles)
    return nu, transmittance(od, mu_view), od

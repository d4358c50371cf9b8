"""``compute_transmittance`` — drop-in for pylblrtm/forward.py:67-152.

Same signature, defaults, validation and exceptions as the reference, plus
one keyword the reference does not have:

``precision`` — arithmetic of the Voigt line sum on the GPU:
  * ``"fp64"``  (default) FP64 throughout; matches the reference's numbers;
  * ``"mixed"`` FP32 line sum with compensated accumulation, FP64 centre
    offsets and Boltzmann exponent (|dT| <= 1e-5 against FP64);
  * ``"surrogate"`` ``"mixed"`` plus the learned Voigt surrogate.

The work runs in libpylblrtm_amd (HIP, gfx950); there is no CPU fallback.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np

from .atmosphere import AtmosphericProfile
from .device import OpticalDepthProblem
from .spectral_data import SpectralData

__all__ = ["compute_transmittance"]


def _check_spectral(spectral_data) -> None:
    if not isinstance(spectral_data, SpectralData):
        raise TypeError("spectral_data must be a SpectralData instance")


def _check_profile(profile) -> None:
    if not isinstance(profile, AtmosphericProfile):
        raise TypeError("profile must be an AtmosphericProfile instance")


def compute_transmittance(
    spectral_data: SpectralData,
    profile: AtmosphericProfile,
    nu_min_cm1: float,
    nu_max_cm1: float,
    *,
    step_cm1: float = 0.001,
    line_cutoff_cm1: float = 25.0,
    tile_width_cm1: float = 2.0,
    mu_view: float = 1.0,
    precision: str = "fp64",
) -> Tuple[np.ndarray, np.ndarray]:
    """Monochromatic transmittance exp(-sum_k od_k / mu_view) on the grid
    nu_min + i * step, i = 0 .. ceil((nu_max - nu_min)/step).

    Returns ``(wavenumber_cm1, transmittance)``, both float64 arrays.
    """
    _check_spectral(spectral_data)
    _check_profile(profile)
    if nu_min_cm1 >= nu_max_cm1:
        raise ValueError("nu_min_cm1 must be strictly less than nu_max_cm1")
    if step_cm1 <= 0:
        raise ValueError("step_cm1 must be positive")
    if not (0.0 < mu_view <= 1.0):
        raise ValueError("mu_view must be in (0, 1]")
    layers = profile.build_layers()
    prob = OpticalDepthProblem(spectral_data, layers, nu_min_cm1, nu_max_cm1, step_cm1,
                               line_cutoff_cm1, tile_width_cm1, stacklevel=4)
    out = prob.run(precision=precision, mus=(float(mu_view),))
    return prob.nu, out["transmittance"][0].double().cpu().numpy()

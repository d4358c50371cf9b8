"""The reference's own seam, with the reference's signature.

``build_optical_depth(nu, sd, layers, line_cutoff_cm1, tile_width_cm1)`` is
a drop-in body for ``_build_optical_depth`` (pylblrtm/forward.py:157-317): it
takes the wavenumber array the seam receives, the reference's ``SpectralData``
and ``_Layers`` (or this package's, which have the same attributes), and
returns the vertical optical-depth stack ``od[n_layers, nu.size]`` as float64
numpy, computed by the HIP kernels through the C ABI.  A maintainer of the
reference binds it with one line in pylblrtm/forward.py::

    from pylblrtm_amd.seam import build_optical_depth as _build_optical_depth

Behaviour follows the seam: tile size from ``nu[1] - nu[0]``
(forward.py:175-176), the same warnings (same text, reported at the same
caller frame), the same ``KeyError`` / ``FloatingPointError`` messages, zero
rows for layers and groups that the reference skips.  The grid is the
caller's, point for point (``lbl_grid_from_nu``); an array that no evenly
spaced grid reproduces raises ValueError (the reference's callers only build
evenly spaced ones, forward.py:141-142, config.py:246-247).
"""

from __future__ import annotations

import numpy as np

from .device import OpticalDepthProblem


def build_optical_depth(nu, sd, layers, line_cutoff_cm1: float, tile_width_cm1: float,
                        precision: str = "fp64") -> np.ndarray:
    """od[n_layers, nu.size] (float64) of forward.py:157-317 on the GPU.

    ``precision="fp64"`` reproduces the reference's optical depths to 1e-9
    relative; ``"mixed"`` runs the FP32 line sum (max |dT| <= 1e-5)."""
    nu = np.asarray(nu, dtype=np.float64)
    if layers.n_layers == 0:
        return np.zeros((0, nu.size), dtype=np.float64)
    # warnings are reported at the seam's caller's caller, as forward.py's
    # stacklevel=3 inside _build_optical_depth does: prepare_groups (1),
    # OpticalDepthProblem (2), this function (3), its caller (4), theirs (5)
    prob = OpticalDepthProblem(sd, layers, None, None, None, line_cutoff_cm1, tile_width_cm1,
                               stacklevel=5, nu=nu)
    out = prob.run(precision, mus=(1.0,), want_layers=True)
    return out["optical_depth_layers"].double().cpu().numpy()

"""pylblrtm_amd — MI355X-native line-by-line optical depth.

Drop-in for PyLBLRTM's forward path (pylblrtm/__init__.py:18-24 names):
``SpectralData``, ``AtmosphericProfile``, ``compute_transmittance``,
``run_forward``, ``load_config``, ``ForwardResult``, ``MOLECULE_IDS`` and,
for HITRAN files on disk, ``from_files`` / ``HITRANDataError`` /
``get_molar_mass``.  The optical-depth work runs in hand-written HIP kernels
for gfx950 (libpylblrtm_amd.so); the .par reader is host C++ in the same
library; downloading from HITRANonline is not part of this package.
"""

from .spectral_data import SpectralData
from .atmosphere import AtmosphericProfile
from .forward import compute_transmittance
from .config import run_forward, load_config, ForwardResult
from .molecules import MOLECULE_IDS
from .data_loader import from_files
from .hitran_manager import HITRANDataError, get_molar_mass

__version__ = "0.1.0"

__all__ = [
    "SpectralData",
    "AtmosphericProfile",
    "compute_transmittance",
    "run_forward",
    "load_config",
    "ForwardResult",
    "MOLECULE_IDS",
    "from_files",
    "get_molar_mass",
    "HITRANDataError",
]

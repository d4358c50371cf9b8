"""YAML-driven runner — drop-in for pylblrtm/config.py:1-319.

``load_config`` / ``run_forward`` / ``ForwardResult`` keep the reference's
keys, defaults, validation messages, output assembly and file writers.  One
key is added: ``calculation.precision`` (``"fp64"`` default, ``"mixed"``,
``"surrogate"``), passed to the GPU line sum.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Union

import numpy as np

from .atmosphere import AtmosphericProfile
from .device import OpticalDepthProblem
from .spectral_data import SpectralData

_DEFAULTS: dict = {
    "spectral": {"nu_min_cm1": 700.0, "nu_max_cm1": 750.0, "step_cm1": 0.001},
    "calculation": {"mu_view": 1.0, "line_cutoff_cm1": 25.0, "tile_width_cm1": 2.0,
                    "precision": "fp64"},
    "output": {
        "transmittance": True,
        "optical_depth": False,
        "optical_depth_layers": False,
        "save": {"enabled": False, "path": "results/", "format": "npz", "prefix": "forward"},
    },
}

_PRECISIONS = ("fp64", "mixed", "surrogate")


def _merge(base: dict, override: dict) -> dict:
    """Recursive dict merge; neither argument is modified (config.py:50-58)."""
    merged = dict(base)
    for key, val in override.items():
        if isinstance(merged.get(key), dict) and isinstance(val, dict):
            merged[key] = _merge(merged[key], val)
        else:
            merged[key] = val
    return merged


def load_config(source: Union[str, Path, dict]) -> dict:
    """Resolve a YAML path or dict against the defaults and validate it
    (config.py:61-96)."""
    if isinstance(source, dict):
        raw = source
    else:
        try:
            import yaml
        except ImportError:
            raise ImportError("PyYAML is required to load YAML config files. "
                              "Install it with:  pip install pyyaml") from None
        path = Path(source)
        if not path.exists():
            raise FileNotFoundError(f"Config file not found: {path}")
        with open(path, "r", encoding="utf-8") as fh:
            raw = yaml.safe_load(fh) or {}
    cfg = _merge(_DEFAULTS, raw)
    _validate(cfg)
    return cfg


def _validate(cfg: dict) -> None:
    """config.py:99-124, plus the precision key."""
    sp, ca, out = cfg["spectral"], cfg["calculation"], cfg["output"]
    if sp["nu_min_cm1"] >= sp["nu_max_cm1"]:
        raise ValueError("spectral.nu_min_cm1 must be strictly less than nu_max_cm1")
    if sp["step_cm1"] <= 0:
        raise ValueError("spectral.step_cm1 must be positive")
    if ca["line_cutoff_cm1"] <= 0:
        raise ValueError("calculation.line_cutoff_cm1 must be positive")
    if not (0.0 < ca["mu_view"] <= 1.0):
        raise ValueError("calculation.mu_view must be in (0, 1]")
    if not any([out["transmittance"], out["optical_depth"], out["optical_depth_layers"]]):
        raise ValueError("At least one output must be enabled: "
                         "transmittance, optical_depth, or optical_depth_layers")
    fmt = out["save"]["format"]
    if fmt not in {"npz", "csv", "txt"}:
        raise ValueError(f"output.save.format must be 'npz', 'csv', or 'txt'; got '{fmt}'")
    if ca.get("precision", "fp64") not in _PRECISIONS:
        raise ValueError(f"calculation.precision must be one of {_PRECISIONS}; "
                         f"got '{ca.get('precision')}'")


@dataclass
class ForwardResult:
    """Requested outputs of :func:`run_forward` (config.py:129-191)."""

    wavenumber_cm1: np.ndarray
    config: dict
    transmittance: Optional[np.ndarray] = field(default=None)
    optical_depth: Optional[np.ndarray] = field(default=None)
    optical_depth_layers: Optional[np.ndarray] = field(default=None)

    def save(self, path: Optional[str] = None, fmt: Optional[str] = None,
             prefix: Optional[str] = None) -> Path:
        sv = self.config["output"]["save"]
        return _write(self, path=path or sv["path"], fmt=fmt or sv["format"],
                      prefix=prefix or sv["prefix"])

    def __repr__(self) -> str:
        nu = self.wavenumber_cm1
        names = []
        if self.transmittance is not None:
            names.append("transmittance")
        if self.optical_depth is not None:
            names.append("optical_depth")
        if self.optical_depth_layers is not None:
            names.append(f"optical_depth_layers{list(self.optical_depth_layers.shape)}")
        return (f"ForwardResult(nu=[{nu[0]:.2f}, {nu[-1]:.2f}] cm-1 | "
                f"step={nu[1]-nu[0]:.4g} cm-1 | outputs=[{', '.join(names)}])")


def run_forward(spectral_data: SpectralData, profile: AtmosphericProfile,
                config: Union[str, Path, dict]) -> ForwardResult:
    """config.py:196-279 on the GPU path."""
    if not isinstance(spectral_data, SpectralData):
        raise TypeError("spectral_data must be a SpectralData instance")
    if not isinstance(profile, AtmosphericProfile):
        raise TypeError("profile must be an AtmosphericProfile instance")
    cfg = load_config(config)
    sp, ca, out = cfg["spectral"], cfg["calculation"], cfg["output"]
    layers = profile.build_layers()
    prob = OpticalDepthProblem(spectral_data, layers, float(sp["nu_min_cm1"]),
                               float(sp["nu_max_cm1"]), float(sp["step_cm1"]),
                               float(ca["line_cutoff_cm1"]), float(ca["tile_width_cm1"]),
                               stacklevel=4)
    mu = float(ca["mu_view"])
    dev = prob.run(precision=ca.get("precision", "fp64"), mus=(mu,),
                   want_tau=bool(out["optical_depth"]),
                   want_layers=bool(out["optical_depth_layers"]))
    res = ForwardResult(wavenumber_cm1=prob.nu, config=cfg)
    as_np = lambda t: t.double().cpu().numpy()
    if out["transmittance"]:
        res.transmittance = as_np(dev["transmittance"][0])
    if out["optical_depth"]:
        res.optical_depth = as_np(dev["optical_depth"])
    if out["optical_depth_layers"]:
        res.optical_depth_layers = as_np(dev["optical_depth_layers"])
    if out["save"]["enabled"]:
        res.save()
    return res


def _write(result: ForwardResult, path: str, fmt: str, prefix: str) -> Path:
    """npz / csv / txt writer with the reference's names and headers
    (config.py:284-319)."""
    folder = Path(path)
    folder.mkdir(parents=True, exist_ok=True)
    if fmt == "npz":
        arrays = {"wavenumber_cm1": result.wavenumber_cm1}
        for name in ("transmittance", "optical_depth", "optical_depth_layers"):
            val = getattr(result, name)
            if val is not None:
                arrays[name] = val
        target = folder / f"{prefix}.npz"
        np.savez(target, **arrays)
    else:
        sep = "," if fmt == "csv" else "\t"
        names = ["wavenumber_cm1"]
        cols = [result.wavenumber_cm1]
        for name in ("transmittance", "optical_depth"):
            val = getattr(result, name)
            if val is not None:
                names.append(name)
                cols.append(val)
        target = folder / f"{prefix}.{fmt}"
        np.savetxt(target, np.column_stack(cols), delimiter=sep, header=sep.join(names),
                   comments="")
    print(f"Saved: {target}")
    return target

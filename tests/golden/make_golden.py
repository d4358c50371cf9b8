"""Write golden vectors by running the reference implementation itself.

Runs ONLY in the development container where the reference tree is mounted
(its location is the first argument, default /root/reference).  It imports the
reference's ``pylblrtm`` package and evaluates ``compute_transmittance``,
``_build_optical_depth``, ``build_layers`` and ``run_forward`` on a handful of
tiny seeded inputs, then stores inputs and outputs as plain ``.npz`` files
next to this script.  Those are the only things committed; the reference
never travels with them.

These functions' integrated behaviour (tile arithmetic, line inclusion per
tile, group order, radiation-term placement, output assembly, warnings and
error messages) is not covered by any test of the reference, so only running
it pins them.
"""

from __future__ import annotations

import hashlib
import io
import os
import sys
import tempfile
import warnings
from contextlib import redirect_stdout
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from golden_io import pack_case  # noqa: E402


def _sha(path: Path) -> str:
    return hashlib.sha256(path.read_bytes()).hexdigest()[:16]


def _lines(rng, n, lo, hi, mol, iso=1):
    return {
        "nu0": rng.uniform(lo, hi, n),
        "Sref": 10.0 ** rng.uniform(-27.0, -19.0, n),
        "g_air": rng.uniform(0.05, 0.10, n),
        "g_self": rng.uniform(0.07, 0.12, n),
        "n_air": rng.uniform(0.6, 0.8, n),
        "shift": rng.normal(-0.002, 0.001, n),
        "Elow": rng.uniform(0.0, 3000.0, n),
        "mol_id": np.full(n, mol, np.int32),
        "iso_id": np.full(n, iso, np.int32),
    }


def _cat(parts):
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def _tips(q296=286.1):
    t = np.arange(70.0, 401.0, 1.0)
    return t, q296 * (t / 296.0) ** 1.5


def _levels(n):
    p = np.logspace(np.log10(1013.25), 0.0, n)[::-1].copy()
    t = np.linspace(216.65, 288.0, n)
    return p, t


def main(ref_root: str) -> None:
    sys.path.insert(0, ref_root)
    import pylblrtm  # noqa: F401  (the reference)
    from pylblrtm.spectral_data import SpectralData
    from pylblrtm.atmosphere import AtmosphericProfile
    from pylblrtm.forward import compute_transmittance, _build_optical_depth
    from pylblrtm.config import run_forward

    prov = {f: _sha(Path(ref_root) / "pylblrtm" / f)
            for f in ("forward.py", "atmosphere.py", "spectral_data.py", "config.py")}

    cases = []

    # A: one CO2 isotopologue, lines well outside the window included by the cutoff.
    rng = np.random.default_rng(101)
    lines = _lines(rng, 200, 700.0, 750.0, 2)
    p, t = _levels(5)
    cases.append(dict(name="co2_mini", lines=lines, tips={(2, 1): _tips()},
                      mm={(2, 1): 43.99}, p=p, t=t, vmr={"CO2": np.full(5, 4.15e-4)},
                      grid=(720.0, 724.0, 0.005), cutoff=25.0, tile=1.0, mus=(1.0, 0.5)))

    # B: three gases, origin-dependent tile_samples (199), skip/fallback paths:
    # an unknown mol_id (50), a gas absent from the profile (CO), a missing
    # molar mass (CO2 iso 2), an O3 layer with zero column.
    rng = np.random.default_rng(202)
    parts = [_lines(rng, 60, 2390.0, 2410.0, 1), _lines(rng, 60, 2390.0, 2410.0, 2),
             _lines(rng, 60, 2390.0, 2410.0, 3), _lines(rng, 10, 2399.0, 2402.0, 2, 2),
             _lines(rng, 5, 2399.0, 2402.0, 50), _lines(rng, 5, 2399.0, 2402.0, 5)]
    lines = _cat(parts)
    tips = {(1, 1): _tips(174.6), (2, 1): _tips(), (3, 1): _tips(3483.7), (2, 2): _tips(578.4),
            (5, 1): _tips(107.4)}
    mm = {(1, 1): 18.01, (2, 1): 43.99, (3, 1): 47.98, (5, 1): 27.99}
    p, t = _levels(6)
    o3 = np.array([2e-6, 7e-6, 0.0, 0.0, 3e-8, 3e-8])
    cases.append(dict(name="three_gas_edge", lines=lines, tips=tips, mm=mm, p=p, t=t,
                      vmr={"H2O": np.geomspace(4e-6, 8e-3, 6), "CO2": np.full(6, 4.15e-4), "O3": o3},
                      grid=(2400.0, 2401.0, 0.001), cutoff=5.0, tile=0.2, mus=(1.0, 0.3)))

    # C: tile narrower than the step -> one point per tile (per-point cutoff).
    rng = np.random.default_rng(303)
    lines = _lines(rng, 50, 699.5, 701.5, 2)
    p, t = _levels(3)
    cases.append(dict(name="tile1", lines=lines, tips={(2, 1): _tips()}, mm={(2, 1): 43.99},
                      p=p, t=t, vmr={"CO2": np.full(3, 4.15e-4)},
                      grid=(700.0, 701.0, 0.01), cutoff=0.3, tile=0.001, mus=(1.0,)))

    # D: a non-finite line strength in layer 0 -> FloatingPointError.
    rng = np.random.default_rng(404)
    lines = _lines(rng, 20, 700.0, 702.0, 2)
    lines["Sref"][7] = np.nan
    p, t = _levels(3)
    cases.append(dict(name="nan_strength", lines=lines, tips={(2, 1): _tips()}, mm={(2, 1): 43.99},
                      p=p, t=t, vmr={"CO2": np.full(3, 4.15e-4)},
                      grid=(700.0, 701.0, 0.01), cutoff=2.0, tile=2.0, mus=(1.0,)))

    for c in cases:
        sd = SpectralData(lines=c["lines"], tips=c["tips"], molar_masses=c["mm"])
        prof = AtmosphericProfile(c["p"], c["t"], c["vmr"])
        lo, hi, step = c["grid"]
        out = {}
        meta = dict(name=c["name"], grid=list(c["grid"]), cutoff=c["cutoff"], tile=c["tile"],
                    mus=list(c["mus"]), provenance=prov)
        layers = prof.build_layers()
        out["lay_p_atm"] = layers.pressure_atm
        out["lay_t_k"] = layers.temperature_k
        for g in layers.vmr:
            out[f"lay_vmr_{g}"] = layers.vmr[g]
            out[f"lay_col_{g}"] = layers.column_cm2[g]
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            try:
                for j, mu in enumerate(c["mus"]):
                    nu, T = compute_transmittance(sd, prof, lo, hi, step_cm1=step,
                                                  line_cutoff_cm1=c["cutoff"],
                                                  tile_width_cm1=c["tile"], mu_view=mu)
                    out[f"T_mu{j}"] = T
                out["nu"] = nu
                out["od_layers"] = _build_optical_depth(nu, sd, layers, c["cutoff"], c["tile"])
                meta["error"] = None
            except FloatingPointError as e:
                meta["error"] = ["FloatingPointError", str(e)]
        meta["warnings"] = sorted({str(x.message) for x in w})
        if meta["error"] is None and c["name"] == "co2_mini":
            with tempfile.TemporaryDirectory() as d:
                headers = {}
                for fmt in ("csv", "txt", "npz"):
                    cfg = {"spectral": {"nu_min_cm1": lo, "nu_max_cm1": hi, "step_cm1": step},
                           "calculation": {"mu_view": 0.5, "line_cutoff_cm1": c["cutoff"],
                                           "tile_width_cm1": c["tile"]},
                           "output": {"transmittance": True, "optical_depth": True,
                                      "optical_depth_layers": True,
                                      "save": {"enabled": True, "path": d, "format": fmt,
                                               "prefix": "gold"}}}
                    buf = io.StringIO()
                    with redirect_stdout(buf):
                        res = run_forward(sd, prof, cfg)
                    if fmt != "npz":
                        headers[fmt] = (Path(d) / f"gold.{fmt}").read_text().splitlines()[0]
                    headers[f"stdout_{fmt}"] = buf.getvalue().replace(d, "<DIR>")
                out["rf_transmittance"] = res.transmittance
                out["rf_optical_depth"] = res.optical_depth
                out["rf_optical_depth_layers"] = res.optical_depth_layers
                meta["run_forward"] = headers
                meta["repr"] = repr(res)
        arrays = pack_case(c["lines"], c["tips"], c["mm"], c["p"], c["t"], c["vmr"], meta)
        arrays.update({f"out_{k}": v for k, v in out.items()})
        np.savez_compressed(HERE / f"{c['name']}.npz", **arrays)
        print(c["name"], "error" if meta["error"] else "ok", len(meta["warnings"]), "warnings")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

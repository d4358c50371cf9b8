"""Flatten / restore the inputs of a forward-model case into a plain npz
(no pickles), so fixtures load with ``numpy.load(allow_pickle=False)``."""

from __future__ import annotations

import json
from typing import Dict, Tuple

import numpy as np

LINE_KEYS = ("nu0", "Sref", "g_air", "g_self", "Elow", "n_air", "shift", "mol_id", "iso_id")


def pack_case(lines, tips, molar_masses, p_hpa, t_k, vmr, meta: dict) -> Dict[str, np.ndarray]:
    out = {}
    for k in LINE_KEYS:
        out[f"line_{k}"] = np.asarray(lines[k])
    tip_keys = sorted(tips)
    out["tips_keys"] = np.array(tip_keys, np.int64).reshape(-1, 2)
    for m, i in tip_keys:
        out[f"tips_T_{m}_{i}"] = np.asarray(tips[(m, i)][0], np.float64)
        out[f"tips_Q_{m}_{i}"] = np.asarray(tips[(m, i)][1], np.float64)
    mm_keys = sorted(molar_masses)
    out["mm_keys"] = np.array(mm_keys, np.int64).reshape(-1, 2)
    out["mm_vals"] = np.array([molar_masses[k] for k in mm_keys], np.float64)
    out["p_hpa"] = np.asarray(p_hpa, np.float64)
    out["t_k"] = np.asarray(t_k, np.float64)
    gases = list(vmr)
    out["vmr_gases"] = np.array(gases)
    for g in gases:
        out[f"vmr_{g}"] = np.asarray(vmr[g], np.float64)
    out["meta_json"] = np.array(json.dumps(meta))
    return out


def unpack_case(z) -> Tuple[dict, dict, dict, np.ndarray, np.ndarray, dict, dict]:
    lines = {k: z[f"line_{k}"] for k in LINE_KEYS}
    tips = {}
    for m, i in z["tips_keys"].reshape(-1, 2):
        tips[(int(m), int(i))] = (z[f"tips_T_{m}_{i}"], z[f"tips_Q_{m}_{i}"])
    mm = {(int(m), int(i)): float(v) for (m, i), v in zip(z["mm_keys"].reshape(-1, 2), z["mm_vals"])}
    vmr = {str(g): z[f"vmr_{g}"] for g in z["vmr_gases"]}
    meta = json.loads(str(z["meta_json"]))
    return lines, tips, mm, z["p_hpa"], z["t_k"], vmr, meta

"""``from_files``: a SpectralData from a HITRAN .par file and TIPS tables
(pylblrtm/data_loader.py:92-175).  ``from_hapi`` (download from
HITRANonline) is out of scope for this package: it needs the network."""

from __future__ import annotations

from pathlib import Path
from typing import Dict, Optional, Tuple, Union

import numpy as np

from . import hitran_manager as hm
from .molecules import MOL_ID_TO_NAME
from .spectral_data import SpectralData


def from_files(par_path: Union[str, Path],
               tips: Dict[tuple, Union[str, Path, Tuple[np.ndarray, np.ndarray]]],
               molar_masses: Optional[Dict[tuple, float]] = None) -> SpectralData:
    """Lines from ``par_path`` (any mix of molecules and isotopologues), TIPS
    per (mol_id, iso_id) as a path to a two-column file or a (T, Q) pair.
    Every isotopologue in the file needs a TIPS entry; entries for absent
    ones are ignored.  ``molar_masses`` [g/mol] per (mol_id, iso_id); without
    it the masses come from ``hitran_manager.get_molar_mass``."""
    lines = hm.par_to_dict(par_path)
    present = set(zip(lines["mol_id"].tolist(), lines["iso_id"].tolist()))
    missing = present - set(tips.keys())
    if missing:
        raise ValueError(
            f"Missing TIPS for isotopologue(s): {sorted(missing)}. "
            "Provide a path or (T, Q) arrays for each (mol_id, iso_id) in the .par file.")
    loaded_tips: Dict[tuple, Tuple[np.ndarray, np.ndarray]] = {}
    for key, val in tips.items():
        if key not in present:
            continue
        if isinstance(val, (str, Path)):
            loaded_tips[key] = hm.txt_to_tips(val)
        else:
            t_arr, q_arr = val
            loaded_tips[key] = (np.asarray(t_arr, dtype=np.float64),
                                np.asarray(q_arr, dtype=np.float64))
    if molar_masses is not None:
        masses = {k: float(v) for k, v in molar_masses.items()}
    else:
        masses = {}
        for mol_id, iso in present:
            name = MOL_ID_TO_NAME.get(mol_id)
            if name is None:
                raise ValueError(f"mol_id={mol_id} not in the internal catalogue. "
                                 "Pass molar_masses manually.")
            masses[(mol_id, iso)] = hm.get_molar_mass(name, iso)
    mol_ids = set(k[0] for k in present)
    molecule = MOL_ID_TO_NAME.get(next(iter(mol_ids))) if len(mol_ids) == 1 else None
    return SpectralData(lines=lines, tips=loaded_tips, molar_masses=masses, molecule=molecule)

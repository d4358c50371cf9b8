"""ctypes binding of libpylblrtm_amd.so (the C ABI declared in
include/pylblrtm_amd.h) and the mapping of its status codes onto the
exceptions the reference raises.

There is no CPU fallback: if the library is missing or no ROCm device is
visible, every compute entry point raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional

LIB_PATH = Path(__file__).resolve().parent / "libpylblrtm_amd.so"

LBL_OK, LBL_EINVAL, LBL_ESTRENGTH, LBL_ENONFINITE, LBL_EHIP = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": 0, "mixed": 1, "surrogate": 2}

EXPORTED = ("lbl_workspace_bytes", "lbl_optical_depth", "lbl_voigt_eval", "lbl_last_error",
            "lbl_enable_timing", "lbl_last_timings", "lbl_last_timings_split", "lbl_vjp_workspace_bytes",
            "lbl_transmittance_vjp", "lbl_mlp_num_params", "lbl_mlp_make_data", "lbl_mlp_pack",
            "lbl_mlp_train_workspace_bytes", "lbl_mlp_train_step", "lbl_mlp_eval",
            "lbl_optical_depth_surrogate", "lbl_par_read", "lbl_optical_depth_noalloc",
            "lbl_grid_from_nu", "lbl_enable_work_count", "lbl_last_work_count",
            "lbl_last_timing_epilogue")

_vp = ctypes.c_void_p


class LblLines(ctypes.Structure):
    _fields_ = [("n_lines", ctypes.c_int64), ("nu0", _vp), ("sref", _vp), ("g_air", _vp),
                ("g_self", _vp), ("elow", _vp), ("n_air", _vp), ("shift", _vp), ("group", _vp)]


class LblGroups(ctypes.Structure):
    _fields_ = [("n_groups", ctypes.c_int32), ("gas", _vp), ("molar_mass", _vp),
                ("tips_offset", _vp), ("tips_t", _vp), ("tips_q", _vp)]


class LblLayers(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int32), ("n_gases", ctypes.c_int32), ("p_atm", _vp),
                ("t_k", _vp), ("vmr", _vp), ("col", _vp)]


class LblGrid(ctypes.Structure):
    _fields_ = [("nu_min", ctypes.c_double), ("step", ctypes.c_double), ("n_nu", ctypes.c_int64),
                ("tile_samples", ctypes.c_int64), ("cutoff", ctypes.c_double),
                ("shift_bound", ctypes.c_double)]


class LblSurrogate(ctypes.Structure):
    _fields_ = [("params", _vp), ("wbf16", _vp), ("gate", ctypes.c_double),
                ("n_network", _vp), ("n_exact_near", _vp)]


class LblOutputs(ctypes.Structure):
    _fields_ = [("n_mu", ctypes.c_int32), ("mu", ctypes.POINTER(ctypes.c_double)),
                ("transmittance", _vp), ("optical_depth", _vp), ("optical_depth_layers", _vp)]


_lib: Optional[ctypes.CDLL] = None


class LibraryMissing(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises LibraryMissing if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LibraryMissing(
            f"{LIB_PATH.name} is not built; run `python -m pylblrtm_amd.build` "
            f"(the HIP path has no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    lib.lbl_workspace_bytes.argtypes = [P(LblLines), P(LblLayers), P(LblGrid), ctypes.c_int,
                                        P(ctypes.c_size_t)]
    lib.lbl_workspace_bytes.restype = ctypes.c_int
    lib.lbl_optical_depth.argtypes = [P(LblLines), P(LblGroups), P(LblLayers), P(LblGrid),
                                      ctypes.c_int, P(LblOutputs), _vp, ctypes.c_size_t,
                                      P(ctypes.c_int32), P(ctypes.c_int32), _vp]
    lib.lbl_optical_depth.restype = ctypes.c_int
    lib.lbl_optical_depth_noalloc.argtypes = [P(LblLines), P(LblGroups), P(LblLayers), P(LblGrid),
                                              ctypes.c_int, P(LblOutputs), P(ctypes.c_int32),
                                              P(ctypes.c_int32), _vp]
    lib.lbl_optical_depth_noalloc.restype = ctypes.c_int
    lib.lbl_optical_depth_surrogate.argtypes = [P(LblLines), P(LblGroups), P(LblLayers),
                                                P(LblGrid), P(LblOutputs), P(LblSurrogate), _vp,
                                                ctypes.c_size_t, P(ctypes.c_int32),
                                                P(ctypes.c_int32), _vp]
    lib.lbl_optical_depth_surrogate.restype = ctypes.c_int
    lib.lbl_voigt_eval.argtypes = [ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int, _vp]
    lib.lbl_voigt_eval.restype = ctypes.c_int
    lib.lbl_vjp_workspace_bytes.argtypes = [P(LblLines), P(LblLayers), P(LblGrid),
                                            P(ctypes.c_size_t)]
    lib.lbl_vjp_workspace_bytes.restype = ctypes.c_int
    lib.lbl_transmittance_vjp.argtypes = [P(LblLines), P(LblGroups), P(LblLayers), P(LblGrid),
                                          ctypes.c_int32, P(ctypes.c_double), _vp, _vp, _vp, _vp,
                                          _vp, _vp, ctypes.c_size_t, _vp]
    lib.lbl_transmittance_vjp.restype = ctypes.c_int
    lib.lbl_mlp_num_params.argtypes = []
    lib.lbl_mlp_num_params.restype = ctypes.c_int
    lib.lbl_mlp_make_data.argtypes = [ctypes.c_int64, ctypes.c_uint64, _vp, _vp, _vp, _vp]
    lib.lbl_mlp_make_data.restype = ctypes.c_int
    lib.lbl_mlp_pack.argtypes = [_vp, _vp, _vp]
    lib.lbl_mlp_pack.restype = ctypes.c_int
    lib.lbl_mlp_train_workspace_bytes.argtypes = [ctypes.c_int, P(ctypes.c_size_t)]
    lib.lbl_mlp_train_workspace_bytes.restype = ctypes.c_int
    lib.lbl_mlp_train_step.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64,
                                       ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                       ctypes.c_int, ctypes.c_float, ctypes.c_int, _vp, _vp, _vp,
                                       ctypes.c_size_t, _vp]
    lib.lbl_mlp_train_step.restype = ctypes.c_int
    lib.lbl_mlp_eval.argtypes = [ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp]
    lib.lbl_mlp_eval.restype = ctypes.c_int
    lib.lbl_enable_timing.argtypes = [ctypes.c_int]
    lib.lbl_enable_timing.restype = None
    lib.lbl_last_timings.argtypes = [P(ctypes.c_double), P(ctypes.c_double)]
    lib.lbl_last_timings.restype = ctypes.c_int
    lib.lbl_last_timings_split.argtypes = [P(ctypes.c_double), P(ctypes.c_double)]
    lib.lbl_last_timings_split.restype = ctypes.c_int
    lib.lbl_par_read.argtypes = [ctypes.c_char_p, ctypes.c_int64, P(ctypes.c_int64), _vp, _vp,
                                 _vp]
    lib.lbl_par_read.restype = ctypes.c_int
    lib.lbl_grid_from_nu.argtypes = [_vp, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                     P(LblGrid)]
    lib.lbl_grid_from_nu.restype = ctypes.c_int
    lib.lbl_last_timing_epilogue.argtypes = [P(ctypes.c_double)]
    lib.lbl_last_timing_epilogue.restype = ctypes.c_int
    lib.lbl_enable_work_count.argtypes = [ctypes.c_int]
    lib.lbl_enable_work_count.restype = None
    lib.lbl_last_work_count.argtypes = [P(ctypes.c_uint64), ctypes.c_int32]
    lib.lbl_last_work_count.restype = ctypes.c_int
    lib.lbl_last_error.argtypes = []
    lib.lbl_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def last_error() -> str:
    return load().lbl_last_error().decode(errors="replace")


def raise_for(status: int, what: str) -> None:
    """Map a non-strength status code to the reference's exception types."""
    if status == LBL_OK:
        return
    msg = f"{what}: {last_error()}"
    if status == LBL_EINVAL:
        raise ValueError(msg)
    if status in (LBL_ESTRENGTH, LBL_ENONFINITE):
        raise FloatingPointError(msg)
    raise RuntimeError(msg)


def grid_from_nu(nu, tile_width: float, cutoff: float) -> LblGrid:
    """lbl_grid_from_nu on a host float64 array: the grid struct whose points
    are exactly ``nu`` (ValueError if no evenly spaced grid reproduces it)."""
    import numpy as np
    a = np.ascontiguousarray(nu, dtype=np.float64)
    g = LblGrid()
    st = load().lbl_grid_from_nu(a.ctypes.data if a.size else None, int(a.size), float(tile_width),
                                 float(cutoff), ctypes.byref(g))
    raise_for(st, "lbl_grid_from_nu")
    return g


WORK_COUNT_NAMES = ("near_evals", "slice_series", "wave_series", "slice_exceptions",
                    "wave_exceptions", "exception_evals", "queued", "core_points", "rescans",
                    "wave_from_slice_list", "group_series", "group_deferred", "group_deferred_evals",
                    "group_series_middle", "series_terms", "series_terms_2")


def enable_work_count(on: bool = True) -> None:
    load().lbl_enable_work_count(1 if on else 0)


def last_work_count() -> dict:
    """The line-sum kernels' own work counts of the calling thread's last
    lbl_optical_depth call made with work counting enabled (names above)."""
    arr = (ctypes.c_uint64 * len(WORK_COUNT_NAMES))()
    raise_for(load().lbl_last_work_count(arr, len(WORK_COUNT_NAMES)), "lbl_last_work_count")
    return {n: int(v) for n, v in zip(WORK_COUNT_NAMES, arr)}


def enable_timing(on: bool = True) -> None:
    load().lbl_enable_timing(1 if on else 0)


def last_timings():
    """(ms of the per-(layer, line) kernel, ms of the line-sum kernel) of the
    calling thread's last lbl_optical_depth call with timing enabled."""
    a, b = ctypes.c_double(0), ctypes.c_double(0)
    load().lbl_last_timings(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def last_timings_split():
    """(ms of the far-field series kernel, ms of the near-field line sum and
    epilogue) of the calling thread's last lbl_optical_depth call with timing
    enabled; together they make up last_timings()[1]."""
    a, b = ctypes.c_double(0), ctypes.c_double(0)
    load().lbl_last_timings_split(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def last_timing_epilogue() -> float:
    """ms of the layer-group epilogue (inside last_timings_split()[1])."""
    a = ctypes.c_double(0)
    load().lbl_last_timing_epilogue(ctypes.byref(a))
    return a.value

"""Device-resident problem: the host prelude of the reference's
``_build_optical_depth`` plus the upload of everything the HIP kernels read.

``OpticalDepthProblem`` is built once per (spectral data, layers, grid,
cutoff, tile width) and can be run any number of times (the benchmark times
``run``; ``compute_transmittance`` / ``run_forward`` build and run once).

Host prelude = pylblrtm/forward.py:171-245: tile size from the first two grid
points, the per-line shift bound, groups in sorted (mol_id, iso_id) order
with the same warnings and skips, the conservative band mask, the 28.97 g/mol
molar-mass fallback, and the KeyError of a missing partition-sum table.  The
partition-sum ratios themselves are interpolated on the device.
"""

from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi
from .molecules import MOL_ID_TO_NAME

T_REF = 296.0


def make_grid(nu_min: float, nu_max: float, step: float) -> np.ndarray:
    """forward.py:141-142."""
    n = int(np.ceil((nu_max - nu_min) / step)) + 1
    return nu_min + np.arange(n, dtype=np.float64) * step


def tile_samples_for(nu: np.ndarray, tile_width: float) -> int:
    """forward.py:175-176 (depends on the grid origin through nu[1] - nu[0])."""
    step = float(nu[1] - nu[0]) if nu.size > 1 else 0.001
    return max(1, int(np.floor(tile_width / step + 1e-8)))


@dataclass
class Group:
    mol_id: int
    iso_id: int
    gas: str
    index: np.ndarray        # positions in SpectralData.lines
    molar_mass: float
    tips: Tuple[np.ndarray, np.ndarray]


def prepare_groups(nu: np.ndarray, sd, layers, cutoff: float, stacklevel: int = 4) -> List[Group]:
    """forward.py:185-245 without the q-ratio arithmetic (done on the device)."""
    lines = sd.lines
    nu0 = np.asarray(lines["nu0"], np.float64)
    delta = np.asarray(lines["shift"], np.float64)
    mol = np.asarray(lines["mol_id"])
    iso = np.asarray(lines["iso_id"])
    p_max = float(layers.pressure_atm.max()) if layers.n_layers > 0 else 1.0
    bound = np.abs(delta) * p_max
    band = (nu0 >= nu[0] - cutoff - bound) & (nu0 <= nu[-1] + cutoff + bound)
    # (mol_id, iso_id) pairs in sorted order: one int64 key, monotonic in the
    # pair for any 32-bit ids, so np.unique sorts exactly as sorted(set(...))
    pair_key = (mol.astype(np.int64) << 32) + (iso.astype(np.int64) + (1 << 31))
    uniq = np.unique(pair_key)
    groups: List[Group] = []
    for kv in uniq.tolist():
        m, i = kv >> 32, (kv & 0xFFFFFFFF) - (1 << 31)
        gas = MOL_ID_TO_NAME.get(int(m))
        if gas is None:
            warnings.warn(f"mol_id={m} is not in the HITRAN catalogue; lines skipped.",
                          RuntimeWarning, stacklevel=stacklevel)
            continue
        if gas not in layers.column_cm2:
            warnings.warn(
                f"Gas '{gas}' (mol_id={m}, iso_id={i}) is in the spectral data but not in the "
                f"atmospheric profile vmr; its lines are skipped.",
                RuntimeWarning, stacklevel=stacklevel)
            continue
        sel = band & (pair_key == kv)
        if not sel.any():
            continue
        key = (int(m), int(i))
        mm = sd.molar_masses.get(key)
        if mm is None:
            warnings.warn(f"No molar mass found for (mol_id={m}, iso_id={i}); "
                          f"using dry-air mass as fallback.", RuntimeWarning, stacklevel=stacklevel)
            mm = 28.97
        if key not in sd.tips:
            raise KeyError(f"No TIPS data for (mol_id={m}, iso_id={i}). "
                           f"Available: {sorted(sd.tips.keys())}")
        t_nodes, q_nodes = sd.tips[key]
        # the reference's first use of the table, q_ref = Q(296) (forward.py:236,
        # spectral_data.py:52), on the host: an empty table or one whose T and Q
        # differ in length (txt_to_tips keeps a row's T when its Q fails) raises
        # numpy's own ValueError here, as there; the kernels only see valid tables
        np.interp(T_REF, np.asarray(t_nodes, np.float64), np.asarray(q_nodes, np.float64))
        groups.append(Group(int(m), int(i), gas, np.flatnonzero(sel), float(mm),
                            (np.asarray(t_nodes, np.float64), np.asarray(q_nodes, np.float64))))
    return groups


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pylblrtm_amd needs a ROCm GPU: torch.cuda.is_available() is False "
                           "(the HIP path has no CPU fallback)")
    return torch


class OpticalDepthProblem:
    """Inputs of the device path, uploaded once."""

    def __init__(self, sd, layers, nu_min: Optional[float], nu_max: Optional[float],
                 step: Optional[float], cutoff: float, tile_width: float, device=None,
                 stacklevel: int = 4, nu: Optional[np.ndarray] = None):
        """The grid is either (nu_min, nu_max, step), built as
        compute_transmittance builds it (forward.py:141-142), or ``nu``: the
        wavenumber array the seam _build_optical_depth receives
        (forward.py:157-176), used as given (its size, its tile size from
        nu[1] - nu[0], and a step that reproduces its points bit for bit;
        lbl_grid_from_nu); nu_min, nu_max and step are then ignored."""
        torch = _torch()
        self.lib = _capi.load()
        self.device = torch.device(device if device is not None else f"cuda:{torch.cuda.current_device()}")
        if nu is not None:
            self.nu = np.array(nu, dtype=np.float64, copy=True).reshape(-1)
            g = _capi.grid_from_nu(self.nu, tile_width, cutoff)
            self.nu_min, self.step = float(g.nu_min), float(g.step)
            self.tile_samples = int(g.tile_samples)
        else:
            self.nu = make_grid(nu_min, nu_max, step)
            self.nu_min, self.step = float(nu_min), float(step)
            self.tile_samples = tile_samples_for(self.nu, tile_width)
        self.n_nu = self.nu.size
        self.cutoff = float(cutoff)
        self.n_layers = layers.n_layers
        self.groups = prepare_groups(self.nu, sd, layers, cutoff, stacklevel=stacklevel)

        gases: List[str] = []
        for g in self.groups:
            if g.gas not in gases:
                gases.append(g.gas)
        self.gases = gases
        lines = sd.lines

        def dev(a, dtype):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).to(self.device)

        f8 = np.float64
        self._t = {}
        # The surviving lines, ordered by centre (stable: lines with equal
        # centres keep the group order), gathered on the device.
        if self.groups:
            idx = dev(np.concatenate([g.index for g in self.groups]), np.int64)
            gid = dev(np.concatenate([np.full(g.index.size, j, np.int32)
                                      for j, g in enumerate(self.groups)]), np.int32)
            full = {key: dev(lines[key], f8) for key in
                    ("nu0", "Sref", "g_air", "g_self", "Elow", "n_air", "shift")}
            order = torch.sort(full["nu0"][idx], stable=True).indices
            idx, gid = idx[order], gid[order]
            for key, arr in full.items():
                self._t[key] = arr[idx].contiguous()
            del full
        else:
            idx = torch.zeros(0, dtype=torch.int64, device=self.device)
            gid = torch.zeros(0, dtype=torch.int32, device=self.device)
            for key in ("nu0", "Sref", "g_air", "g_self", "Elow", "n_air", "shift"):
                self._t[key] = torch.zeros(0, dtype=torch.float64, device=self.device)
        self._t["group"] = gid.contiguous()
        self.n_lines = int(idx.numel())
        self.line_index = idx.cpu().numpy()
        self.layers = layers
        p_max = float(layers.pressure_atm.max()) if self.n_layers > 0 else 1.0
        self.shift_bound = (float(self._t["shift"].abs().max().item()) * p_max
                            if self.n_lines else 0.0)
        self._host = None
        offs = np.zeros(len(self.groups) + 1, np.int32)
        for j, g in enumerate(self.groups):
            offs[j + 1] = offs[j] + g.tips[0].size
        cat = lambda k: (np.concatenate([g.tips[k] for g in self.groups]) if self.groups
                         else np.zeros(1))
        self._t["gas"] = dev(np.array([gases.index(g.gas) for g in self.groups] or [0], np.int32),
                             np.int32)
        self._t["mm"] = dev(np.array([g.molar_mass for g in self.groups] or [1.0]), f8)
        self._t["toff"] = dev(offs, np.int32)
        self._t["tips_t"] = dev(cat(0), f8)
        self._t["tips_q"] = dev(cat(1), f8)
        nl = max(self.n_layers, 1)
        vmr = np.zeros((max(len(gases), 1), nl))
        col = np.zeros((max(len(gases), 1), nl))
        for j, gname in enumerate(gases):
            vmr[j, : self.n_layers] = layers.vmr[gname]
            col[j, : self.n_layers] = layers.column_cm2[gname]
        # air column per layer (col = vmr * air, atmosphere.py:133-139); used when
        # layer mixing ratios are changed and for the gradient at vmr = 0.  This
        # package's layers carry it; for others (the reference's _Layers) it is
        # recovered from the gases' columns, and stays NaN in a layer where
        # every mixing ratio is 0 (that layer's columns cannot be changed then)
        air = getattr(layers, "air_cm2", None)
        if air is not None:
            self.air_cm2 = np.asarray(air, np.float64).reshape(-1)[: self.n_layers].copy()
        else:
            num = np.zeros(nl)
            den = np.zeros(nl)
            for gname in layers.vmr:
                if gname in layers.column_cm2:
                    num[: self.n_layers] += layers.column_cm2[gname]
                    den[: self.n_layers] += layers.vmr[gname]
            self.air_cm2 = np.where(den > 0, num / np.where(den > 0, den, 1.0),
                                    np.nan)[: self.n_layers]
        self._t["p_atm"] = dev(layers.pressure_atm if self.n_layers else np.zeros(1), f8)
        self._t["t_k"] = dev(layers.temperature_k if self.n_layers else np.ones(1), f8)
        self._t["vmr"] = dev(vmr, f8)
        self._t["col"] = dev(col, f8)

        p = lambda name: self._t[name].data_ptr()
        self.c_lines = _capi.LblLines(self.n_lines, p("nu0"), p("Sref"), p("g_air"), p("g_self"),
                                      p("Elow"), p("n_air"), p("shift"), p("group"))
        self.c_groups = _capi.LblGroups(len(self.groups), p("gas"), p("mm"), p("toff"),
                                        p("tips_t"), p("tips_q"))
        self.c_layers = _capi.LblLayers(self.n_layers, len(gases), p("p_atm"), p("t_k"),
                                        p("vmr"), p("col"))
        self.c_grid = _capi.LblGrid(self.nu_min, self.step, self.n_nu, self.tile_samples,
                                    self.cutoff, self.shift_bound)
        self._ws = {}

    # host copies of the sorted line centres, shifts and groups, fetched once
    # on first use (the evaluation count needs them; the line sums do not)
    def _host_lines(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        if self._host is None:
            self._host = (self._t["nu0"].cpu().numpy(), self._t["shift"].cpu().numpy(),
                          self._t["group"].cpu().numpy())
        return self._host

    # ------------------------------------------------------------ layers --
    def set_layers(self, temperature_k, vmr) -> None:
        """Replace the layer temperatures [n_layers] and mixing ratios
        [n_gases, n_layers] (rows in ``self.gases`` order) on the device; the
        columns become vmr * air column.  Accepts numpy arrays or device
        tensors."""
        torch = _torch()
        t = torch.as_tensor(temperature_k, dtype=torch.float64, device=self.device).reshape(-1)
        v = torch.as_tensor(vmr, dtype=torch.float64, device=self.device)
        if t.numel() != self.n_layers or tuple(v.shape) != (len(self.gases), self.n_layers):
            raise ValueError(f"expected temperature [{self.n_layers}] and vmr "
                             f"[{len(self.gases)}, {self.n_layers}]")
        unknown = ~np.isfinite(self.air_cm2)
        if unknown.any() and bool((v[:, torch.as_tensor(unknown, device=self.device)] != 0).any()):
            raise ValueError("a layer whose every mixing ratio was 0 has no known air column; "
                             "build the layers with AtmosphericProfile.build_layers")
        air = torch.as_tensor(np.nan_to_num(self.air_cm2, nan=0.0), dtype=torch.float64,
                              device=self.device)
        self._t["t_k"][: self.n_layers].copy_(t)
        self._t["vmr"][: len(self.gases), : self.n_layers].copy_(v)
        self._t["col"][: len(self.gases), : self.n_layers].copy_(v * air[None, :])

    def layer_state(self):
        """(temperature [n_layers], vmr [n_gases, n_layers]) currently on the device."""
        return (self._t["t_k"][: self.n_layers].clone(),
                self._t["vmr"][: len(self.gases), : self.n_layers].clone())

    def vjp(self, transmittance, optical_depth_layers, grad_transmittance, mus: Sequence[float]):
        """dL/dT_k [n_layers] and dL/dvmr [n_gases, n_layers] (float64 device
        tensors) for a mixed-precision forward (``transmittance`` [n_mu, n_nu]
        and ``optical_depth_layers`` [n_layers, n_nu] of one
        ``run("mixed", want_layers=True)``) and the upstream gradient
        ``grad_transmittance`` [n_mu, n_nu]."""
        torch = _torch()
        T = transmittance.to(torch.float32).contiguous()
        OD = optical_depth_layers.to(torch.float32).contiguous()
        G = grad_transmittance.to(device=self.device, dtype=torch.float32).contiguous()
        if tuple(T.shape) != (len(mus), self.n_nu) or tuple(G.shape) != tuple(T.shape):
            raise ValueError("transmittance / gradient must be [n_mu, n_nu]")
        if tuple(OD.shape) != (self.n_layers, self.n_nu):
            raise ValueError("optical_depth_layers must be [n_layers, n_nu]")
        if "vjp" not in self._ws:
            nbytes = ctypes.c_size_t(0)
            st = self.lib.lbl_vjp_workspace_bytes(ctypes.byref(self.c_lines),
                                                  ctypes.byref(self.c_layers),
                                                  ctypes.byref(self.c_grid), ctypes.byref(nbytes))
            _capi.raise_for(st, "lbl_vjp_workspace_bytes")
            self._ws["vjp"] = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8,
                                          device=self.device)
        ws = self._ws["vjp"]
        gt = torch.zeros(max(self.n_layers, 1), dtype=torch.float64, device=self.device)
        gv = torch.zeros((max(len(self.gases), 1), max(self.n_layers, 1)), dtype=torch.float64,
                         device=self.device)
        mu_arr = (ctypes.c_double * len(mus))(*[float(m) for m in mus])
        stream = torch.cuda.current_stream(self.device).cuda_stream
        with torch.cuda.device(self.device):
            st = self.lib.lbl_transmittance_vjp(
                ctypes.byref(self.c_lines), ctypes.byref(self.c_groups), ctypes.byref(self.c_layers),
                ctypes.byref(self.c_grid), len(mus), mu_arr, T.data_ptr(), OD.data_ptr(),
                G.data_ptr(), gt.data_ptr(), gv.data_ptr(), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                ctypes.c_void_p(stream))
        _capi.raise_for(st, "lbl_transmittance_vjp")
        gt, gv = gt[: self.n_layers], gv[: len(self.gases), : self.n_layers]
        self._vjp_at_zero_vmr(T, G, mus, gv)
        return gt, gv

    def _vjp_at_zero_vmr(self, T, G, mus: Sequence[float], gv) -> None:
        """dL/dvmr_{k,g} where vmr_{k,g} = 0 (into ``gv``).

        The forward skips such a gas in that layer (col = 0, forward.py:253-255)
        but is continuous there: the strength is proportional to col =
        vmr * air (forward.py:270-278), and every other factor of its optical
        depth is multiplied by it.  So the one-sided derivative at vmr = 0 is
        d tau(nu)/d vmr = od_unit_k(nu), the layer optical depth of that gas
        with col = air_k and no self-broadening or self-shift (P_self = 0),
        and dL/dvmr = sum_nu W(nu) od_unit_k(nu), W = -sum_m G_m T_m / mu_m.
        One mixed forward per gas that has such layers, on the device."""
        torch = _torch()
        ng, nl = len(self.gases), self.n_layers
        if ng == 0 or nl == 0:
            return
        vmr = self._t["vmr"][:ng, :nl]
        air = torch.as_tensor(self.air_cm2, dtype=torch.float64, device=self.device)
        zero = (vmr == 0) & torch.isfinite(air)[None, :] & (air > 0)[None, :]
        if not bool(zero.any()):
            return
        inv_mu = torch.tensor([1.0 / float(m) for m in mus], dtype=torch.float64,
                              device=self.device)
        W = -(G.double() * T.double() * inv_mu[:, None]).sum(0)
        col = self._t["col"][:ng, :nl]
        keep = col.clone()
        try:
            for g in range(ng):
                if not bool(zero[g].any()):
                    continue
                unit = torch.zeros_like(keep)
                unit[g] = torch.where(zero[g], air, torch.zeros_like(air))
                col.copy_(unit)
                od = self.run("mixed", mus=(1.0,), want_layers=True)["optical_depth_layers"]
                d = od.double() @ W
                gv[g] = torch.where(zero[g], d, gv[g])
        finally:
            col.copy_(keep)

    # ------------------------------------------------------------ work count --
    def tile_edges(self):
        """(first point, end point, window low, window high) per reference tile."""
        ts = self.tile_samples
        t0 = np.arange(0, self.n_nu, ts, dtype=np.int64)
        t1 = np.minimum(t0 + ts, self.n_nu)
        return t0, t1, self.nu[t0] - self.cutoff, self.nu[t1 - 1] + self.cutoff

    def count_evaluations(self, layers: Optional[Sequence[int]] = None,
                          tiles: Optional[Sequence[int]] = None) -> int:
        """Exact number of (layer, line, grid point) Voigt evaluations the
        reference performs: sum over layers and tiles of the lines inside the
        tile window times the tile's points (forward.py:248-309), for groups
        with a non-zero column."""
        t0, t1, lo_b, hi_b = self.tile_edges()
        if tiles is not None:
            sel = np.asarray(tiles, np.int64)
            t0, t1, lo_b, hi_b = t0[sel], t1[sel], lo_b[sel], hi_b[sel]
        npts = t1 - t0
        total = 0
        h_nu0, h_shift, h_group = self._host_lines()
        ks = range(self.n_layers) if layers is None else layers
        lay = self.layers
        for k in ks:
            P = float(lay.pressure_atm[k])
            for j, g in enumerate(self.groups):
                if float(lay.column_cm2[g.gas][k]) == 0.0:
                    continue
                m = h_group == j
                p_self = float(lay.vmr[g.gas][k]) * P
                c = np.sort(h_nu0[m] + h_shift[m] * (P - p_self))
                cnt = np.searchsorted(c, hi_b, side="right") - np.searchsorted(c, lo_b, side="left")
                total += int(np.sum(np.maximum(cnt, 0) * npts))
        return total

    def count_series_work(self, precision: str = "mixed") -> dict:
        """Work units of the line sums, split as the kernels do.  A slice of
        SLICE_POINTS points is four wavefronts of WAVE_POINTS, a wavefront
        eight groups of GROUP_POINTS; the near range of a slice, wavefront or
        group is the lines whose nu0 lies within SERIES_RATIO times its
        half-span (+ the shift bound) of its middle point.  slice_pairs:
        series summed once per (layer, line, slice) outside the slice's near
        range; wave_pairs: once per (layer, line, wavefront) between the two
        near ranges; group_pairs = (layer, line, group) with the line in the
        wavefront's near range but not the group's, which the group series
        takes unless its form fails at the group (the one-node and middle-tier
        forms mixed, the four-node form FP64; the pairs it leaves to the near
        field only the kernels count); far_pairs: their total.  near_evals =
        (layer, line in the group's near range, point of the group) triples,
        evaluated point by point.  Groups of lines with a zero column are
        left out (they add 0)."""
        wave_pts, slice_pts, group_pts, ratio = 512, 2048, 64, 3.0
        t0, t1, lo_b, hi_b = self.tile_edges()
        h_nu0, h_shift, h_group = self._host_lines()
        margin = abs(self.shift_bound) * (1 + 1e-9) + 1e-9
        spt = (self.tile_samples + slice_pts - 1) // slice_pts

        def units(pts_per_unit):
            offs = np.arange(spt * (slice_pts // pts_per_unit), dtype=np.int64) * pts_per_unit
            first = (t0[:, None] + offs[None, :]).ravel()
            tile_of = np.repeat(np.arange(len(t0)), len(offs))
            live = first < t1[tile_of]
            first, tile_of = first[live], tile_of[live]
            mid = self.nu[first] + 0.5 * (pts_per_unit - 1) * self.step
            half = ratio * 0.5 * pts_per_unit * self.step + margin
            return first, tile_of, mid, half

        w_first, w_tile, w_mid, w_half = units(wave_pts)
        s_first, s_tile, s_mid, s_half = units(slice_pts)
        g_first, g_tile, g_mid, g_half = units(group_pts)
        g_pts = np.minimum(t1[g_tile] - g_first, group_pts)
        # slice of each wavefront, wavefront of each group
        w_slice = np.searchsorted(s_first, w_first, side="right") - 1
        g_wave = np.searchsorted(w_first, g_first, side="right") - 1
        near_evals = slice_pairs = wave_pairs = group_pairs = 0
        lay = self.layers
        for k in range(self.n_layers):
            P = float(lay.pressure_atm[k])
            win = np.zeros(len(t0), np.int64)
            nu0_on = []
            for j, g in enumerate(self.groups):
                if float(lay.column_cm2[g.gas][k]) == 0.0:
                    continue
                m = h_group == j
                p_self = float(lay.vmr[g.gas][k]) * P
                c = np.sort(h_nu0[m] + h_shift[m] * (P - p_self))
                win += np.maximum(np.searchsorted(c, hi_b, side="right")
                                  - np.searchsorted(c, lo_b, side="left"), 0)
                nu0_on.append(h_nu0[m])
            if not nu0_on:
                continue
            nu0 = np.sort(np.concatenate(nu0_on))
            s_near = np.minimum(np.searchsorted(nu0, s_mid + s_half, side="right")
                                - np.searchsorted(nu0, s_mid - s_half, side="left"), win[s_tile])
            w_near = np.minimum(np.searchsorted(nu0, w_mid + w_half, side="right")
                                - np.searchsorted(nu0, w_mid - w_half, side="left"), s_near[w_slice])
            slice_pairs += int(np.sum(win[s_tile] - s_near))
            wave_pairs += int(np.sum(s_near[w_slice] - w_near))
            g_near = np.minimum(np.searchsorted(nu0, g_mid + g_half, side="right")
                                - np.searchsorted(nu0, g_mid - g_half, side="left"),
                                w_near[g_wave])
            near_evals += int(np.sum(g_near * g_pts))
            group_pairs += int(np.sum(w_near[g_wave] - g_near))
        return {"near_evals": near_evals, "slice_pairs": slice_pairs, "wave_pairs": wave_pairs,
                "group_pairs": group_pairs, "far_pairs": slice_pairs + wave_pairs + group_pairs}

    def kernel_work(self, precision: str = "mixed") -> dict:
        """The line-sum kernels' own work counts for one run of this problem
        (their counting variants, lbl_enable_work_count): near-field point
        evaluations, series pairs per level, the records each level passes on
        or evaluates point by point, queued near-centre pairs and core points.
        ``count_series_work`` predicts the first three from the host; the
        tests hold the two to equality."""
        _capi.enable_work_count(True)
        try:
            self.run(precision, mus=(1.0,))
            return _capi.last_work_count()
        finally:
            _capi.enable_work_count(False)

    # ------------------------------------------------------------------ run --
    def workspace(self, precision: int):
        torch = _torch()
        if precision not in self._ws:
            nbytes = ctypes.c_size_t(0)
            st = self.lib.lbl_workspace_bytes(ctypes.byref(self.c_lines), ctypes.byref(self.c_layers),
                                              ctypes.byref(self.c_grid), precision, ctypes.byref(nbytes))
            _capi.raise_for(st, "lbl_workspace_bytes")
            self._ws[precision] = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8,
                                              device=self.device)
        return self._ws[precision]

    def alloc_outputs(self, precision: str, n_mu: int, want_tau: bool, want_layers: bool):
        torch = _torch()
        dt = torch.float64 if precision == "fp64" else torch.float32
        out = {"transmittance": torch.empty((n_mu, self.n_nu), dtype=dt, device=self.device)}
        if want_tau:
            out["optical_depth"] = torch.empty(self.n_nu, dtype=dt, device=self.device)
        if want_layers:
            out["optical_depth_layers"] = torch.empty((self.n_layers, self.n_nu), dtype=dt,
                                                      device=self.device)
        return out

    def run(self, precision: str = "fp64", mus: Sequence[float] = (1.0,), want_tau: bool = False,
            want_layers: bool = False, outputs: Optional[dict] = None, surrogate=None,
            gate: Optional[float] = None, census: Optional[dict] = None) -> dict:
        """Run the device path; returns device tensors (transmittance [n_mu, n_nu], ...).

        precision "surrogate" uses ``surrogate`` (a VoigtSurrogate; default:
        the committed network) for near-centre points that pass ``gate``
        (default SURROGATE_GATE).  If ``census`` is a dict, it receives the
        number of network and exact near-centre evaluations."""
        torch = _torch()
        if precision not in _capi.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_capi.PRECISIONS)}; got {precision!r}")
        code = _capi.PRECISIONS[precision]
        ws = self.workspace(_capi.PRECISIONS["mixed"] if precision == "surrogate" else code)
        out = outputs or self.alloc_outputs(precision, len(mus), want_tau, want_layers)
        mu_arr = (ctypes.c_double * len(mus))(*[float(m) for m in mus])
        ptr = lambda k: out[k].data_ptr() if k in out else None
        c_out = _capi.LblOutputs(len(mus), mu_arr, ptr("transmittance"), ptr("optical_depth"),
                                 ptr("optical_depth_layers"))
        err_layer = ctypes.c_int32(-1)
        err_group = ctypes.c_int32(-1)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        with torch.cuda.device(self.device):
            if precision == "surrogate":
                from .surrogate import SURROGATE_GATE, default_surrogate
                net = surrogate if surrogate is not None else default_surrogate(self.device)
                counts = torch.zeros(2, dtype=torch.int64, device=self.device)
                c_sur = _capi.LblSurrogate(net.params.data_ptr(), net.wbf.data_ptr(),
                                           float(SURROGATE_GATE if gate is None else gate),
                                           counts.data_ptr() if census is not None else None,
                                           counts.data_ptr() + 8 if census is not None else None)
                st = self.lib.lbl_optical_depth_surrogate(
                    ctypes.byref(self.c_lines), ctypes.byref(self.c_groups),
                    ctypes.byref(self.c_layers), ctypes.byref(self.c_grid), ctypes.byref(c_out),
                    ctypes.byref(c_sur), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                    ctypes.byref(err_layer), ctypes.byref(err_group), ctypes.c_void_p(stream))
            else:
                st = self.lib.lbl_optical_depth(
                    ctypes.byref(self.c_lines), ctypes.byref(self.c_groups),
                    ctypes.byref(self.c_layers), ctypes.byref(self.c_grid), code,
                    ctypes.byref(c_out), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                    ctypes.byref(err_layer), ctypes.byref(err_group), ctypes.c_void_p(stream))
        if st == _capi.LBL_ESTRENGTH:
            gas = self.groups[err_group.value].gas
            raise FloatingPointError(
                f"Layer {err_layer.value}, gas '{gas}': non-finite or negative line strength")
        if st == _capi.LBL_ENONFINITE:
            raise FloatingPointError(
                f"Layer {err_layer.value}: non-finite optical depth after radiation term")
        _capi.raise_for(st, "lbl_optical_depth")
        if precision == "surrogate" and census is not None:
            c = counts.cpu().tolist()
            census.update(network=int(c[0]), exact_near=int(c[1]))
        return out


def voigt_eval(x: np.ndarray, y: np.ndarray, precision: str = "fp64", device=None) -> np.ndarray:
    """Device K(x, y) = Re w(x + iy) with the line sum's region dispatch."""
    torch = _torch()
    lib = _capi.load()
    dt = np.float64 if precision == "fp64" else np.float32
    dev = torch.device(device if device is not None else f"cuda:{torch.cuda.current_device()}")
    xt = torch.as_tensor(np.ascontiguousarray(x, dt)).to(dev)
    yt = torch.as_tensor(np.ascontiguousarray(y, dt)).to(dev)
    kt = torch.empty_like(xt)
    stream = torch.cuda.current_stream(dev).cuda_stream
    st = lib.lbl_voigt_eval(xt.numel(), xt.data_ptr(), yt.data_ptr(), kt.data_ptr(),
                            _capi.PRECISIONS[precision], ctypes.c_void_p(stream))
    _capi.raise_for(st, "lbl_voigt_eval")
    return kt.cpu().numpy()

"""FP64 CPU analytic adjoint of the oracle's transmittance.

TEST INFRASTRUCTURE ONLY (checker of the HIP adjoint kernel).

The reference has no gradients.  This differentiates the reference's
forward (pylblrtm/forward.py:236-240 partition-sum ratio, :257-292 line
scaling, :51-62 Voigt, :312 radiation term, :151 transmittance) with respect
to layer temperatures T_k and layer mixing ratios v_{k,g}, the column of gas
g in layer k being v_{k,g} * air_k.  Voigt derivatives use both parts of
w = K + iL from scipy.special.wofz:
    dK/dx = -2 (x K - y L),   dK/dy = 2 (x L + y K) - 2/sqrt(pi).
The partition sum is np.interp (piecewise linear), so d ln q/dT uses the
slope of the segment holding T.  At v_{k,g} = 0 (where the forward skips the
gas) the one-sided derivative is the gas's layer optical depth with column
air_k and no self-pressure.  Pinned by central finite differences of the
oracle's own forward, and at v = 0 by one-sided ones
(tests/test_adjoint_oracle.py).
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import wofz

from . import lbl_oracle as O


def _interp_slope(T, xp, fp):
    if T < xp[0] or T > xp[-1] or len(xp) < 2:
        return 0.0
    j = min(int(np.searchsorted(xp, T, side="right")) - 1, len(xp) - 2)
    return (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j])


def forward(nu, lines, tips, mm, p_atm, t_lay, vmr, air, cutoff, tile_width, mus):
    """Transmittance [n_mu, n_nu] with col = vmr * air (FP64)."""
    col = {g: np.asarray(v) * air for g, v in vmr.items()}
    od = O.optical_depth(nu, lines, tips, mm, p_atm, t_lay, vmr, col, cutoff, tile_width)
    tau = od.sum(0)
    return np.stack([np.exp(-tau / m) for m in mus])


def weight(Tm, G, mus):
    """W(nu) = dL/d tau(nu) = -sum_m G_m T_m / mu_m for L = sum G * T."""
    return -np.sum(np.asarray(G) * np.asarray(Tm) / np.asarray(mus, np.float64)[:, None], axis=0)


def vjp(nu, lines, tips, mm, p_atm, t_lay, vmr, air, cutoff, tile_width, mus, G):
    """(dL/dT_k [n_layers], {gas: dL/dv_{k,gas} [n_layers]}) for L = sum G * T."""
    Tm = forward(nu, lines, tips, mm, p_atm, t_lay, vmr, air, cutoff, tile_width, mus)
    return vjp_layers(nu, lines, tips, mm, p_atm, t_lay, vmr, air, cutoff, tile_width,
                      weight(Tm, G, mus), range(len(p_atm)))


def vjp_layers(nu, lines, tips, mm, p_atm, t_lay, vmr, air, cutoff, tile_width, W, layers,
               block: int = 256):
    """The layers ``layers`` of vjp for a given W(nu) (``weight``); the other
    entries stay 0.  Layers are independent given W, so a caller may split
    them over processes.  Lines are taken ``block`` at a time per tile."""
    gases = list(vmr)
    col = {g: np.asarray(vmr[g]) * air for g in gases}
    nl = len(p_atm)
    gT = np.zeros(nl)
    gV = {g: np.zeros(nl) for g in gases}
    ts = O.tile_samples(nu, tile_width)
    bounds = O.tile_bounds(nu, ts)
    groups = O.prepare_groups(nu, lines, tips, mm, p_atm, t_lay, gases, cutoff)
    sl, sp = O.SQRT_LN2, math.sqrt(math.pi)

    def in_tile(c, t0, t1):
        sel = np.flatnonzero((c >= nu[t0] - cutoff) & (c <= nu[t1 - 1] + cutoff))
        return [sel[b0:b0 + block] for b0 in range(0, sel.size, block)]

    for k in layers:
        T, P = float(t_lay[k]), float(p_atm[k])
        R = O.radiation_term(nu, T)
        u_grid = O.C2 * nu / T
        dlnR_grid = -(u_grid / T) / np.sinh(u_grid)
        lay = np.zeros_like(nu)
        for g in groups:
            gas = g["gas"]
            v = float(vmr[gas][k])
            idx = g["idx"]
            if v == 0.0 and float(air[k]) > 0.0:
                # forward.py:253-255 skips the gas here, but tau is continuous in
                # v (strength ~ col = v * air, forward.py:270-278) and every
                # other v-dependence multiplies that strength: the one-sided
                # derivative is the optical depth with col = air, P_self = 0
                c, S, gl, ad = O.line_params(lines, idx, T, P, 0.0, float(air[k]),
                                             g["q_ratio"][k], g["molar_mass"])
                A1, s1, y1 = S * O.VOIGT_K / ad, sl / ad, sl * gl / ad
                for t0, t1 in bounds:
                    for m in in_tile(c, t0, t1):
                        x = (nu[t0:t1, None] - c[m]) * s1[m]
                        K = wofz(x + 1j * y1[m]).real
                        gV[gas][k] += np.sum((W[t0:t1] * R[t0:t1])[:, None] * A1[m] * K)
                continue
            if float(col[gas][k]) == 0.0 or v == 0.0:
                continue
            c, S, gl, ad = O.line_params(lines, idx, T, P, v, float(col[gas][k]),
                                         g["q_ratio"][k], g["molar_mass"])
            nu0 = lines["nu0"][idx]
            delta = lines["shift"][idx]
            n_air = lines["n_air"][idx]
            tn, qn = tips[(g["mol"], g["iso"])]
            dlnq = -_interp_slope(T, tn, qn) / O.partition_sum(T, tn, qn)
            a = O.C2 * nu0 / T
            u = O.C2 * c / T
            A = S * O.VOIGT_K / ad
            s = sl / ad
            y = sl * gl / ad
            at = dlnq + O.C2 * lines["Elow"][idx] / T**2 - (a / T) / np.expm1(a) + (u / T) / np.sinh(u) - 0.5 / T
            dyt = -y * (n_air + 0.5) / T
            av = 1.0 / v + delta * P * (1.0 / c + (O.C2 / T) / np.sinh(u))
            bxv = s * delta * P
            gscale = (O.T_REF / T) ** n_air
            dyv = sl * gscale * P * (lines["g_self"][idx] - lines["g_air"][idx]) / ad
            for t0, t1 in bounds:
                wr = (W[t0:t1] * R[t0:t1])[:, None]
                for m in in_tile(c, t0, t1):
                    x = (nu[t0:t1, None] - c[m]) * s[m]
                    w = wofz(x + 1j * y[m])
                    K, L = w.real, w.imag
                    Kx = -2.0 * (x * K - y[m] * L)
                    Ky = 2.0 * (x * L + y[m] * K) - 2.0 / sp
                    AK = A[m] * K
                    lay[t0:t1] += AK.sum(1)
                    gT[k] += np.sum(wr * (at[m] * AK + A[m] * Kx * (-0.5 / T) * x + A[m] * Ky * dyt[m]))
                    gV[gas][k] += np.sum(wr * (av[m] * AK + A[m] * Kx * bxv[m] + A[m] * Ky * dyv[m]))
        gT[k] += np.sum(W * R * dlnR_grid * lay)
    return gT, gV

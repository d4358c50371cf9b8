"""The FP64 analytic adjoint oracle against central finite differences of the
oracle's own forward (CPU only)."""

import numpy as np
import pytest

from oracle import adjoint_oracle as A
from oracle import lbl_oracle as O
from pylblrtm_amd import synth


def translucent_case(seed=7):
    """Two gases, 4 layers, line strengths scaled so 0.05 < T < 0.95."""
    rng = np.random.default_rng(seed)
    parts = [synth.random_lines(rng, 30, 699.0, 703.0, 2), synth.random_lines(rng, 30, 699.0, 703.0, 1)]
    lines = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    lines["Sref"] = lines["Sref"] * 1e-4
    tips = {(2, 1): synth.tips_table(), (1, 1): synth.tips_table(174.6)}
    mm = {(2, 1): 43.99, (1, 1): 18.01}
    p = np.array([10.0, 100.0, 300.0, 600.0, 1000.0])
    lay_p = np.sqrt(p[:-1] * p[1:]) / 1013.25
    lay_t = np.array([220.3, 235.7, 258.4, 281.1])   # away from the 1 K TIPS nodes
    air = (6.02214086e23 / (9.80665 * 0.0289644)) * (np.diff(p) * 100.0) / 1e4
    vmr = {"CO2": np.full(4, 4.15e-4), "H2O": np.array([5e-6, 2e-4, 2e-3, 8e-3])}
    nu = O.grid(700.0, 702.0, 0.004)
    return nu, lines, tips, mm, lay_p, lay_t, vmr, air


def test_adjoint_oracle_matches_finite_differences():
    nu, lines, tips, mm, p, t, vmr, air = translucent_case()
    mus = (1.0, 0.5)
    args = (nu, lines, tips, mm, p, t, vmr, air, 25.0, 0.5, mus)
    T0 = A.forward(*args)
    assert 0.01 < T0.min() and T0.max() < 0.999
    rng = np.random.default_rng(3)
    G = rng.normal(size=T0.shape)
    gT, gV = A.vjp(*args, G)

    def loss(tt, vv):
        return float(np.sum(G * A.forward(nu, lines, tips, mm, p, tt, vv, air, 25.0, 0.5, mus)))

    for k in range(4):
        h = 1e-4
        tp, tm = t.copy(), t.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss(tp, vmr) - loss(tm, vmr)) / (2 * h)
        assert gT[k] == pytest.approx(fd, rel=2e-6, abs=1e-12), (k, gT[k], fd)
        for gas in vmr:
            hv = vmr[gas][k] * 1e-3   # smaller steps move centres by ~ulp(nu)
            vp = {g: v.copy() for g, v in vmr.items()}
            vm = {g: v.copy() for g, v in vmr.items()}
            vp[gas][k] += hv
            vm[gas][k] -= hv
            fdv = (loss(t, vp) - loss(t, vm)) / (2 * hv)
            assert gV[gas][k] == pytest.approx(fdv, rel=2e-6, abs=1e-12), (gas, k, gV[gas][k], fdv)


def test_adjoint_oracle_at_zero_mixing_ratio():
    """At vmr = 0 the forward skips the gas (forward.py:253-255) but is
    continuous there; the oracle's gradient is the one-sided derivative.
    Pinned by one-sided differences at h and h/2 (Richardson: 2 D(h/2) - D(h)
    removes the O(h) term), and by continuity with the gradient at a tiny
    positive vmr."""
    nu, lines, tips, mm, p, t, vmr, air = translucent_case()
    mus = (1.0,)
    vmr0 = {g: v.copy() for g, v in vmr.items()}
    vmr0["H2O"][1] = 0.0
    vmr0["CO2"][3] = 0.0
    args = (nu, lines, tips, mm, p, t, vmr0, air, 25.0, 0.5, mus)
    T0 = A.forward(*args)
    G = np.random.default_rng(9).normal(size=T0.shape)
    gT, gV = A.vjp(*args, G)

    def loss(vv):
        return float(np.sum(G * A.forward(nu, lines, tips, mm, p, t, vv, air, 25.0, 0.5, mus)))

    L0 = loss(vmr0)
    for gas, k, h in (("H2O", 1, 2e-6), ("CO2", 3, 2e-6)):
        assert gV[gas][k] != 0.0
        d = []
        for hh in (h, h / 2):
            vp = {g: v.copy() for g, v in vmr0.items()}
            vp[gas][k] = hh
            d.append((loss(vp) - L0) / hh)
        fd = 2 * d[1] - d[0]
        assert gV[gas][k] == pytest.approx(fd, rel=1e-5), (gas, k, gV[gas][k], fd, d)
        # continuity: the analytic gradient at vmr = 1e-12 (the general path)
        vt = {g: v.copy() for g, v in vmr0.items()}
        vt[gas][k] = 1e-12
        _, gVt = A.vjp(nu, lines, tips, mm, p, t, vt, air, 25.0, 0.5, mus, G)
        assert gV[gas][k] == pytest.approx(gVt[gas][k], rel=1e-7)

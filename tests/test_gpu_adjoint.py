"""HIP adjoint (lbl_transmittance_vjp) against the FP64 analytic adjoint
oracle, central finite differences of the FP64 GPU forward, and through
torch.autograd.  Needs a ROCm GPU."""

import sys

import numpy as np
import pytest

from oracle import adjoint_oracle as A
from pylblrtm_amd import SpectralData, synth
from pylblrtm_amd.atmosphere import _Layers
from pylblrtm_amd.device import OpticalDepthProblem
from test_adjoint_oracle import translucent_case

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-5      # FP32 adjoint vs FP64 references, relative to the gradient norm


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(t_override=None):
    nu, lines, tips, mm, p, t, vmr, air = translucent_case()
    if t_override is not None:
        t = t_override
    sd = SpectralData(lines=lines, tips=tips, molar_masses=mm)
    lay = _Layers(p, t, dict(vmr), {g: v * air for g, v in vmr.items()}, air)
    prob = OpticalDepthProblem(sd, lay, 700.0, 702.0, 0.004, 25.0, 0.5)
    return prob, (nu, lines, tips, mm, p, t, vmr, air)


def _vmr_matrix(prob, vmr):
    return np.stack([vmr[g] for g in prob.gases])


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_vjp_matches_fp64_analytic_adjoint():
    say("case: vjp vs oracle adjoint, translucent 2-gas 4-layer")
    import torch
    prob, (nu, lines, tips, mm, p, t, vmr, air) = _problem()
    mus = (1.0, 0.5)
    np.testing.assert_allclose(prob.air_cm2, air, rtol=1e-14)
    fw = prob.run("mixed", mus=mus, want_layers=True)
    T = fw["transmittance"]
    rng = np.random.default_rng(3)
    G = rng.normal(size=tuple(T.shape))
    gt, gv = prob.vjp(T, fw["optical_depth_layers"],
                      torch.as_tensor(G, dtype=torch.float32, device=T.device), mus)
    rt, rv = A.vjp(nu, lines, tips, mm, p, t, vmr, air, 25.0, 0.5, mus, G)
    ref_v = np.stack([rv[g] for g in prob.gases])
    et, ev = _rel(gt.cpu().numpy(), rt), _rel(gv.cpu().numpy(), ref_v)
    say(f"  rel error: temperature {et:.2e}, mixing ratio {ev:.2e}")
    assert et < GRAD_RTOL
    assert ev < GRAD_RTOL


def test_vjp_matches_central_differences_of_fp64_forward():
    say("case: vjp vs central differences (h = 1e-3 relative) of the FP64 GPU forward")
    import torch
    t = np.array([220.5, 235.5, 258.5, 281.5])   # +-0.1% stays inside one 1 K TIPS segment
    prob, (nu, lines, tips, mm, p, _, vmr, air) = _problem(t)
    mus = (1.0,)
    V = _vmr_matrix(prob, vmr)
    rng = np.random.default_rng(4)
    G = rng.normal(size=(1, prob.n_nu))

    def loss(tt, vv):
        prob.set_layers(tt, vv)
        T = prob.run("fp64", mus=mus)["transmittance"].cpu().numpy()
        return float(np.sum(G * T))

    fd_t = np.zeros(4)
    for k in range(4):
        h = 1e-3 * t[k]
        tp, tm = t.copy(), t.copy()
        tp[k] += h
        tm[k] -= h
        fd_t[k] = (loss(tp, V) - loss(tm, V)) / (2 * h)
    fd_v = np.zeros_like(V)
    for g in range(V.shape[0]):
        for k in range(4):
            h = 1e-3 * V[g, k]
            vp, vm = V.copy(), V.copy()
            vp[g, k] += h
            vm[g, k] -= h
            fd_v[g, k] = (loss(t, vp) - loss(t, vm)) / (2 * h)
    prob.set_layers(t, V)
    fw = prob.run("mixed", mus=mus, want_layers=True)
    T = fw["transmittance"]
    gt, gv = prob.vjp(T, fw["optical_depth_layers"],
                      torch.as_tensor(G, dtype=torch.float32, device=T.device), mus)
    et, ev = _rel(gt.cpu().numpy(), fd_t), _rel(gv.cpu().numpy(), fd_v)
    say(f"  rel error: temperature {et:.2e}, mixing ratio {ev:.2e}")
    assert et < GRAD_RTOL
    assert ev < GRAD_RTOL


def test_vjp_at_zero_mixing_ratio():
    """A retrieval that starts a gas from nothing: at vmr = 0 the forward
    skips the gas in that layer (forward.py:253-255) but is continuous there
    (strength ~ vmr * air, forward.py:270-278), so dL/dvmr is the one-sided
    derivative, not 0.  Against the FP64 adjoint oracle (pinned by one-sided
    differences on the CPU) and one-sided differences of the FP64 GPU
    forward (Richardson over h and h/2)."""
    say("case: vjp at vmr = 0, translucent 2-gas 4-layer")
    import torch
    prob, (nu, lines, tips, mm, p, t, vmr, air) = _problem()
    mus = (1.0, 0.5)
    vmr0 = {g: v.copy() for g, v in vmr.items()}
    vmr0["H2O"][1] = 0.0
    vmr0["CO2"][3] = 0.0
    V0 = _vmr_matrix(prob, vmr0)
    prob.set_layers(t, V0)
    fw = prob.run("mixed", mus=mus, want_layers=True)
    T = fw["transmittance"]
    G = np.random.default_rng(21).normal(size=tuple(T.shape))
    gt, gv = prob.vjp(T, fw["optical_depth_layers"],
                      torch.as_tensor(G, dtype=torch.float32, device=T.device), mus)
    gv = gv.cpu().numpy()
    rt, rv = A.vjp(nu, lines, tips, mm, p, t, vmr0, air, 25.0, 0.5, mus, G)
    ref_v = np.stack([rv[g] for g in prob.gases])
    et, ev = _rel(gt.cpu().numpy(), rt), _rel(gv, ref_v)
    say(f"  rel error vs oracle: temperature {et:.2e}, mixing ratio {ev:.2e}")
    assert et < GRAD_RTOL and ev < GRAD_RTOL

    def loss(vv):
        prob.set_layers(t, vv)
        return float(np.sum(G * prob.run("fp64", mus=mus)["transmittance"].cpu().numpy()))

    L0 = loss(V0)
    for gas, k in (("H2O", 1), ("CO2", 3)):
        g = prob.gases.index(gas)
        d = []
        for h in (2e-6, 1e-6):
            vp = V0.copy()
            vp[g, k] = h
            d.append((loss(vp) - L0) / h)
        fd = 2 * d[1] - d[0]
        rel_o = abs(gv[g, k] - ref_v[g, k]) / abs(ref_v[g, k])
        rel_f = abs(gv[g, k] - fd) / abs(fd)
        say(f"  d/dvmr[{gas}][{k}] at 0: gpu {gv[g, k]:.6e}, oracle {ref_v[g, k]:.6e} ({rel_o:.1e}), "
            f"one-sided fd {fd:.6e} ({rel_f:.1e})")
        assert gv[g, k] != 0.0
        assert rel_o < GRAD_RTOL and rel_f < 1e-4
    prob.set_layers(t, _vmr_matrix(prob, vmr))


def test_autograd_function_and_batched_profiles():
    say("case: autograd over 3 perturbed profiles")
    import torch
    from pylblrtm_amd.autograd import layer_transmittance
    prob, (nu, lines, tips, mm, p, t, vmr, air) = _problem()
    mus = (1.0,)
    V0 = _vmr_matrix(prob, vmr)
    rng = np.random.default_rng(9)
    for i in range(3):
        tk = t + rng.normal(0.0, 2.0, 4)
        vk = V0 * rng.lognormal(0.0, 0.1, V0.shape)
        tt = torch.tensor(tk, dtype=torch.float64, device=prob.device, requires_grad=True)
        vv = torch.tensor(vk, dtype=torch.float64, device=prob.device, requires_grad=True)
        T = layer_transmittance(prob, tt, vv, mus)
        w = torch.linspace(-1.0, 1.0, T.shape[1], device=T.device, dtype=torch.float32)
        (T[0] * w).sum().backward()
        Gw = w.cpu().numpy()[None, :].astype(np.float64)
        rt, rv = A.vjp(nu, lines, tips, mm, p, tk, {g: vk[j] for j, g in enumerate(prob.gases)},
                       air, 25.0, 0.5, mus, Gw)
        et = _rel(tt.grad.cpu().numpy(), rt)
        ev = _rel(vv.grad.cpu().numpy(), np.stack([rv[g] for g in prob.gases]))
        say(f"  profile {i} rel error: temperature {et:.2e}, mixing ratio {ev:.2e}")
        assert et < GRAD_RTOL, i
        assert ev < GRAD_RTOL, i


# Levels whose layer means are the translucent case's layers: temperatures
# 220.3, 235.7, 258.4, 281.1 K (a 0.1 K level step moves a layer by 0.05 K,
# inside its 1 K TIPS segment) and H2O 5e-6, 2e-4, 2e-3, 8e-3.
T_LEVELS = np.array([215.0, 225.6, 245.8, 271.0, 291.2])
H2O_LEVELS = np.array([2e-6, 8e-6, 3.92e-4, 3.608e-3, 1.2392e-2])


def _level_state(prob):
    rows = {"CO2": np.full(5, 4.15e-4), "H2O": H2O_LEVELS}
    return T_LEVELS.copy(), np.stack([rows[g] for g in prob.gases])


def test_level_gradients_chain_rule_and_central_differences():
    say("case: level gradients through the device layer means")
    import torch
    from pylblrtm_amd.autograd import level_layers, level_transmittance
    prob, _ = _problem()
    mus = (1.0,)
    tl, vl = _level_state(prob)
    dev = prob.device
    tt = torch.tensor(tl, dtype=torch.float64, device=dev, requires_grad=True)
    vv = torch.tensor(vl, dtype=torch.float64, device=dev, requires_grad=True)
    T = level_transmittance(prob, tt, vv, mus)
    rng = np.random.default_rng(11)
    G = rng.normal(size=tuple(T.shape))
    Gt = torch.as_tensor(G, dtype=torch.float32, device=dev)
    (T * Gt).sum().backward()
    g_tl, g_vl = tt.grad.cpu().numpy(), vv.grad.cpu().numpy()

    # closed form: level i feeds layers i-1 and i with weight 1/2 each
    t_lay, v_lay = level_layers(torch.as_tensor(tl), torch.as_tensor(vl))
    prob.set_layers(t_lay.numpy(), v_lay.numpy())
    fw = prob.run("mixed", mus=mus, want_layers=True)
    gt, gv = prob.vjp(fw["transmittance"], fw["optical_depth_layers"], Gt, mus)
    gt, gv = gt.cpu().numpy(), gv.cpu().numpy()
    ct = np.zeros(5)
    cv = np.zeros_like(vl)
    ct[:-1] += 0.5 * gt
    ct[1:] += 0.5 * gt
    cv[:, :-1] += 0.5 * gv
    cv[:, 1:] += 0.5 * gv
    # (two separate forward + adjoint runs: equal up to the rounding of the
    # atomic reductions)
    assert _rel(g_tl, ct) < 1e-6
    assert _rel(g_vl, cv) < 1e-6

    # central differences of the FP64 forward through the level means
    def loss(t_lev, v_lev):
        a, b = level_layers(torch.as_tensor(t_lev), torch.as_tensor(v_lev))
        prob.set_layers(a.numpy(), b.numpy())
        return prob.run("fp64", mus=mus)["transmittance"]

    fd_t, ad_t = [], []
    for i in range(5):
        tp, tm = tl.copy(), tl.copy()
        tp[i] += 0.1
        tm[i] -= 0.1
        fd_t.append(float(((loss(tp, vl) - loss(tm, vl)) * Gt.double()).sum()) / 0.2)
        ad_t.append(g_tl[i])
    gi = prob.gases.index("H2O")
    fd_v, ad_v = [], []
    for i in range(5):
        h = 1e-3 * vl[gi, i]
        vp, vm = vl.copy(), vl.copy()
        vp[gi, i] += h
        vm[gi, i] -= h
        fd_v.append(float(((loss(tl, vp) - loss(tl, vm)) * Gt.double()).sum()) / (2 * h))
        ad_v.append(g_vl[gi, i])
    et, ev = _rel(ad_t, fd_t), _rel(ad_v, fd_v)
    say(f"  level gradients vs central differences: temperature {et:.2e}, H2O {ev:.2e}")
    assert et < GRAD_RTOL
    assert ev < GRAD_RTOL


def test_batched_profiles_match_one_by_one():
    say("case: batch of 3 profiles vs one profile at a time")
    import torch
    from pylblrtm_amd.autograd import batch_layer_transmittance, layer_transmittance
    prob, (nu, lines, tips, mm, p, t, vmr, air) = _problem()
    mus = (1.0, 0.5)
    V0 = _vmr_matrix(prob, vmr)
    rng = np.random.default_rng(21)
    tb = np.stack([t + rng.normal(0.0, 2.0, 4) for _ in range(3)])
    vb = np.stack([V0 * rng.lognormal(0.0, 0.1, V0.shape) for _ in range(3)])
    dev = prob.device
    G = torch.as_tensor(rng.normal(size=(3, len(mus), prob.n_nu)), dtype=torch.float32, device=dev)
    tt = torch.tensor(tb, dtype=torch.float64, device=dev, requires_grad=True)
    vv = torch.tensor(vb, dtype=torch.float64, device=dev, requires_grad=True)
    T = batch_layer_transmittance(prob, tt, vv, mus)
    (T * G).sum().backward()
    for b in range(3):
        t1 = torch.tensor(tb[b], dtype=torch.float64, device=dev, requires_grad=True)
        v1 = torch.tensor(vb[b], dtype=torch.float64, device=dev, requires_grad=True)
        T1 = layer_transmittance(prob, t1, v1, mus)
        (T1 * G[b]).sum().backward()
        # same kernels on the same inputs; the near-centre drain (LDS float
        # atomics) and the gradient reduction (FP64 atomics) add in hardware
        # order, so agreement is to rounding, not bitwise
        np.testing.assert_allclose(T[b].detach().cpu().numpy(), T1.detach().cpu().numpy(),
                                   rtol=1e-6, atol=1e-7)
        assert _rel(tt.grad[b].cpu().numpy(), t1.grad.cpu().numpy()) < 1e-6, b
        assert _rel(vv.grad[b].cpu().numpy(), v1.grad.cpu().numpy()) < 1e-6, b


def _vjp_vs_oracle(prob, case, mus, seed, tile_width):
    import torch
    nu, lines, tips, mm, p, t, vmr, air = case
    fw = prob.run("mixed", mus=mus, want_layers=True)
    T = fw["transmittance"]
    G = np.random.default_rng(seed).normal(size=tuple(T.shape))
    gt, gv = prob.vjp(T, fw["optical_depth_layers"],
                      torch.as_tensor(G, dtype=torch.float32, device=T.device), mus)
    rt, rv = A.vjp(nu, lines, tips, mm, p, t, vmr, air, prob.cutoff, tile_width, mus, G)
    ref_v = np.stack([rv[g] for g in prob.gases])
    return _rel(gt.cpu().numpy(), rt), _rel(gv.cpu().numpy(), ref_v)


def test_vjp_several_wavefronts_per_slice():
    # 2 cm-1 tiles at 0.0005: 4001 points, two slices of four live wavefronts,
    # so the slice-level series, the wavefront-level series of each wavefront
    # and the near field's per-wavefront skips all take part
    say("case: vjp vs oracle adjoint, 700-702 cm-1 at 0.0005, 2 cm-1 tiles")
    nu, lines, tips, mm, p, t, vmr, air = translucent_case()
    nu = A.O.grid(700.0, 702.0, 0.0005)
    sd = SpectralData(lines=lines, tips=tips, molar_masses=mm)
    lay = _Layers(p, t, dict(vmr), {g: v * air for g, v in vmr.items()}, air)
    prob = OpticalDepthProblem(sd, lay, 700.0, 702.0, 0.0005, 25.0, 2.0)
    et, ev = _vjp_vs_oracle(prob, (nu, lines, tips, mm, p, t, vmr, air), (1.0, 0.5), 21, 2.0)
    say(f"  rel error: temperature {et:.2e}, mixing ratio {ev:.2e}")
    assert et < GRAD_RTOL
    assert ev < GRAD_RTOL


def dense_broad_case(n=5000):
    """5,000 broad H2O-like lines (g_air, g_self 0.4-0.7 cm-1/atm) within
    2.5 cm-1 of a 0.6 cm-1 window at 1e-4: in the surface layer about 1,500
    records per slice lie outside the slice's near range yet too near for the
    one-node form to hold over the whole slice, more than the 1,024-entry list
    of the series kernels holds, so their rescan path runs (checked below)."""
    rng = np.random.default_rng(17)
    lines = synth.random_lines(rng, n, 2497.5, 2502.5, 1)
    lines["g_air"] = rng.uniform(0.4, 0.7, n)
    lines["g_self"] = rng.uniform(0.4, 0.7, n)
    lines["Sref"] = 10.0 ** rng.uniform(-27.5, -25.5, n)
    tips = {(1, 1): synth.tips_table(174.6)}
    mm = {(1, 1): 18.01}
    p = np.array([10.0, 100.0, 300.0, 600.0, 1000.0])
    lay_p = np.sqrt(p[:-1] * p[1:]) / 1013.25
    lay_t = np.array([220.3, 235.7, 258.4, 281.1])
    air = (6.02214086e23 / (9.80665 * 0.0289644)) * (np.diff(p) * 100.0) / 1e4
    vmr = {"H2O": np.array([5e-6, 2e-4, 2e-3, 8e-3])}
    nu = A.O.grid(2499.7, 2500.3, 1e-4)
    return nu, lines, tips, mm, lay_p, lay_t, vmr, air


def _slice_exceptions(case, k):
    """Records of layer k outside each slice's near range (3 half-spans) for
    which the one-node form fails somewhere in the slice (float64 version of
    the kernels' test)."""
    nu, lines, tips, mm, p, t, vmr, air = case
    idx = np.arange(len(lines["nu0"]))
    c, _, lor, dop = A.O.line_params(lines, idx, t[k], p[k], vmr["H2O"][k], 1.0, 1.0, 18.01)
    s = np.sqrt(np.log(2.0)) / dop
    y2 = (np.sqrt(np.log(2.0)) * lor / dop) ** 2
    x_s1 = np.sqrt(np.maximum(np.maximum(625.0, np.sqrt(164062.5 * (2.0 * y2 + 1.5))) - y2, 0.0))
    step = nu[1] - nu[0]
    hs = 1024 * step
    out = []
    for i0 in range(0, len(nu), 2048):
        mid = nu[0] + (i0 + 1023.5) * step
        far = np.abs(lines["nu0"] - mid) > 3.0 * hs
        fails = np.abs(mid - c) < x_s1 / s + hs
        out.append(int(np.sum(far & fails)))
    return out


def test_slice_exception_overflow_forward_and_vjp():
    say("case: dense broad lines, slice-level exception list overflows (forward and vjp)")
    case = dense_broad_case()
    nu, lines, tips, mm, p, t, vmr, air = case
    assert min(_slice_exceptions(case, 3)) > 1024   # the rescan path is taken
    sd = SpectralData(lines=lines, tips=tips, molar_masses=mm)
    lay = _Layers(p, t, dict(vmr), {g: v * air for g, v in vmr.items()}, air)
    prob = OpticalDepthProblem(sd, lay, 2499.7, 2500.3, 1e-4, 25.0, 2.0)
    ref = A.forward(nu, lines, tips, mm, p, t, vmr, air, 25.0, 2.0, (1.0,))[0]
    T64 = prob.run("fp64", mus=(1.0,))["transmittance"][0].cpu().numpy()
    T32 = prob.run("mixed", mus=(1.0,))["transmittance"][0].double().cpu().numpy()
    d64, d32 = float(np.max(np.abs(T64 - ref))), float(np.max(np.abs(T32 - ref)))
    say(f"  forward max |dT|: fp64 {d64:.2e}, mixed {d32:.2e}")
    assert d64 < 1e-9
    assert d32 < 1e-5
    et, ev = _vjp_vs_oracle(prob, case, (1.0,), 23, 2.0)
    say(f"  vjp rel error: temperature {et:.2e}, mixing ratio {ev:.2e}")
    assert et < GRAD_RTOL
    assert ev < GRAD_RTOL

"""HIP adjoint (lbl_transmittance_vjp) against the FP64 analytic adjoint
oracle, central finite differences of the FP64 GPU forward, and through
torch.autograd.  Needs a ROCm GPU."""

import sys

import numpy as np
import pytest

from oracle import adjoint_oracle as A
from pylblrtm_amd import SpectralData
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
    lay = _Layers(p, t, dict(vmr), {g: v * air for g, v in vmr.items()})
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

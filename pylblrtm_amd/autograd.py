"""Differentiable transmittance for retrievals.

``LayerTransmittance.apply(t_layers, vmr_layers, problem, mus)`` returns the
mixed-precision GPU transmittance [n_mu, n_nu] as a function of the layer
temperatures [n_layers] and layer mixing ratios [n_gases, n_layers] (rows in
``problem.gases`` order; each layer column is vmr x the layer's air column).
Its backward is the hand-written adjoint kernel (libpylblrtm_amd,
``lbl_transmittance_vjp``), which recomputes the Voigt values instead of
storing anything of size lines x grid.

``level_transmittance`` takes level values instead (the layer means of
atmosphere.py:133-139 taken on the device, so gradients reach the levels),
and ``batch_layer_transmittance`` stacks B profiles into one autograd node.

The reference has no gradients; the forward it differentiates is
pylblrtm/forward.py:257-312 and :151.
"""

from __future__ import annotations

from typing import Sequence

import torch

from .device import OpticalDepthProblem


class LayerTransmittance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t_layers: torch.Tensor, vmr_layers: torch.Tensor,
                problem: OpticalDepthProblem, mus: Sequence[float] = (1.0,)):
        problem.set_layers(t_layers.detach(), vmr_layers.detach())
        out = problem.run("mixed", mus=tuple(mus), want_layers=True)
        T = out["transmittance"]
        ctx.problem = problem
        ctx.mus = tuple(mus)
        ctx.save_for_backward(t_layers.detach().clone(), vmr_layers.detach().clone(), T,
                              out["optical_depth_layers"])
        return T

    @staticmethod
    def backward(ctx, grad_T: torch.Tensor):
        t_layers, vmr_layers, T, od_layers = ctx.saved_tensors
        prob = ctx.problem
        prob.set_layers(t_layers, vmr_layers)
        gt, gv = prob.vjp(T, od_layers, grad_T, ctx.mus)
        return (gt.to(t_layers.dtype), gv.to(vmr_layers.dtype), None, None)


def layer_transmittance(problem: OpticalDepthProblem, t_layers: torch.Tensor,
                        vmr_layers: torch.Tensor, mus: Sequence[float] = (1.0,)) -> torch.Tensor:
    return LayerTransmittance.apply(t_layers, vmr_layers, problem, tuple(mus))


class BatchTransmittance(torch.autograd.Function):
    """B profiles in one autograd node: T[b] [n_mu, n_nu] for
    t_layers[b] [n_layers] and vmr_layers[b] [n_gases, n_layers].  Each
    profile's forward keeps its transmittance and per-layer optical depth
    (B x n_layers x n_nu floats) for the backward; the profiles run one after
    another, each launch filling the GPU on its own."""

    @staticmethod
    def forward(ctx, t_layers: torch.Tensor, vmr_layers: torch.Tensor,
                problem: OpticalDepthProblem, mus: Sequence[float] = (1.0,)):
        mus = tuple(mus)
        Ts, ods = [], []
        for b in range(t_layers.shape[0]):
            problem.set_layers(t_layers[b].detach(), vmr_layers[b].detach())
            out = problem.run("mixed", mus=mus, want_layers=True)
            Ts.append(out["transmittance"])
            ods.append(out["optical_depth_layers"])
        T = torch.stack(Ts)
        ctx.problem, ctx.mus = problem, mus
        ctx.save_for_backward(t_layers.detach().clone(), vmr_layers.detach().clone(), T,
                              torch.stack(ods))
        return T

    @staticmethod
    def backward(ctx, grad_T: torch.Tensor):
        t_layers, vmr_layers, T, ods = ctx.saved_tensors
        prob = ctx.problem
        gts, gvs = [], []
        for b in range(t_layers.shape[0]):
            prob.set_layers(t_layers[b], vmr_layers[b])
            gt, gv = prob.vjp(T[b], ods[b], grad_T[b], ctx.mus)
            gts.append(gt)
            gvs.append(gv)
        return (torch.stack(gts).to(t_layers.dtype), torch.stack(gvs).to(vmr_layers.dtype),
                None, None)


def batch_layer_transmittance(problem: OpticalDepthProblem, t_layers: torch.Tensor,
                              vmr_layers: torch.Tensor,
                              mus: Sequence[float] = (1.0,)) -> torch.Tensor:
    """T [B, n_mu, n_nu] for a batch of layer states [B, n_layers] and
    [B, n_gases, n_layers]."""
    return BatchTransmittance.apply(t_layers, vmr_layers, problem, tuple(mus))


def level_layers(t_levels: torch.Tensor, vmr_levels: torch.Tensor):
    """Layer temperature and mixing ratios from level values, on the device
    and differentiable: the arithmetic means of the bounding levels that
    AtmosphericProfile.build_layers takes (atmosphere.py:133-139).  Works on
    [n_levels] / [n_gases, n_levels] and on batches with a leading axis."""
    t_lay = 0.5 * (t_levels[..., :-1] + t_levels[..., 1:])
    v_lay = 0.5 * (vmr_levels[..., :-1] + vmr_levels[..., 1:])
    return t_lay, v_lay


def level_transmittance(problem: OpticalDepthProblem, t_levels: torch.Tensor,
                        vmr_levels: torch.Tensor, mus: Sequence[float] = (1.0,)) -> torch.Tensor:
    """Transmittance as a function of LEVEL temperatures [n_levels] (or
    [B, n_levels]) and level mixing ratios [n_gases, n_levels] (or
    [B, n_gases, n_levels]), rows in ``problem.gases`` order.  The level
    pressures are the ones ``problem`` was built from (its layer pressures
    and air columns stay fixed); gradients reach the levels through the
    layer means."""
    t_lay, v_lay = level_layers(t_levels, vmr_levels)
    if t_lay.dim() == 2:
        return batch_layer_transmittance(problem, t_lay, v_lay, mus)
    return layer_transmittance(problem, t_lay, v_lay, mus)

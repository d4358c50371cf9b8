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
    t_layers[b] [n_layers] and vmr_layers[b] [n_gases, n_layers].  The
    profiles run one after another, each launch filling the GPU on its own.

    With ``save_layers`` each profile's per-layer optical depth (B x n_layers
    x n_nu floats, 12.8 GB for 64 profiles of 50 layers on 1,000,001 points)
    is kept for the backward; without it only T (B x n_mu x n_nu) is kept and
    the backward runs each profile's forward again for its layer optical
    depth (one forward per profile more, n_layers x n_nu floats at a time)."""

    @staticmethod
    def forward(ctx, t_layers: torch.Tensor, vmr_layers: torch.Tensor,
                problem: OpticalDepthProblem, mus: Sequence[float] = (1.0,),
                save_layers: bool = True):
        mus = tuple(mus)
        Ts, ods = [], []
        for b in range(t_layers.shape[0]):
            problem.set_layers(t_layers[b].detach(), vmr_layers[b].detach())
            out = problem.run("mixed", mus=mus, want_layers=save_layers)
            Ts.append(out["transmittance"])
            if save_layers:
                ods.append(out["optical_depth_layers"])
            del out
        T = torch.stack(Ts)
        del Ts
        ctx.problem, ctx.mus = problem, mus
        saved = [t_layers.detach().clone(), vmr_layers.detach().clone(), T]
        if save_layers:
            saved.append(torch.stack(ods))
        ctx.save_for_backward(*saved)
        return T

    @staticmethod
    def backward(ctx, grad_T: torch.Tensor):
        t_layers, vmr_layers, T = ctx.saved_tensors[:3]
        ods = ctx.saved_tensors[3] if len(ctx.saved_tensors) > 3 else None
        prob = ctx.problem
        gts, gvs = [], []
        for b in range(t_layers.shape[0]):
            prob.set_layers(t_layers[b], vmr_layers[b])
            if ods is not None:
                od = ods[b]
            else:
                od = prob.run("mixed", mus=ctx.mus, want_layers=True)["optical_depth_layers"]
            gt, gv = prob.vjp(T[b], od, grad_T[b], ctx.mus)
            del od
            gts.append(gt)
            gvs.append(gv)
        return (torch.stack(gts).to(t_layers.dtype), torch.stack(gvs).to(vmr_layers.dtype),
                None, None, None)


def batch_layer_transmittance(problem: OpticalDepthProblem, t_layers: torch.Tensor,
                              vmr_layers: torch.Tensor,
                              mus: Sequence[float] = (1.0,),
                              save_layers: bool = True) -> torch.Tensor:
    """T [B, n_mu, n_nu] for a batch of layer states [B, n_layers] and
    [B, n_gases, n_layers]; ``save_layers=False`` trades one more forward
    per profile in the backward for B x n_layers x n_nu floats of memory."""
    return BatchTransmittance.apply(t_layers, vmr_layers, problem, tuple(mus), bool(save_layers))


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

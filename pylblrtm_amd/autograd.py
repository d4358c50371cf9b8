"""Differentiable transmittance for retrievals.

``LayerTransmittance.apply(t_layers, vmr_layers, problem, mus)`` returns the
mixed-precision GPU transmittance [n_mu, n_nu] as a function of the layer
temperatures [n_layers] and layer mixing ratios [n_gases, n_layers] (rows in
``problem.gases`` order; each layer column is vmr x the layer's air column).
Its backward is the hand-written adjoint kernel (libpylblrtm_amd,
``lbl_transmittance_vjp``), which recomputes the Voigt values instead of
storing anything of size lines x grid.

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

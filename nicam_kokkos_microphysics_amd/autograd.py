"""``torch.autograd`` binding of the FP64 NSW6 step and its hand-written reverse mode.

:class:`MPStep` is a ``torch.autograd.Function``: its forward runs one mp_driver step
(nsw6_mp_driver_fields) on copies of the differentiable inputs, its backward calls
nsw6_mp_driver_adj with the saved input state.  Differentiable inputs are the
tensors of ``fields.ADJ_INPUTS`` and the 8 tunable coefficients; differentiable
outputs are the tensors of ``fields.ADJ_OUTPUTS`` (freshly allocated per call).
The per-tracer surface mass fluxes have a cotangent argument in
``NSW6.adjoint(cot_sflux=...)`` but are not outputs of this Function.

The coefficients a step runs with are those of the context (``NSW6.set_params``);
their tensor in the call carries the gradient only, and must equal the context's
(a different tensor raises instead of returning a gradient at the wrong point).
Under OPT_INDIR (``params.opt_indir``) with ``unccn`` in the state, ``unccn`` is a
differentiable input too.
"""
from __future__ import annotations

import torch

from . import _native as N
from .fields import ADJ_INPUTS, ADJ_OUTPUTS


def _context_coef(nsw6, device) -> torch.Tensor:
    p = nsw6.params
    return torch.tensor([getattr(p, n) for n in N.TUNABLE_COEFFS], dtype=torch.float64, device=device)


def _diff_inputs(nsw6, state: dict) -> tuple:
    """Names of the differentiable inputs of a step on `state`."""
    if nsw6.params.opt_indir != 0.0 and state.get("unccn") is not None:
        return ADJ_INPUTS + ("unccn",)
    return ADJ_INPUTS


class MPStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, nsw6, metrics: dict, dt: float, coef: torch.Tensor, names: tuple, *inputs):
        if coef.shape != (N.NCOEF,) or not torch.equal(coef.detach().to(torch.float64),
                                                      _context_coef(nsw6, coef.device)):
            raise ValueError("coef must equal the context's tunable coefficients (set them with "
                             "NSW6.set_params first): the step runs with the context's values")
        state_in = dict(zip(names, inputs))
        saved = {k: v.detach().clone() for k, v in state_in.items()}
        work = dict(metrics)
        work.update({k: torch.empty_like(metrics[k]) for k in ADJ_OUTPUTS if k not in saved})
        work.update({k: v.clone() for k, v in saved.items()})
        nsw6.step(work, dt)
        outs = tuple(work[k] for k in ADJ_OUTPUTS)
        ctx.nsw6, ctx.metrics, ctx.dt, ctx.names = nsw6, metrics, dt, names
        ctx.save_for_backward(*saved.values())
        return outs

    @staticmethod
    def backward(ctx, *gouts):
        saved = dict(zip(ctx.names, ctx.saved_tensors))
        state = dict(ctx.metrics)
        state.update(saved)
        cot = {k: g.contiguous() for k, g in zip(ADJ_OUTPUTS, gouts) if g is not None}
        grad, gcoef = ctx.nsw6.adjoint(state, cot, ctx.dt)
        return (None, None, None, gcoef, None) + tuple(grad[k] for k in ctx.names)


def mp_step(nsw6, state: dict, dt: float, coef: torch.Tensor | None = None) -> dict:
    """One differentiable FP64 step: `state` holds every mp_driver array as device
    tensors (the differentiable ones may require grad); returns a dict of the
    ADJ_OUTPUTS tensors after the step.  `coef` (8 doubles, equal to the context's
    coefficients; default: a tensor of them) receives the coefficient gradient."""
    if coef is None:
        coef = _context_coef(nsw6, state["rhog"].device)
    names = _diff_inputs(nsw6, state)
    metrics = {k: v for k, v in state.items() if k not in names}
    outs = MPStep.apply(nsw6, metrics, dt, coef, names, *[state[k] for k in names])
    return dict(zip(ADJ_OUTPUTS, outs))

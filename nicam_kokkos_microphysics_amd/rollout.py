"""Multi-step rollout with step-level checkpoints and the reverse mode over all steps.

The forward runs `steps` mp_driver steps in place on a device-resident state (either
precision form) and keeps a checkpoint of the differentiable inputs every `every`
steps.  The backward walks the segments from the last to the first.  For each
segment it re-runs the forward from the segment's checkpoint, keeping every step's
input.  The forward is deterministic (no atomics in any value path), so these inputs
equal the original ones bit for bit.  It then runs the FP64 reverse mode
(nsw6_mp_driver_adj) step by step from the segment's last step back.

The reverse mode runs `chunk_regions` regions at a time: the step is separable by
region, so a chunk's reverse mode needs records for that chunk only.  The gradient
is that of the FP64 step taken along the trajectory the forward produced (the mixed
form's trajectory when the forward is mixed).

Rollout loss (the multi-step form of the single-step adjoint check):
    L = sum_t sum_{l,ij} |precip_t|^2 + sum_{l,k,ij} (rhoge_T - rhoge_0) / (T dt)
with precip_t the surface precipitation (rain and solid rows) of step t and T the
number of steps.
"""
from __future__ import annotations

import torch

from . import _native as N
from .driver import NSW6
from .fields import ADJ_INPUTS, ADJ_OUTPUTS

# differentiable inputs carried from one step to the next (the rest, vx/vy/vz, are
# inputs of every step that the step does not change)
CARRIED = tuple(k for k in ADJ_INPUTS if k in ADJ_OUTPUTS)
CONSTANT = tuple(k for k in ADJ_INPUTS if k not in ADJ_OUTPUTS)


def _regions(d: dict, r0: int, r1: int) -> dict:
    """Views of regions [r0, r1) of every array (the region axis is the slowest)."""
    return {k: (None if v is None else v[r0:r1]) for k, v in d.items()}


class Rollout:
    """Rollout of one batch of regions: `fwd` runs the forward steps (its precision is
    the forward's), `adj` the FP64 reverse mode with records for `chunk_regions`
    regions of `ijdim` columns."""

    def __init__(self, gz, gzh, lall: int, ijdim: int, *, precision: int = N.NSW6_MIXED,
                 chunk_regions: int = 10, params=None):
        self.lall, self.ijdim = int(lall), int(ijdim)
        self.chunk = max(1, min(int(chunk_regions), self.lall))
        self.fwd = NSW6(gz, gzh, params=params, max_columns=self.lall * self.ijdim, precision=precision)
        self.adj = NSW6(gz, gzh, params=params, max_columns=self.chunk * self.ijdim)

    def close(self):
        self.fwd.close()
        self.adj.close()

    # -- forward ---------------------------------------------------------------------
    def forward(self, state: dict, dt: float, steps: int, every: int, *, stream=None):
        """Run `steps` steps in place on `state`; returns (checkpoints, precip list).
        checkpoints[s] holds clones of the differentiable inputs of step s for
        s = 0, every, 2 every, ...; precip[t] is step t's surface precipitation."""
        ckpt, precip = {}, []
        for t in range(steps):
            if t % every == 0:
                ckpt[t] = {k: state[k].clone() for k in ADJ_INPUTS}
            self.fwd.step(state, dt, stream=stream)
            precip.append(state["precip"].clone())
        return ckpt, precip

    def _segment_inputs(self, work: dict, ckpt_inputs: dict, dt: float, n: int, stream=None):
        """Re-run n steps from a checkpoint in the scratch state `work`; returns the n
        step inputs (the first is the checkpoint itself) and precips."""
        for k, v in ckpt_inputs.items():
            work[k].copy_(v)
        inputs, precs = [ckpt_inputs], []
        for j in range(n):
            self.fwd.step(work, dt, stream=stream)
            precs.append(work["precip"].clone())
            if j + 1 < n:
                inputs.append({k: work[k].clone() for k in ADJ_INPUTS})
        return inputs, precs

    # -- reverse -----------------------------------------------------------------------
    def _step_adjoint(self, metrics: dict, inputs: dict, cot: dict, dt: float, stream=None):
        """The FP64 reverse mode of one step over all regions, chunk by chunk."""
        dev = inputs["rhog"].device
        grad = {k: torch.empty_like(inputs[k]) for k in ADJ_INPUTS}
        gcoef = torch.zeros(N.NCOEF, dtype=torch.float64, device=dev)
        part = torch.empty(N.NCOEF, dtype=torch.float64, device=dev)
        state = dict(metrics)
        state.update(inputs)
        for r0 in range(0, self.lall, self.chunk):
            r1 = min(self.lall, r0 + self.chunk)
            self.adj.adjoint(_regions(state, r0, r1), _regions(cot, r0, r1), dt, grad=_regions(grad, r0, r1),
                             grad_coef=part, stream=stream)
            gcoef += part
        return grad, gcoef

    def loss(self, rhoge0, state_final, precip, dt):
        T = len(precip)
        s = sum(float((p * p).sum()) for p in precip)
        return s + float(((state_final["rhoge"] - rhoge0) / (T * dt)).sum())

    def backward(self, state: dict, ckpt: dict, precip: list, dt: float, steps: int, every: int, *,
                 stream=None):
        """Gradient of the rollout loss with respect to the initial differentiable
        inputs and the 8 coefficients.  `state` is the state dict forward() ran on: its
        metric arrays are read and its other arrays serve as scratch for the
        recomputation (their contents are overwritten).  `ckpt` and `precip` come from
        forward()."""
        T = steps
        dev = state["rhog"].device
        metrics = {k: v for k, v in state.items() if k not in ADJ_INPUTS}
        cot_state = {k: torch.zeros_like(state[k]) for k in CARRIED}
        cot_state["rhoge"].fill_(1.0 / (T * dt))            # d L / d rhoge_T
        cot_const = {k: torch.zeros_like(state[k]) for k in CONSTANT}
        gcoef = torch.zeros(N.NCOEF, dtype=torch.float64, device=dev)
        starts = sorted(ckpt)
        work = state
        for si in reversed(range(len(starts))):
            s0 = starts[si]
            n = (starts[si + 1] if si + 1 < len(starts) else T) - s0
            inputs, precs = self._segment_inputs(work, ckpt[s0], dt, n, stream=stream)
            for j in reversed(range(n)):
                t = s0 + j
                if not torch.equal(precs[j], precip[t]):
                    raise RuntimeError(f"rollout recomputation differs from the forward at step {t}")
                cot = dict(cot_state)
                cot["precip"] = 2.0 * precs[j]
                grad, gc = self._step_adjoint(metrics, inputs[j], cot, dt, stream=stream)
                for k in CARRIED:
                    cot_state[k] = grad[k]
                for k in CONSTANT:
                    cot_const[k] += grad[k]
                gcoef += gc
                del grad
            del inputs, precs
        out = dict(cot_state)
        out.update(cot_const)
        out["rhoge"] = out["rhoge"] - 1.0 / (T * dt)          # explicit -rhoge_0 term
        return out, gcoef

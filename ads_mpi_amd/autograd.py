"""Differentiable heat rollouts: torch.autograd.Function over the C ABI.

    solver = HeatSolver(n, p)                       # one GPU context
    cT = heat_rollout(c0, kappa, solver, dt, steps, checkpoint_every=10)
    loss = 0.5 * ((cT - c_obs) ** 2).sum()
    loss.backward()                                 # c0.grad, kappa.grad

Forward runs ``steps`` explicit heat steps (ads_step_d) and keeps the state
every ``checkpoint_every`` steps.  Backward walks the segments last to first:
it recomputes the segment's states from its checkpoint, then applies the
hand-written adjoint step (ads_step_adjoint_d: transposed banded sweeps, the
assembly applied to mu, and the per-element diffusivity gradient).

All device work is ordered on torch's current stream (the context is bound to
it with ads_set_stream), so no host synchronisation happens inside.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib, adss

__all__ = ["HeatSolver", "HeatRollout", "heat_rollout"]


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class HeatSolver:
    """One ADS context (ads_init) used through its device-buffer entry points."""

    def __init__(self, n, p):
        self.ads, self.data = adss.initialize(n, p)
        self.h = self.data.handle
        self.L = _lib.load()
        self.D = self.data.D
        self.E = int(self.ads.nelem.prod())
        self.dofs = self.data.dofs

    def _bind(self):
        self.L.ads_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def step(self, c, kappa, dt, out=None, prec=_lib.ADS_PREC_FP64, refine=2):
        self._bind()
        out = torch.empty_like(c) if out is None else out
        _lib.check(self.L.ads_step_d(self.h, _vp(c), _vp(out), _vp(kappa), float(dt), int(prec), int(refine), None),
                   self.h)
        return out

    def adjoint(self, c_n, lam_next, kappa, dt, dkappa=None, out=None, prec=_lib.ADS_PREC_FP64, refine=2):
        self._bind()
        out = torch.empty_like(lam_next) if out is None else out
        _lib.check(self.L.ads_step_adjoint_d(self.h, _vp(c_n), _vp(lam_next), _vp(kappa), float(dt), int(prec),
                                             int(refine), _vp(out), _vp(dkappa)), self.h)
        return out

    def close(self):
        adss.Cleanup(self.ads, self.data)


class HeatRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c0, kappa, solver, dt, steps, checkpoint_every, prec, refine, adjoint_prec, adjoint_refine):
        k = max(1, int(checkpoint_every))
        c0d = c0.detach().contiguous()
        kd = kappa.detach().contiguous() if kappa is not None else None
        ckpts = []
        c = c0d
        a = torch.empty_like(c0d)
        b = torch.empty_like(c0d)
        for n in range(int(steps)):
            if n % k == 0:
                ckpts.append(c.clone())
            nxt = a if (c is not a) else b
            solver.step(c, kd, dt, out=nxt, prec=prec, refine=refine)
            c = nxt
        ctx.solver = solver
        ctx.prec, ctx.refine = int(prec), int(refine)
        ctx.adjoint_prec, ctx.adjoint_refine = int(adjoint_prec), int(adjoint_refine)
        ctx.dt = float(dt)
        ctx.steps = int(steps)
        ctx.k = k
        ctx.ckpts = ckpts
        ctx.has_kappa = kappa is not None
        ctx.save_for_backward(kd if kd is not None else torch.empty(0, device=c0.device))
        return c.clone()

    @staticmethod
    def backward(ctx, grad_out):
        solver, dt, steps, k = ctx.solver, ctx.dt, ctx.steps, ctx.k
        (kd,) = ctx.saved_tensors
        kd = kd if ctx.has_kappa else None
        lam = grad_out.detach().contiguous().clone()
        dk = torch.zeros(solver.E, dtype=lam.dtype, device=lam.device) if ctx.has_kappa else None
        nseg = len(ctx.ckpts)
        for s in range(nseg - 1, -1, -1):
            n0 = s * k
            n1 = min(n0 + k, steps)
            states = [ctx.ckpts[s]]
            for n in range(n0, n1 - 1):
                states.append(solver.step(states[-1], kd, dt, prec=ctx.prec, refine=ctx.refine))
            for j in range(n1 - n0 - 1, -1, -1):
                lam = solver.adjoint(states[j], lam, kd, dt, dkappa=dk, out=lam, prec=ctx.adjoint_prec,
                                     refine=ctx.adjoint_refine)
        ctx.ckpts = None
        return lam, dk, None, None, None, None, None, None, None, None


def heat_rollout(c0, kappa, solver, dt, steps, checkpoint_every=10, prec=_lib.ADS_PREC_FP64, refine=2,
                 adjoint_prec=None, adjoint_refine=None):
    """c_T after ``steps`` explicit heat steps, differentiable in c0 and kappa.

    prec/refine select the forward form (ads_step_d); adjoint_prec and
    adjoint_refine the adjoint form (ads_step_adjoint_d), by default the same.
    The backward pass recomputes each segment with the forward form."""
    ap = prec if adjoint_prec is None else adjoint_prec
    ar = refine if adjoint_refine is None else adjoint_refine
    return HeatRollout.apply(c0, kappa, solver, dt, steps, checkpoint_every, prec, refine, ap, ar)

"""torch.autograd bindings of the primal step and of a checkpointed rollout.

  Step.apply(u, source, ctx, dt) -> (u_next, J)
      one time step (the primal Function, density.py:97-105); J is the
      step's objective (FP64, shape (1,)).  backward runs the hand-written
      adjoint step kernels (the reference obtains them from
      primal.map.grad(), apps/adjoint.py:101).

  rollout(ctx, u0, dts, source=None, schedule='revolve', snapshots=..)
      nSteps steps as one autograd node: forward keeps only u0; backward is
      a reverse sweep under the chosen checkpoint schedule.  Returns
      (u_N, J) with J = sum_n J_n in FP64.

The source argument may be None (no source term, no source gradient).
"""
import torch

from . import adjoint as _adjoint


class Step(torch.autograd.Function):

    @staticmethod
    def forward(fctx, u, source, ctx, dt):
        out = u.detach().clone()
        obj = torch.zeros(1, dtype=torch.float64, device=u.device)
        ctx.primal_step(out, dt, source=None if source is None else source.detach(), obj_out=obj)
        fctx.save_for_backward(u.detach(), None if source is None else source.detach())
        fctx.ctx, fctx.dt = ctx, dt
        fctx.has_source = source is not None
        return out, obj

    @staticmethod
    def backward(fctx, g_out, g_obj):
        u, source = fctx.saved_tensors
        ctx = fctx.ctx
        ua = ctx.state(0.) if g_out is None else g_out.to(ctx.tdtype).contiguous().clone()
        obja = 0. if g_obj is None else float(g_obj.reshape(-1)[0])
        sg = ctx.state(0.) if (fctx.has_source and fctx.needs_input_grad[1]) else None
        ctx.adjoint_step(u, fctx.dt, ua, obja=obja, source=source, source_grad=sg)
        return ua, sg, None, None


class Rollout(torch.autograd.Function):

    @staticmethod
    def forward(fctx, u0, source, ctx, dts, schedule, interval, snapshots, tapes):
        u = u0.detach().clone()
        src = None if source is None else source.detach()
        J = torch.zeros(1, dtype=torch.float64, device=u0.device)
        obj = torch.zeros(1, dtype=torch.float64, device=u0.device)
        for dt in dts:
            ctx.primal_step(u, float(dt), source=src, obj_out=obj)
            J += obj
        fctx.save_for_backward(u0.detach(), src)
        fctx.ctx, fctx.dts = ctx, [float(d) for d in dts]
        fctx.opts = (schedule, interval, snapshots, tapes)
        fctx.has_source = source is not None
        return u, J

    @staticmethod
    def backward(fctx, g_u, g_J):
        u0, src = fctx.saved_tensors
        ctx = fctx.ctx
        schedule, interval, snapshots, tapes = fctx.opts
        ua = ctx.state(0.) if g_u is None else g_u.to(ctx.tdtype).contiguous().clone()
        obja = 0. if g_J is None else float(g_J.reshape(-1)[0])
        sg = ctx.state(0.) if (fctx.has_source and fctx.needs_input_grad[1]) else None
        _adjoint.sweep(ctx, u0, fctx.dts, schedule=schedule, interval=interval, snapshots=snapshots, obja=obja,
                       ua=ua, source=src, source_grad=sg, tapes=tapes)
        return ua, sg, None, None, None, None, None, None


def step(ctx, u, dt, source=None):
    return Step.apply(u, source, ctx, float(dt))


def rollout(ctx, u0, dts, source=None, schedule='revolve', interval=None, snapshots=None, tapes=True):
    """(u_N, J) over len(dts) steps; differentiable in u0 and source."""
    if schedule == 'revolve' and snapshots is None:
        snapshots = max(1, int(len(dts) ** 0.5))
    if schedule == 'uniform' and interval is None:
        interval = max(1, int(len(dts) ** 0.5))
    return Rollout.apply(u0, source, ctx, list(dts), schedule, interval, snapshots, tapes)

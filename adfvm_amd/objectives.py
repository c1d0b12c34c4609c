"""Objective functionals evaluated inside the step, in FP64.

adFVM's case files supply `objective(fields, solver)`, which the primal
Function evaluates on the first intermediate RK state (adFVM/density.py:
354-357, :412-413) and whose reverse adpy derives (apps/adjoint.py:101).
Here an objective is one of:

  kinetic_energy()   library kernel: sum_c 0.5*rho_c |U_c|^2 V_c
  drag(patch)        library kernel: sum_f (p_g n_x - mu(T_g)(U_g,x - U_o,x)/delta) A
                     over a patch's faces (templates/cylinder_test.py:9-19),
                     g the ghost, o the owner
  a callable         objective(fields, solver) -> scalar torch tensor, with
                     fields = (U (nCells,3), T (nCells,1), p (nCells,1)),
                     float64 torch tensors on the device holding the
                     primitive fields of the objective stage, ghost rows
                     included, and `solver` the RCF (mesh arrays on the
                     device: solver.mesh.tensor(name)).  The library
                     computes the fields and chains their cotangent to the
                     state (adfvm_primitive_ghosts / _vjp); torch autograd
                     differentiates only the objective itself.
"""
from . import _lib


class Objective(object):
    code = _lib.OBJ_NONE
    patch = None

    def __repr__(self):
        return '%s(%s)' % (self.__class__.__name__, self.patch or '')


class kinetic_energy(Objective):
    code = _lib.OBJ_KINETIC_ENERGY


class drag(Objective):
    code = _lib.OBJ_DRAG

    def __init__(self, patch):
        self.patch = patch


class PythonObjective(Objective):
    """A case file's objective(fields, solver) as a torch function of the
    primitive fields (cells and ghosts) of the objective stage."""
    code = _lib.OBJ_NONE

    def __init__(self, fn):
        self.fn = fn

    def __repr__(self):
        return 'PythonObjective(%s)' % getattr(self.fn, '__name__', repr(self.fn))

    def _fields(self, ctx, u, requires_grad):
        prim = ctx.primitive_ghosts(u).double()
        if requires_grad:
            prim.requires_grad_(True)
        return prim, (prim[0:3].T, prim[3:4].T, prim[4:5].T)

    def _value(self, solver, fields):
        J = self.fn(fields, solver)
        torch = solver._torch()
        if not torch.is_tensor(J):
            raise TypeError('objective(fields, solver) must return a torch tensor, got %r' % type(J))
        J = J.sum() if J.dim() else J
        if J.dtype != torch.float64:
            raise TypeError('objective must evaluate in float64 (got %s)' % J.dtype)
        return J

    def value(self, ctx, solver, u):
        """J on the state u (5, nIC): a float64 device scalar."""
        torch = solver._torch()
        with torch.no_grad():
            _, fields = self._fields(ctx, u, False)
            return self._value(solver, fields).detach().reshape(1)

    def cotangent(self, ctx, solver, u, bc_grad=None, weight=1.0):
        """(J, weight * dJ/du): J a float64 device scalar (1,), the
        gradient (5, nIC) in the context precision; weight times the ghost
        rows' dependence on boundary values is added to bc_grad
        (Context.bc_state) when given."""
        torch = solver._torch()
        with torch.enable_grad():
            prim, fields = self._fields(ctx, u, True)
            J = self._value(solver, fields)
            (pbar,) = torch.autograd.grad(J, prim, allow_unused=True)
        if pbar is None:
            pbar = torch.zeros_like(prim)
        pbar = (pbar * float(weight)).to(ctx.tdtype).contiguous()
        ubar = ctx.primitive_ghosts_vjp(u, pbar, bc_grad=bc_grad)
        return J.detach().reshape(1), ubar


def stage_cotangent(ctx, solver, u_n, dt, source, tape, weight, bc_grad=None):
    """weight * dJ_n/du at the objective stage of the step from u_n, for
    the solver's case-file objective (None when the weight is 0).  The stage
    state comes from the step's tape, or from a re-run of the step into a
    scratch tape."""
    if weight == 0:
        return None
    if ctx.n_stages == 1:
        us = u_n
    else:
        if tape is None:
            tape = solver._scratch_tape()
            ctx.primal_step(u_n.clone(), dt, source=source, tape=tape)
        us = ctx.stage_state(u_n, tape)
    return solver.objective.cotangent(ctx, solver, us, bc_grad=bc_grad, weight=weight)[1]


def resolve(obj):
    """None, a descriptor, a callable objective(fields, solver), or the
    shorthand 'kinetic_energy' / ('drag', patch)."""
    if obj is None:
        return None
    if isinstance(obj, Objective):
        return obj
    if obj == 'kinetic_energy':
        return kinetic_energy()
    if isinstance(obj, (tuple, list)) and obj and obj[0] == 'drag':
        return drag(obj[1])
    if callable(obj):
        return PythonObjective(obj)
    raise ValueError('objective must be an adfvm_amd.objectives descriptor or a callable objective(fields, '
                     'solver), got %r' % (obj,))

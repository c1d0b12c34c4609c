"""Objective functionals evaluated inside the step, in FP64.

adFVM's case files supply `objective(fields, solver)` as adpy-symbolic code
that the primal Function evaluates on the first intermediate RK state
(adFVM/density.py:354-357, :412-413) and whose reverse adpy derives.  Here an
objective is a descriptor naming one of the library's objective kernels,
each with a hand-written adjoint:

  kinetic_energy()   sum_c 0.5*rho_c |U_c|^2 V_c
  drag(patch)        sum_f (p_g n_x - mu(T_g)(U_g,x - U_o,x)/delta) A over a
                     patch's faces (templates/cylinder_test.py:9-19), g the
                     ghost, o the owner
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


def resolve(obj):
    """None, a descriptor, or the shorthand 'kinetic_energy' /
    ('drag', patch)."""
    if obj is None:
        return None
    if isinstance(obj, Objective):
        return obj
    if obj == 'kinetic_energy':
        return kinetic_energy()
    if isinstance(obj, (tuple, list)) and obj and obj[0] == 'drag':
        return drag(obj[1])
    raise ValueError('objective must be an adfvm_amd.objectives descriptor, got %r (Python objective '
                     'functions are not compiled: use kinetic_energy() or drag(patch))' % (obj,))

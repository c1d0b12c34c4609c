"""Cylinder case file whose objective is a Python function of the fields,
as the reference's case files write theirs (templates/cylinder_test.py:
9-37): objective(fields, solver) with fields = (U, T, p), here float64
torch tensors of the primitive fields on every cell, ghost rows included,
of the objective stage.  The library evaluates the fields and chains the
objective's torch gradient to the state; no objective kernel is involved.

ADFVM_CYL_OBJECTIVE selects the objective:
  drag      (default) objectiveDrag of the reference, restated in torch
  pressure  the area-mean pressure on the cylinder wall

Otherwise as templates/cylinder_test.py (64 x 32 O-grid, Mach 0.2,
Re = 100, LF at the boundaries, fixed time step, source perturbation).
"""
import os

import numpy as np

from adfvm_amd import cases
from adfvm_amd.density import RCF

nt = int(os.environ.get('ADFVM_CYL_NT', '64'))
nr = int(os.environ.get('ADFVM_CYL_NR', '32'))
ratio = float(os.environ.get('ADFVM_CYL_RATIO', '1.1'))
setup = None


def _wall(solver, name='cylinder'):
    mesh = solver.mesh
    b = mesh.boundary[name]
    f = slice(b['startFace'], b['startFace'] + b['nFaces'])
    return mesh, f


def _mu(solver, T):
    mu = solver.mu
    return mu(T) if callable(mu) else T * 0 + mu


def objectiveDrag(fields, solver):
    """sum over the wall faces of (p_g n_x - mu(T_g) (U_g,x - U_o,x)/delta) A,
    g the ghost (neighbour), o the owner (templates/cylinder_test.py:9-19)."""
    U, T, p = fields
    mesh, f = _wall(solver)
    nb, ow = mesh.tensor('neighbour')[f], mesh.tensor('owner')[f]
    U0, U0i = U[nb, 0], U[ow, 0]
    p0, T0 = p[nb, 0], T[nb, 0]
    nx = mesh.tensor('normals')[f, 0]
    mungUx = _mu(solver, T0) * (U0 - U0i) / mesh.tensor('deltas')[f, 0]
    return ((p0 * nx - mungUx) * mesh.tensor('areas')[f, 0]).sum()


def objectivePressure(fields, solver):
    """Area-mean pressure on the wall: sum p_g A / sum A."""
    U, T, p = fields
    mesh, f = _wall(solver)
    A = mesh.tensor('areas')[f, 0]
    return (p[mesh.tensor('neighbour')[f], 0] * A).sum() / A.sum()


def _make():
    global setup
    setup = cases.cylinder(nt, nr, ratio=ratio)
    return setup


caseDir = cases.ensure_case(os.environ.get('ADFVM_CASE_DIR', 'cases/cylinder%dx%d' % (nt, nr)), _make)
setup = setup or cases.cylinder(nt, nr, ratio=ratio)
objective = objectivePressure if os.environ.get('ADFVM_CYL_OBJECTIVE') == 'pressure' else objectiveDrag

primal = RCF(caseDir,
             mu=setup.mu,
             boundaryRiemannSolver='eulerLaxFriedrichs',
             objective=objective,
             fixedTimeStep=True,
             )


def perturb(fields, mesh, t):
    mid = np.array([-1.0, 0.3, 0.])
    r2 = ((mesh.cellCentres[:mesh.nInternalCells] - mid) ** 2).sum(1, keepdims=True)
    G = 1e-2 * np.exp(-r2 / 0.1)
    rho = G
    rhoU = np.zeros((mesh.nInternalCells, 3))
    rhoU[:, 0] += G.flatten() * 100
    rhoE = G * 2e5
    return rho, rhoU, rhoE


parameters = 'source'

nSteps = 20
writeInterval = 10
startTime = 0.0
dt = float(os.environ.get('ADFVM_CYL_DT', '2e-6'))

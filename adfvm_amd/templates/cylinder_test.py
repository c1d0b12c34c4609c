"""Cylinder case file: drag on the cylinder wall (templates/cylinder_test.py
of the reference: objectiveDrag, Lax-Friedrichs at the boundary, fixed time
step, Gaussian source perturbation upstream of the cylinder).

O-grid around a cylinder of diameter 1 m, far field at 20 diameters,
Mach 0.2, Re = 100, isothermal no-slip wall, CBC_UPT far field.

Environment: ADFVM_CASE_DIR (default cases/cylinder<NT>x<NR>),
ADFVM_CYL_NT / ADFVM_CYL_NR (default 64 x 32), ADFVM_CYL_RATIO (1.1).
"""
import os

import numpy as np

from adfvm_amd import cases
from adfvm_amd.density import RCF
from adfvm_amd.objectives import drag

nt = int(os.environ.get('ADFVM_CYL_NT', '64'))
nr = int(os.environ.get('ADFVM_CYL_NR', '32'))
ratio = float(os.environ.get('ADFVM_CYL_RATIO', '1.1'))
setup = None


def _make():
    global setup
    setup = cases.cylinder(nt, nr, ratio=ratio)
    return setup


caseDir = cases.ensure_case(os.environ.get('ADFVM_CASE_DIR', 'cases/cylinder%dx%d' % (nt, nr)), _make)
setup = setup or cases.cylinder(nt, nr, ratio=ratio)

primal = RCF(caseDir,
             mu=setup.mu,
             boundaryRiemannSolver='eulerLaxFriedrichs',
             objective=drag('cylinder'),
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

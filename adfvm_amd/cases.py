"""Generated cases: mesh, initial primitive fields, boundary conditions and
physical constants of the standard workloads, and a writer that lays one out
as an OpenFOAM case directory (constant/polyMesh + 0/U, T, p) for the
case-file apps.

  taylor_green(n)    periodic [0, 2pi]^3 box, Mach 0.1, rho0 = 1.2 kg/m^3,
                     T0 = 300 K, Re = 1600 (mu = rho0 U0 L / Re, L = 1)
  shock_tube(nx)     Sod problem on [x0, x0+length]: left p = 1e5 Pa,
                     T = 348.4 K; right p = 1e4 Pa, T = 278.7 K
                     (cases/shockTube/system/setFieldsDict)
  cylinder(...)      O-grid around a cylinder of diameter D: isothermal
                     no-slip wall, characteristic (CBC_UPT) far field, free
                     stream at Mach `mach`, Re = rho U D / mu
"""
import os

import numpy as np

from . import foam, meshgen
from .mesh import Mesh

GAMMA, CP = 1.4, 1004.5
R = CP - CP / GAMMA


def conservative(U, T, p):
    e = (CP / GAMMA) * T
    rho = p / (e * (GAMMA - 1))
    return rho, rho * U, rho * (e + 0.5 * (U * U).sum(-1, keepdims=True))


class CaseSetup(object):
    """md: MeshData; mesh: built Mesh; U (nIC,3), T, p (nIC,1); boundary
    {field: {patch: spec}}; mu: viscosity; extra keys in params."""

    def __init__(self, md, U, T, p, boundary, mu, mesh=None, **params):
        self.md, self.U, self.T, self.p = md, U, T, p
        self.boundary = boundary
        self.mu = mu
        self.mesh = mesh
        self.params = params

    def state(self):
        """(5, nIC) conservative rows rho, rhoU, rhoE in FP64."""
        rho, rhoU, rhoE = conservative(self.U, self.T, self.p)
        return np.concatenate([rho.T, rhoU.T, rhoE.T])

    def context(self, precision, riemann='eulerRoe', boundaryRiemann='eulerRoe', objective=None):
        """A device context for this case without a case directory (what
        RCF.compile builds from the case files).  objective: None,
        ('kinetic_energy',) or ('drag', patch)."""
        from . import BCs, _lib, config
        from .context import Context
        flux = {'eulerRoe': _lib.RIEMANN_ROE, 'eulerLaxFriedrichs': _lib.RIEMANN_LF}
        ph = dict(gamma=GAMMA, Cp=CP, Pr=0.7, mu=(_lib.MU_CONSTANT, float(self.mu), 0., 0.),
                  riemann=flux[riemann], boundaryRiemann=flux[boundaryRiemann], small=config.small_for(precision),
                  integrator=_lib.RK_SSPRK3)
        if objective is not None:
            ph['objective'] = {'kinetic_energy': _lib.OBJ_KINETIC_ENERGY, 'drag': _lib.OBJ_DRAG}[objective[0]]
            if objective[0] == 'drag':
                ph['objectivePatch'] = objective[1]
        return Context(self.mesh, BCs.validate(self.mesh, self.boundary), ph, precision)


def taylor_green(n, seed=0, rcm=True, mach=0.1, rho0=1.2, T0=300., Re=1600.):
    md = meshgen.permute_and_rcm(meshgen.periodic_box(n), seed=seed, rcm=rcm)
    mesh = Mesh.from_data(md)
    c0 = np.sqrt(GAMMA * R * T0)
    U0 = mach * c0
    x, y, z = [mesh.cellCentres[:mesh.nInternalCells, i:i + 1] for i in range(3)]
    U = np.concatenate([U0 * np.sin(x) * np.cos(y) * np.cos(z), -U0 * np.cos(x) * np.sin(y) * np.cos(z), 0 * x], 1)
    p0 = rho0 * R * T0
    p = p0 + rho0 * U0 ** 2 / 16 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
    T = T0 + 0 * x
    return CaseSetup(md, U, T, p, {}, rho0 * U0 / Re, mesh, U0=U0, rho0=rho0, T0=T0, c0=c0)


def shock_tube(nx=1000, length=1.0, origin=0.0, dy=None, split=None):
    md = meshgen.shock_tube(nx, length, dy, origin=origin)
    mesh = Mesh.from_data(md)
    x = mesh.cellCentres[:mesh.nInternalCells, 0:1]
    split = origin + 0.5 * length if split is None else split
    left = x < split
    p = np.where(left, 1e5, 1e4)
    T = np.where(left, 348.4, 278.7)
    U = np.zeros((len(x), 3))
    bnd = {f: {'sides': {'type': 'zeroGradient'}} for f in ('U', 'T', 'p')}
    return CaseSetup(md, U, T, p, bnd, 0.0, mesh)


def cylinder(n_theta=512, n_r=256, D=1.0, r_far=20.0, ratio=1.02, mach=0.2, Re=100., T_inf=300., p_inf=1e5,
             T_wall=300.):
    """r from D/2 to r_far*D; wall 'cylinder', far field 'farfield'."""
    md = meshgen.o_grid(n_theta, n_r, 0.5 * D, r_far * D, ratio, wall='cylinder', farfield='farfield')
    mesh = Mesh.from_data(md)
    nIC = mesh.nInternalCells
    c = np.sqrt(GAMMA * R * T_inf)
    Uinf = mach * c
    rho = p_inf / (R * T_inf)
    mu = rho * Uinf * D / Re
    U = np.zeros((nIC, 3))
    U[:, 0] = Uinf
    T = np.full((nIC, 1), T_inf)
    p = np.full((nIC, 1), p_inf)
    nw = mesh.boundary['cylinder']['nFaces']
    bnd = {'U': {'cylinder': {'type': 'fixedValue', 'value': np.zeros((nw, 3))},
                 'farfield': {'type': 'calculated', 'value': np.tile([Uinf, 0., 0.], (mesh.boundary['farfield']['nFaces'], 1))}},
           'T': {'cylinder': {'type': 'fixedValue', 'value': np.full((nw, 1), T_wall)},
                 'farfield': {'type': 'calculated', 'value': np.full((mesh.boundary['farfield']['nFaces'], 1), T_inf)}},
           'p': {'cylinder': {'type': 'zeroGradient'},
                 'farfield': {'type': 'CBC_UPT', 'U0': [Uinf, 0., 0.], 'T0': T_inf, 'p0': p_inf}}}
    return CaseSetup(md, U, T, p, bnd, mu, mesh, Uinf=Uinf, rho=rho, D=D)


def _spec_for_file(spec, ncomp_value):
    out = {}
    for k, v in spec.items():
        if isinstance(v, str) or k == 'type':
            out[k] = v
        else:
            a = np.asarray(v, np.float64)
            kk = 3 if (k in ('U0', 'direction') or (k == 'value' and ncomp_value == 3)) else 1
            a = a.reshape(-1, kk)
            if (a == a[0]).all():
                out[k] = 'uniform ' + ('%.17g' % a[0, 0] if kk == 1 else '(%s)' % ' '.join('%.17g' % x for x in a[0]))
            else:
                out[k] = a
    return out


def write_case(caseDir, setup):
    """Lay out constant/polyMesh and 0/U, T, p of a CaseSetup."""
    foam.write_polymesh(setup.md, caseDir)
    mesh = Mesh.from_data(setup.md, caseDir)
    bnd = {}
    for field, k in (('U', 3), ('T', 1), ('p', 1)):
        bnd[field] = {patch: _spec_for_file(spec, k) for patch, spec in setup.boundary.get(field, {}).items()}
    foam.write_fields(mesh, 0, {'U': setup.U, 'T': setup.T, 'p': setup.p}, bnd, mesh.defaultBoundary)
    return mesh


def ensure_case(caseDir, make):
    """Write the case made by make() unless caseDir already holds a mesh."""
    if not os.path.exists(os.path.join(caseDir, 'constant', 'polyMesh', 'owner')):
        write_case(caseDir, make())
    return caseDir if caseDir.endswith('/') else caseDir + '/'

"""RCF: the compressible Navier-Stokes solver of adFVM (adFVM/density.py)
on the MI355X kernels.

Same configuration keys and defaults as RCF.defaultConfig
(density.py:22-39).  `compile()` creates the device context (mesh, boundary
values, physics) and binds `self.map`, a callable with the reference
primal Function's positional contract (density.py:101-105, called at
solver.py:312-323):

    rho', rhoU', rhoE', dtc, obj = primal.map(rho, rhoU, rhoE, dt,
                                              *mesh.getTensor(), *mesh.getScalar(),
                                              *sources, *bc_values)

(mesh and boundary arguments are accepted and ignored: the context holds
them).  The host loop, `run`, keeps the state on the device instead.
"""
import numpy as np

from . import config, _lib
from . import BCs
from . import objectives
from .solver import Solver


def mu_model(mu):
    """Viscosity spec -> (model, mu0, A, Ts) for the kernels.  Accepts a
    number, ('sutherland', A, Ts), or a callable mu(T) that is constant or of
    Sutherland form A*T^1.5/(T+Ts) (RCF's default, density.py:27)."""
    if mu is None:
        return (_lib.MU_SUTHERLAND, 0., 1.4792e-06, 116.)
    if np.isscalar(mu):
        return (_lib.MU_CONSTANT, float(mu), 0., 0.)
    if isinstance(mu, tuple) and mu and mu[0] == 'sutherland':
        return (_lib.MU_SUTHERLAND, 0., float(mu[1]), float(mu[2]))
    if callable(mu):
        Ts = np.array([150., 300., 600.])
        vals = np.array([float(np.asarray(mu(T)).reshape(-1)[0]) for T in Ts])
        if np.allclose(vals, vals[0], rtol=1e-12, atol=0):
            return (_lib.MU_CONSTANT, float(vals[0]), 0., 0.)
        # fit A, S from two points of A T^1.5/(T+S) and check the third
        r = vals[0] / vals[1] * (Ts[1] ** 1.5 / Ts[0] ** 1.5)   # = (T0+S)^-1 (T1+S)
        S = (Ts[1] - r * Ts[0]) / (r - 1)
        A = vals[0] * (Ts[0] + S) / Ts[0] ** 1.5
        if np.isclose(A * Ts[2] ** 1.5 / (Ts[2] + S), vals[2], rtol=1e-9):
            return (_lib.MU_SUTHERLAND, 0., float(A), float(S))
    raise ValueError('viscosity must be constant or Sutherland-form, got %r' % (mu,))


class RCF(Solver):
    defaultConfig = dict(Solver.defaultConfig)
    defaultConfig.update({
        'Cp': 1004.5,
        'gamma': 1.4,
        'mu': None,                 # None = Sutherland 1.4792e-06*T**1.5/(T+116.)
        'Pr': 0.7,
        'CFL': 1.2,
        'stepFactor': 1.2,
        'timeIntegrator': 'SSPRK', 'nStages': 3,
        'riemannSolver': 'eulerRoe',
        'objectivePLInfo': None,
        'boundaryRiemannSolver': 'eulerRoe',
        'readConservative': False,
        'faceReconstructor': 'secondOrder',   # or 'secondOrderLimited' (clamped; not in the reference)
    })

    def __init__(self, case, **userConfig):
        super(RCF, self).__init__(case, **userConfig)
        self.Cv = self.Cp / self.gamma
        self.R = self.Cp - self.Cv
        self.names = ['rho', 'rhoU', 'rhoE']
        self.dimensions = [(1,), (3,), (1,)]
        self.boundaries = {'U': {}, 'T': {}, 'p': {}}
        self.objective = objectives.resolve(self.objective)
        for key, allowed in (('riemannSolver', ('eulerRoe', 'eulerLaxFriedrichs')),
                             ('boundaryRiemannSolver', ('eulerRoe', 'eulerLaxFriedrichs')),
                             ('timeIntegrator', ('SSPRK', 'euler')),
                             ('faceReconstructor', ('secondOrder', 'secondOrderLimited'))):
            if getattr(self, key) not in allowed:
                raise ValueError('%s=%r not supported (%s)' % (key, getattr(self, key), ', '.join(allowed)))

    # ------------------------------------------------------- conversions --
    def primitive(self, rho, rhoU, rhoE):
        """adFVM/density.py:162-171 (host arrays)."""
        U = rhoU / rho
        E = rhoE / rho
        e = E - 0.5 * (U * U).sum(-1, keepdims=True)
        p = (self.gamma - 1) * rho * e
        T = e * (1. / self.Cv)
        return U, T, p

    def conservative(self, U, T, p):
        """adFVM/density.py:173-181 (host arrays)."""
        e = self.Cv * T
        rho = p / (e * (self.gamma - 1))
        rhoE = rho * (e + 0.5 * (U * U).sum(-1, keepdims=True))
        rhoU = U * rho
        return rho, rhoU, rhoE

    # ------------------------------------------------------------ fields --
    def setFields(self, U, T, p, boundary=None):
        """Initial primitive fields on the internal cells ((nIC,3), (nIC,1),
        (nIC,1)) and their boundary conditions {field: {patch: spec}}; the
        role of readFields for generated cases."""
        if boundary is not None:
            for k in boundary:
                self.boundaries[k] = dict(boundary[k])
        U = np.asarray(U, np.float64).reshape(-1, 3)
        T = np.asarray(T, np.float64).reshape(-1, 1)
        p = np.asarray(p, np.float64).reshape(-1, 1)
        rho, rhoU, rhoE = self.conservative(U, T, p)
        self.setState(np.concatenate([rho.T, rhoU.T, rhoE.T]))
        return [rho, rhoU, rhoE]

    def setState(self, u):
        """Conservative state (5, nIC) as rows rho, rhoU, rhoE."""
        torch = self._torch()
        dt = torch.float32 if config.precision == np.float32 else torch.float64
        self.u = torch.as_tensor(np.ascontiguousarray(u), dtype=dt).to('cuda').contiguous()
        self.firstRun = False

    def readFields(self, t, suffix=''):
        """Read U, T, p (+ boundaryField) of time directory t from the case
        (adFVM/density.py:113-137).  The conservative fields rho, rhoU, rhoE
        are read instead of converting U, T, p when readConservative is set,
        or when t > 0 and the directory holds them (directories this solver
        wrote): a restart or checkpoint then reproduces the state exactly."""
        import os
        from . import foam
        U, T, p, bnd = foam.read_fields(self.mesh, t, suffix)
        d = foam.time_dir(self.mesh.caseDir, t)
        names = [n + suffix for n in self.names]
        if self.readConservative or (t != 0 and all(os.path.exists(os.path.join(d, n)) for n in names)):
            nIC = self.mesh.nInternalCells
            rho = foam.read_field(self.mesh.caseDir, t, names[0], nIC, 1)[0]
            rhoU = foam.read_field(self.mesh.caseDir, t, names[1], nIC, 3)[0]
            rhoE = foam.read_field(self.mesh.caseDir, t, names[2], nIC, 1)[0]
            for k in bnd:
                self.boundaries[k] = dict(bnd[k])
            self.setState(np.concatenate([rho.T, rhoU.T, rhoE.T]))
            return [rho, rhoU, rhoE]
        return self.setFields(U, T, p, boundary=bnd)

    def getFields(self):
        """Host conservative fields [rho, rhoU, rhoE]."""
        u = self.u.double().cpu().numpy()
        return [u[0:1].T.copy(), u[1:4].T.copy(), u[4:5].T.copy()]

    def writeFields(self, t):
        """Write the primitive and conservative fields at time t."""
        from . import foam
        rho, rhoU, rhoE = self.getFields()
        U, T, p = self.primitive(rho, rhoU, rhoE)
        foam.write_fields(self.mesh, t, {'U': U, 'T': T, 'p': p}, self.boundaries, self.mesh.defaultBoundary)
        foam.write_fields(self.mesh, t, {'rho': rho, 'rhoU': rhoU, 'rhoE': rhoE}, {})

    def updateSource(self, source, perturb=False):
        """Source terms (rho, rhoU, rhoE) added to the residual
        (adFVM/solver.py:116-133)."""
        torch = self._torch()
        s = np.concatenate([np.asarray(source[0], np.float64).reshape(-1, 1).T,
                            np.asarray(source[1], np.float64).reshape(-1, 3).T,
                            np.asarray(source[2], np.float64).reshape(-1, 1).T]) \
            if not np.isscalar(source[0]) else np.zeros((5, self.mesh.nInternalCells)) + \
            np.array([source[0]] + [source[1]] * 3 + [source[2]])[:, None]
        dt = torch.float32 if config.precision == np.float32 else torch.float64
        st = torch.as_tensor(s, dtype=dt).to('cuda')
        if perturb and self.source is not None:
            self.source += st
        else:
            self.source = st.contiguous()
        if not bool((self.source != 0).any()):
            self.source = None

    # ----------------------------------------------------------- compile --
    def physics(self):
        obj = self.objective
        ph = dict(gamma=self.gamma, Cp=self.Cp, Pr=self.Pr, mu=mu_model(self.mu),
                  riemann=_lib.RIEMANN_ROE if self.riemannSolver == 'eulerRoe' else _lib.RIEMANN_LF,
                  boundaryRiemann=_lib.RIEMANN_ROE if self.boundaryRiemannSolver == 'eulerRoe' else _lib.RIEMANN_LF,
                  small=config.small_for(config.precision),
                  integrator=_lib.RK_SSPRK3 if self.timeIntegrator == 'SSPRK' else _lib.RK_EULER,
                  objective=obj.code if obj else _lib.OBJ_NONE,
                  limiter=1 if self.faceReconstructor == 'secondOrderLimited' else 0)
        if obj is not None and obj.patch is not None:
            ph['objectivePatch'] = obj.patch
        return ph

    def compile(self):
        """Create the device context: Function.compile + Function.initialize
        of the reference (solver.py:61-71)."""
        from .context import Context
        specs = BCs.validate(self.mesh, self.boundaries)
        self.ctx = Context(self.mesh, specs, self.physics(), config.precision)
        if self.u is not None and self.u.dtype != self.ctx.tdtype:
            self.u = self.u.to(self.ctx.tdtype).contiguous()
        self.map = PrimalMap(self)
        return self


def bc_inputs(solver, specs=None):
    """The boundary-value arrays of the reference's Function inputs, in its
    order (getBoundaryTensor, solver.py:84-85, field.py:141-146): fields U,
    T, p; patches in sorted order (mesh.py:100-101); per patch the inputs
    its condition creates, in creation order (BCs.py:70-184: fixedValue
    'value'; CBC_UPT U0, T0, p0; CBC_TOTAL_PT Tt, pt[, direction];
    calculated and the copy rules none).  Returns [(field, patch, key,
    components)].  The gradient fields' default conditions (density.py:
    108-110) create none."""
    from . import BCs
    specs = specs if specs is not None else BCs.validate(solver.mesh, solver.boundaries)
    comps = {'value': None, 'U0': 3, 'T0': 1, 'p0': 1, 'Tt': 1, 'pt': 1, 'direction': 3}
    order = {'fixedValue': ('value',), 'CBC_UPT': ('U0', 'T0', 'p0'), 'CBC_TOTAL_PT': ('Tt', 'pt', 'direction')}
    out = []
    for field, k in (('U', 3), ('T', 1), ('p', 1)):
        for patch in solver.mesh.sortedPatches:
            spec = specs[patch][field]
            for key in order.get(spec['type'], ()):
                if key in spec:
                    out.append((field, patch, key, comps[key] or k))
    return out


def _live(specs, field, patch, key):
    """Whether the kernels read this boundary value: a fixedValue's value,
    or a characteristic condition's values on p (density.py:333-337 applies
    the p field's characteristic condition; the U and T copies are unused)."""
    t = specs[patch][field]['type']
    return t == 'fixedValue' or (field == 'p' and t in ('CBC_UPT', 'CBC_TOTAL_PT'))


class PositionalInputs(object):
    """The reference Function's positional inputs after the state (and dt):
    mesh.getTensor() + mesh.getScalar() + source terms + boundary-value
    arrays (density.py:97-105, solver.py:312-317).  The mesh arguments must
    be the context's mesh (a mesh perturbation, solver.py:254-261, is not
    supported and raises); the source and boundary values are used: the
    source is uploaded for the call, boundary values that differ from the
    context's are written to the device (Context.set_bc_value) and kept."""

    def __init__(self, solver):
        from . import BCs
        self.solver = solver
        self.specs = BCs.validate(solver.mesh, solver.boundaries)
        self.bcs = bc_inputs(solver, self.specs)
        m = solver.mesh
        self.mesh_args = m.getTensor() + m.getScalar()
        self.n_mesh = len(self.mesh_args)
        self.n = self.n_mesh + 3 + len(self.bcs)
        from .BCs import parse_value
        self.values = {}
        for f, p, key, k in self.bcs:
            try:
                self.values[(f, p, key)] = parse_value(self.specs[p][f][key], m.boundary[p]['nFaces'], k)
            except ValueError:
                self.values[(f, p, key)] = None
        self._source = None

    def consume(self, rest):
        """Checks and applies rest (everything after the state and dt);
        returns the device source tensor (or None) for the call."""
        s = self.solver
        if len(rest) != self.n:
            raise ValueError('expected %d positional arguments after the state and dt (%d mesh, 3 source, %d '
                             'boundary-value arrays), got %d' % (self.n, self.n_mesh, len(self.bcs), len(rest)))
        for i, (a, b) in enumerate(zip(rest[:self.n_mesh], self.mesh_args)):
            if a is b:
                continue
            if not np.array_equal(np.asarray(a), np.asarray(b)):
                raise NotImplementedError('mesh argument %d differs from the context\'s mesh: mesh perturbations '
                                          'are not supported' % i)
        src = rest[self.n_mesh:self.n_mesh + 3]
        source = None
        if any(np.any(np.asarray(x) != 0) for x in src):
            nIC = s.mesh.nInternalCells
            host = np.concatenate([np.asarray(src[0], np.float64).reshape(nIC, 1).T,
                                   np.asarray(src[1], np.float64).reshape(nIC, 3).T,
                                   np.asarray(src[2], np.float64).reshape(nIC, 1).T])
            if self._source is None or not np.array_equal(self._source[0], host):
                torch = s._torch()
                self._source = (host, torch.as_tensor(host, dtype=s.ctx.tdtype).to('cuda').contiguous())
            source = self._source[1]
        for (field, patch, key, k), a in zip(self.bcs, rest[self.n_mesh + 3:]):
            n = s.mesh.boundary[patch]['nFaces']
            v = np.asarray(a, np.float64)
            if v.size not in (k, n * k):
                raise ValueError('boundary value (%s, %s, %s): %d numbers for %d faces x %d' % (field, patch, key,
                                                                                            v.size, n, k))
            v = np.broadcast_to(v.reshape(-1, k), (n, k))
            cur = self.values[(field, patch, key)]
            if cur is not None and np.array_equal(cur, v):
                continue
            self.values[(field, patch, key)] = np.array(v)
            if _live(self.specs, field, patch, key):
                s.ctx.set_bc_value(field, patch, key, v)
                s.boundaries[field][patch][key] = np.array(v)
        return source


class PrimalMap(object):
    """The primal Function's call contract (density.py:101-105, called at
    solver.py:312-323): positional [rho (nIC,1), rhoU (nIC,3), rhoE (nIC,1),
    dt (1,1)] + mesh.getTensor() + mesh.getScalar() + the 3 source arrays +
    the boundary-value arrays (bc_inputs order); returns [rho', rhoU',
    rhoE', dtc (1,1), obj (1,1)] as host arrays.  The source and boundary
    values passed are the ones used (PositionalInputs).  Called with the
    state and dt alone, it uses the solver's own source and boundary
    values.  Options: return_reusable, replace_reusable (accepted; the
    state always comes back)."""

    OPTIONS = ('return_reusable', 'replace_reusable')

    def __init__(self, solver):
        self.solver = solver
        self.inputs = PositionalInputs(solver)

    def state(self, rho, rhoU, rhoE):
        s = self.solver
        torch = s._torch()
        ctx = s.ctx
        nIC = ctx.nIC
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=ctx.tdtype).to('cuda')
        return ctx.pack(dev(np.reshape(rho, (nIC, 1))), dev(np.reshape(rhoU, (nIC, 3))),
                        dev(np.reshape(rhoE, (nIC, 1))))

    def __call__(self, rho, rhoU, rhoE, dt, *rest, **options):
        for k in options:
            if k not in self.OPTIONS:
                raise TypeError('unknown option %r' % k)
        s = self.solver
        torch = s._torch()
        ctx = s.ctx
        source = self.inputs.consume(rest) if rest else s.source
        u = self.state(rho, rhoU, rhoE)
        dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        s.step(u, float(np.asarray(dt).reshape(-1)[0]), source, dtc, obj)
        r, rU, rE = ctx.unpack(u)
        f = lambda t: t.cpu().numpy()
        return [f(r), f(rU), f(rE), dtc.cpu().numpy().reshape(1, 1), obj.cpu().numpy().reshape(1, 1)]

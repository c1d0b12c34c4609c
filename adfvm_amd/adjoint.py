"""Reverse sweep of an unsteady run: the reference's Adjoint solver
(apps/adjoint.py:30-404) over the device kernels.

Two layers:

  sweep(ctx, u0, dts, ...)   in-process reverse of nSteps steps from the
                             device state u0 with a checkpoint schedule
                             (store_all / uniform / revolve, see
                             adfvm_amd.checkpoint); step tapes keep each
                             re-run segment's stage states and gradients in
                             HBM so that no step is recomputed twice.
  Adjoint(primal)            the case-file workflow of apps/adjoint.py: the
                             primal's written time directories are the
                             checkpoints, segments of writeInterval steps are
                             re-run and reversed from the last, the source
                             sensitivity of every perturbation is accumulated
                             in FP64, status/sensTimeSeries/energyTimeSeries
                             files and the adjoint fields are written.

Adjoint fields are gradients of the objective with respect to the
conservative state, i.e. "times volume" as the reference keeps them
internally (apps/adjoint.py:185-189); they are divided by the cell volumes
when written.
"""
import os
import time

import numpy as np

from . import checkpoint
from . import foam


class DeviceStepper(object):
    """The checkpoint.stepper interface over a Context.

    dts[n] is step n's dt; obja(n) its objective weight; ua the adjoint
    state (updated in place); source_grad and bc_grad, if given, accumulate
    the source and boundary-value gradients; after_reverse(n, u_n) runs after
    step n is reversed."""

    def __init__(self, ctx, dts, ua, source=None, source_grad=None, obja=1.0, tapes=True, after_reverse=None,
                 bc_grad=None, solver=None):
        self.ctx = ctx
        self.dts = dts
        self.ua = ua
        self.source = source
        self.source_grad = source_grad
        self.bc_grad = bc_grad
        self.obja = obja if callable(obja) else (lambda n, w=float(obja): w)
        self.tapes = tapes
        self.after_reverse = after_reverse
        self.forwards = 0
        # a solver whose objective is a case-file function: its cotangent
        # enters each reversed step (the library has no kernel for it)
        from .objectives import PythonObjective
        self.solver = solver if solver is not None and isinstance(solver.objective, PythonObjective) else None

    def clone(self, u):
        return u.clone()

    def forward(self, u, n, keep):
        tape = self.ctx.new_tapes(1)[0] if (keep and self.tapes) else None
        self.ctx.primal_step(u, self.dts[n], source=self.source, tape=tape)
        self.forwards += 1
        return tape

    def reverse(self, u_n, n, tape):
        og, w = None, self.obja(n)
        if self.solver is not None:
            from .objectives import stage_cotangent
            og = stage_cotangent(self.ctx, self.solver, u_n, self.dts[n], self.source, tape, w, self.bc_grad)
            if og is not None:
                w = 1.0
        self.ctx.adjoint_step(u_n, self.dts[n], self.ua, obja=w, source=self.source,
                              source_grad=self.source_grad, bc_grad=self.bc_grad, tape=tape, obj_grad=og)
        if self.after_reverse is not None:
            self.after_reverse(n, u_n)


def sweep(ctx, u0, dts, schedule='store_all', interval=None, snapshots=None, obja=1.0, ua=None, source=None,
          source_grad=None, tapes=True, after_reverse=None, checkpoints=None, bc_grad=None, solver=None):
    """Reverse nSteps = len(dts) steps from the device state u0 (unchanged).

    Returns ua = d/du0 of sum_n obja(n) J_n + <ua_in, u_N> (ua_in = the given
    ua, zero by default), the objective J_n evaluated inside step n as the
    primal step does.  source_grad (device (5, nIC)) and bc_grad (device
    (5, nBoundaryFaces), Context.bc_state) accumulate the source and
    boundary-value gradients of the same functional.  schedule: 'store_all', 'uniform'
    (interval steps per segment; `checkpoints` may hold segment starts
    n -> state) or 'revolve' (`snapshots` extra states).  solver: an RCF
    whose objective is a case-file function objective(fields, solver) (the
    context then has no objective of its own); J_n is then that function's."""
    n = len(dts)
    if ua is None:
        ua = ctx.state(0.)
    st = DeviceStepper(ctx, dts, ua, source, source_grad, obja, tapes, after_reverse, bc_grad, solver)
    if schedule == 'store_all':
        checkpoint.store_all(st, u0, 0, n)
    elif schedule == 'uniform':
        checkpoint.uniform(st, u0, 0, n, int(interval), checkpoints)
    elif schedule == 'revolve':
        checkpoint.revolve(st, u0, 0, n, int(snapshots))
    else:
        raise ValueError('unknown schedule %r' % (schedule,))
    sweep.last_forward_steps = st.forwards
    return ua


def initial_velocity_gradient(ua, u0):
    """Chain rule through the conservative map at fixed rho and T for a
    perturbation of the initial velocity: with rhoU = rho U and
    rhoE = rho (Cv T + |U|^2/2), d/dU = rho (ua_rhoU + ua_rhoE U).  Returns
    (3, nIC) in FP64."""
    a = ua.double()
    u = u0.double()
    rho = u[0]
    U = u[1:4] / rho
    return rho * (a[1:4] + a[4] * U)


class Adjoint(object):
    """apps/adjoint.py's Adjoint solver for a primal RCF whose orig run wrote
    its checkpoints (time directories every writeInterval steps) and its
    time-step log."""

    def __init__(self, primal, scaling=0.0, viscosityType=None, viscosityScaler=None, viscousIterations=1000):
        """scaling, viscosityType, viscosityScaler: the case file's adjParams
        (apps/adjoint.py:46-48).  With a viscosityType ('uniform',
        'abarbanel', 'turkel') every viscousInterval-th reversed step is
        followed by the adjoint-viscosity smoothing (apps/adjoint.py:127-140,
        287-291) of viscousIterations Jacobi sweeps; viscosityScaler is
        used only by the reference's CPU/PETSc variant."""
        self.primal = primal
        self.viscosityType = viscosityType
        self.viscosityScaler = viscosityScaler
        self.viscousIterations = viscousIterations
        self.mesh = primal.mesh
        self.names = [name + 'a' for name in primal.names]
        self.dimensions = primal.dimensions
        self.statusFile = primal.statusFile
        self.sensTimeSeriesFile = self.mesh.case + 'sensTimeSeries.txt'
        self.energyTimeSeriesFile = self.mesh.case + 'energyTimeSeries.txt'
        self.scaling = scaling
        self.forceReadFields = False
        self.homogeneousAdjoint = False
        self.ua = None

    def compileSolver(self, parameters=('source',)):
        """The positional adjoint Functions of apps/adjoint.py:94-141:
        self.map (primal_grad) and self.viscousMap (primal_grad_viscous,
        when a viscosityType is set), for the parameter parameters[0]."""
        param = list(parameters)[0] if parameters else 'source'
        self.map = AdjointMap(self, param)
        self.viscousMap = AdjointMap(self, param, viscous=True) if self.viscosityType else None
        return self

    # ------------------------------------------------------------- fields --
    def createFields(self):
        self.ua = self.primal.ctx.state(0.)
        return self.ua

    def _volumes(self):
        torch = self.primal._torch()
        V = torch.as_tensor(self.mesh.volumes.reshape(-1), dtype=self.primal.ctx.tdtype).to('cuda')
        return V

    def readFields(self, t):
        """Adjoint fields of time t, times the cell volumes (apps/adjoint.py:
        185-189)."""
        nIC = self.mesh.nInternalCells
        rows = []
        for name, (k,) in zip(self.names, self.dimensions):
            v = foam.read_field(self.mesh.caseDir, t, name, nIC, k)[0]
            rows.append(v.T)
        torch = self.primal._torch()
        ua = torch.as_tensor(np.concatenate(rows), dtype=self.primal.ctx.tdtype).to('cuda')
        self.ua = (ua * self._volumes()).contiguous()
        return self.ua

    def writeFields(self, t):
        a = (self.ua / self._volumes()).double().cpu().numpy()
        fields = {self.names[0]: a[0:1].T, self.names[1]: a[1:4].T, self.names[2]: a[4:5].T}
        foam.write_fields(self.mesh, t, fields, {})

    def readStatusFile(self):
        return self.primal.readStatusFile()

    def writeStatusFile(self, data):
        self.primal.writeStatusFile(data)

    def removeStatusFile(self):
        self.primal.removeStatusFile()

    # ----------------------------------------------------------- the run --
    def initPrimalData(self, nSteps, startTime, dt, nPerturb):
        """(result, firstCheckpoint) from the status file and the primal's
        time-step log, (nSteps+1, 2) with the final time appended
        (apps/adjoint.py:146-172)."""
        if os.path.exists(self.statusFile):
            firstCheckpoint, result = self.readStatusFile()
        else:
            firstCheckpoint, result = 0, [0.] * nPerturb
        if not os.path.exists(self.primal.timeStepFile):
            if not self.primal.fixedTimeStep:
                raise RuntimeError('no time-step log %s and the primal time step is not fixed'
                                   % self.primal.timeStepFile)
            timeSteps = np.stack([startTime + np.arange(nSteps) * dt, np.full(nSteps, dt)], 1)
        else:
            timeSteps = np.loadtxt(self.primal.timeStepFile, ndmin=2)
        if timeSteps.shape != (nSteps, 2):
            raise RuntimeError('time-step log has shape %s, expected (%d, 2)' % (timeSteps.shape, nSteps))
        timeSteps = np.concatenate([timeSteps, [[np.sum(timeSteps[-1]).round(12), 0.]]])
        return (result, firstCheckpoint), (timeSteps,)

    def run(self, checkpointData, primalData, nSteps, writeInterval, perturb=(), parameters=('source',),
            source=None, sampleInterval=1, reportInterval=1, avgStart=0, runCheckpoints=2 ** 30, tapes=True,
            verbose=False, viscousInterval=1):
        """The reverse sweep of apps/adjoint.py:176-404 for the parameter
        parameters[0]: 'source', 'mesh' (the arrays of Mesh.gradFields) or a
        boundary-condition value ('BCs', field, patch, key)
        (apps/adjoint.py:104-118); returns the (unaveraged) sensitivities,
        one per perturbation."""
        param = list(parameters)[0] if parameters else 'source'
        if not (param in ('source', 'mesh') or (isinstance(param, tuple) and len(param) == 4 and param[0] == 'BCs')):
            raise NotImplementedError('parameter %r (source, mesh and BC values are supported)' % (param,))
        primal = self.primal
        ctx = primal.ctx
        torch = primal._torch()
        mesh = self.mesh
        result, firstCheckpoint = checkpointData
        result = list(result)
        (timeSteps,) = primalData
        startTime = timeSteps[nSteps - firstCheckpoint * writeInterval][0]
        if firstCheckpoint > 0 or self.forceReadFields:
            self.readFields(startTime)
        elif self.ua is None:
            self.createFields()
        ua = self.ua

        perturbations = []
        bgrad = bc_rows = bc_faces = mgrad = None
        if param == 'mesh':
            mgrad = ctx.mesh_grad_state()
        elif param != 'source':
            _, phi, patchID, key = param
            bc_rows, bc_faces = ctx.bc_param(phi, patchID, key)
            bgrad = ctx.bc_state()
        for p in perturb:
            values = p(None, mesh, 0)
            values = list(values) if isinstance(values, (list, tuple)) else [values]
            if mgrad is not None:        # the Mesh.gradFields arrays, in the gradient's layout
                views = ctx.mesh_grad_arrays(mgrad)
                if len(values) != len(views):
                    raise ValueError('a mesh perturbation has %d arrays (Mesh.gradFields)' % len(views))
                d = np.concatenate([np.broadcast_to(np.asarray(v, np.float64).reshape(-1, *tuple(w.shape[1:])),
                                                    tuple(w.shape)).reshape(-1) for v, w in zip(values, views)])
            elif bgrad is None:
                vals = [np.asarray(v, np.float64).reshape(mesh.nInternalCells, -1) for v in values]
                d = np.concatenate([v.T for v in vals])
            else:                        # one (nFaces, components) array, broadcast
                k = bc_rows.stop - bc_rows.start
                nf = bc_faces.stop - bc_faces.start
                d = np.broadcast_to(np.asarray(values[0], np.float64).reshape(-1, k), (nf, k)).T
            perturbations.append(torch.as_tensor(np.ascontiguousarray(d), dtype=ctx.tdtype).to('cuda').contiguous())
        sgrad = ctx.state(0.)
        dsum = torch.zeros(1, dtype=torch.float64, device='cuda')
        energy = torch.zeros(1, dtype=torch.float64, device='cuda')
        sensTimeSeries, energyTimeSeries = [], []
        obja = 0. if self.homogeneousAdjoint else 1.
        from .objectives import PythonObjective, stage_cotangent
        pyobj = isinstance(primal.objective, PythonObjective)

        totalCheckpoints = nSteps // writeInterval
        nCheckpoints = min(firstCheckpoint + runCheckpoints, totalCheckpoints)
        cp = firstCheckpoint
        t = startTime
        while cp < nCheckpoints:
            primalIndex = nSteps - (cp + 1) * writeInterval
            t = timeSteps[primalIndex, 0]
            dts = timeSteps[primalIndex:primalIndex + writeInterval, 1]
            primal.readFields(t)
            if source is not None:
                primal.updateSource(source(primal.getFields(), mesh, 0))
            # forward re-run of the segment (apps/adjoint.py:217), keeping
            # each step's start state and tape
            start = time.time()
            u = primal.u
            states, tps = [], []
            for k in range(writeInterval):
                states.append(u.clone())
                tps.append(ctx.new_tapes(1)[0] if tapes else None)
                ctx.primal_step(u, dts[k], source=primal.source, tape=tps[-1])
            if cp == 0:
                self.writeFields(timeSteps[-1][0])
            ctx.adjoint_energy(ua, u, out=energy)
            energyTimeSeries.append(energy.item() ** 0.5)
            for r in range(1, writeInterval + 1):              # reverse count, apps/adjoint.py:244-249
                k = writeInterval - r
                og, w = None, obja
                if pyobj:
                    og = stage_cotangent(ctx, primal, states[k], dts[k], primal.source, tps[k], obja, bgrad)
                    w = 1.0 if og is not None else obja
                ctx.adjoint_step(states[k], dts[k], ua, obja=w, source=primal.source, source_grad=sgrad,
                                 bc_grad=bgrad, tape=tps[k], obj_grad=og)
                if self.viscosityType and r % viscousInterval == 0:       # apps/adjoint.py:250,288-289
                    ctx.adjoint_viscosity(states[k], ua, dts[k], self.viscosityType, self.scaling,
                                          iterations=self.viscousIterations)
                if r % reportInterval == 0:
                    ctx.adjoint_energy(ua, states[k], out=energy)
                    energyTimeSeries.append(energy.item() ** 0.5)
                states[k] = tps[k] = None
                if r % sampleInterval == 0:
                    sens = []
                    g = mgrad if mgrad is not None else (sgrad if bgrad is None else
                                                         bgrad[bc_rows, bc_faces].contiguous())
                    for d in perturbations:
                        ctx.dot(g, d, out=dsum)
                        sens.append(dsum.item())
                    sgrad.zero_()
                    if bgrad is not None:
                        bgrad.zero_()
                    if mgrad is not None:
                        mgrad.zero_()
                    if (nSteps - (primalIndex + k)) > avgStart:
                        for i, sv in enumerate(sens):
                            result[i] += sv
                    sensTimeSeries.extend([[sv / sampleInterval for sv in sens]] * sampleInterval)
            if verbose:
                print('adjoint segment %d/%d: %d steps in %.3f s' % (cp + 1, totalCheckpoints, writeInterval,
                                                                   time.time() - start))
            cp += 1
            self.writeStatusFile([cp, result])
            with open(self.sensTimeSeriesFile, 'ab') as f:
                np.savetxt(f, np.array(sensTimeSeries).reshape(len(sensTimeSeries), -1))
            with open(self.energyTimeSeriesFile, 'ab') as f:
                np.savetxt(f, np.array(energyTimeSeries))
            sensTimeSeries, energyTimeSeries = [], []
            self.writeFields(t)
        if mgrad is not None:
            ctx.set_mesh_grad(None)
        self.finished = cp >= totalCheckpoints
        if self.finished:
            self.removeStatusFile()
        return result


class AdjointMap(object):
    """The primal_grad Function's call contract (apps/adjoint.py:94-126,
    called at :268-291): positional primal inputs (as PrimalMap takes them)
    + [rhoa (nIC,1), rhoUa (nIC,3), rhoEa (nIC,1), dtca (1,1), obja (1,1)]
    + [scaling (1,1)]; returns [rhoa', rhoUa', rhoEa'] + the parameter
    gradients: the 3 source arrays' for 'source' (apps/adjoint.py:104-106),
    the 10 arrays of Mesh.gradFields for 'mesh' (:107-110), the value
    array's (nFaces, components) for ('BCs', field, patch, key) (:111-118).  Adjoint fields are times volume, as the reference keeps
    them (apps/adjoint.py:185-189).  The parameter gradients accumulate over
    calls (the reference's static outputs): return_static returns them,
    zero_static zeroes them after (both default to True); with
    return_static False their slots hold None.  dtca must be 0 (the
    reference always seeds it so, :268); a non-zero seed raises.  With
    viscous=True the adjoint-viscosity smoothing of the Adjoint's
    viscosityType follows the step (the viscous map, :127-140) with the
    passed scaling."""

    OPTIONS = ('return_static', 'zero_static', 'return_reusable', 'replace_reusable')

    def __init__(self, adjoint, parameter='source', viscous=False):
        from .density import PrimalMap
        primal = adjoint.primal
        self.adjoint, self.primal, self.ctx = adjoint, primal, primal.ctx
        self.pmap = PrimalMap(primal)
        self.parameter = parameter
        self.viscous = viscous
        if not (parameter in ('source', 'mesh') or (isinstance(parameter, tuple) and len(parameter) == 4
                                                    and parameter[0] == 'BCs')):
            raise NotImplementedError('parameter %r (source, mesh and BC values are supported)' % (parameter,))
        if isinstance(parameter, tuple):
            _, phi, patchID, key = parameter
            self.rows, self.faces = self.ctx.bc_param(phi, patchID, key)
        self.sgrad = self.ctx.state(0.)
        self.bgrad = self.ctx.bc_state()
        self.mgrad = None
        if parameter == 'mesh':               # registered for this map's steps only
            self.mgrad = self.ctx.mesh_grad_state()
            self.ctx.set_mesh_grad(None)

    def __call__(self, *args, **options):
        for k in options:
            if k not in self.OPTIONS:
                raise TypeError('unknown option %r' % k)
        ret_static = options.get('return_static', True)
        zero_static = options.get('zero_static', True)
        pmap, ctx = self.pmap, self.ctx
        n_primal = 4 + pmap.inputs.n
        if len(args) != n_primal + 6:
            raise ValueError('expected %d positional arguments (%d primal inputs, 5 adjoint inputs, scaling), got %d'
                             % (n_primal + 6, n_primal, len(args)))
        rho, rhoU, rhoE, dt = args[:4]
        source = pmap.inputs.consume(args[4:n_primal])
        rhoa, rhoUa, rhoEa, dtca, obja, scaling = args[n_primal:]
        if float(np.asarray(dtca).reshape(-1)[0]) != 0.0:
            raise NotImplementedError('a non-zero dtc seed (the reference seeds 0, apps/adjoint.py:268)')
        obja = float(np.asarray(obja).reshape(-1)[0])
        dt = float(np.asarray(dt).reshape(-1)[0])
        u = pmap.state(rho, rhoU, rhoE)
        ua = pmap.state(rhoa, rhoUa, rhoEa)
        primal = self.primal
        og, w = None, obja
        from .objectives import PythonObjective, stage_cotangent
        if isinstance(primal.objective, PythonObjective):
            og = stage_cotangent(ctx, primal, u, dt, source, None, obja, self.bgrad)
            w = 1.0 if og is not None else obja
        if self.mgrad is not None:
            ctx.set_mesh_grad(self.mgrad)
        try:
            ctx.adjoint_step(u, dt, ua, obja=w, source=source, source_grad=self.sgrad, bc_grad=self.bgrad,
                             obj_grad=og)
        finally:
            if self.mgrad is not None:
                ctx.set_mesh_grad(None)
        if self.viscous and self.adjoint.viscosityType:
            ctx.adjoint_viscosity(u, ua, dt, self.adjoint.viscosityType,
                                  float(np.asarray(scaling).reshape(-1)[0]),
                                  iterations=self.adjoint.viscousIterations)
        r, rU, rE = ctx.unpack(ua)
        out = [r.cpu().numpy(), rU.cpu().numpy(), rE.cpu().numpy()]
        n_static = {'source': 3, 'mesh': 10}.get(self.parameter, 1) if not isinstance(self.parameter, tuple) else 1
        if ret_static:
            if self.parameter == 'source':
                g = self.sgrad.double().cpu().numpy()
                out += [g[0:1].T.copy(), g[1:4].T.copy(), g[4:5].T.copy()]
            elif self.parameter == 'mesh':
                out += [a.double().cpu().numpy().copy() for a in ctx.mesh_grad_arrays(self.mgrad)]
            else:
                out.append(self.bgrad[self.rows, self.faces].double().cpu().numpy().T.copy())
        else:
            out += [None] * n_static
        if zero_static:
            self.sgrad.zero_()
            self.bgrad.zero_()
            if self.mgrad is not None:
                self.mgrad.zero_()
        return out

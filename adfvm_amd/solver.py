"""Solver base class and the host time loop (adFVM/solver.py).

The loop keeps the state resident on the GPU as a (5, nInternalCells) tensor
and calls the library step once per time step, as the reference calls its
compiled `self.map` (solver.py:312-323).  Host synchronisation happens only
where the reference's loop needs a host value: the CFL time-step update
(solver.py:359-366), report and write steps.  The time-integrated objective
is accumulated in FP64 on the device.
"""
import json
import os
import time

import numpy as np

from . import config
from . import foam
from .mesh import Mesh
from .meshgen import MeshData


class Solver(object):
    defaultConfig = {
        'timeIntegrator': 'euler',
        'objective': None,
        'objectiveString': None,
        'adjoint': False,
        'dynamicMesh': False,
        'localTimeStep': False,
        'fixedTimeStep': False,
        'stepFactor': 1.0,
        'postpro': [],
        'sourceTerms': [],
        'timeSeriesAppend': '',
    }

    def __init__(self, case, **userConfig):
        """case: an OpenFOAM case directory, a built Mesh, or generated
        MeshData (adfvm_amd.meshgen).  Unknown configuration keys are
        rejected as in solver.py:37-39 (the class default dict is copied,
        not mutated)."""
        full = dict(self.__class__.defaultConfig)
        for key in userConfig:
            if key not in full:
                raise KeyError('unknown solver option %r' % key)
        full.update(userConfig)
        for key, v in full.items():
            setattr(self, key, v)
        if isinstance(case, Mesh):
            self.mesh = case
        elif isinstance(case, MeshData):
            self.mesh = Mesh.from_data(case)
        else:
            self.mesh = Mesh.create(case)
        caseDir = self.mesh.case or ''
        self.resultFile = caseDir + 'objective.txt'
        self.statusFile = caseDir + 'status.json'
        self.timeSeriesFile = caseDir + 'timeSeries{}.txt'.format(self.timeSeriesAppend)
        self.timeStepFile = caseDir + 'timeSteps.txt'
        self.stage = 0
        self.firstRun = True
        self.ctx = None
        self.u = None                      # device state (5, nIC)
        self.source = None                 # device source (5, nIC) or None
        return

    # ------------------------------------------------------------- status --
    def readStatusFile(self):
        with open(self.statusFile) as f:
            return json.load(f)

    def writeStatusFile(self, data):
        with open(self.statusFile, 'w') as f:
            json.dump(data, f)

    def removeStatusFile(self):
        try:
            os.remove(self.statusFile)
        except OSError:
            pass

    # --------------------------------------------------------- host loop --
    def run(self, endTime=np.inf, writeInterval=config.LARGE, reportInterval=1, startTime=0.0, dt=1e-3,
            nSteps=config.LARGE, startIndex=0, result=0., mode='simulation', source=None, perturbation=None,
            avgStart=0, tapes=None):
        """The reference's Solver.run (solver.py:210-411).

        dt: a number, or an array of per-step dt indexed by the absolute
        step index (replay of a logged run).
        mode 'forward' returns the list of device states u_n at the start of
        each step (and, with `tapes`, fills tapes[n] with step n's tape).
        source(fields, mesh, t) -> (rho, rhoU, rhoE) source terms;
        perturbation = (parameters, perturb) adds perturb(...) to the
        'source' parameter or a BC value ('BCs', field, patch, key) for the
        run and reverts it after (solver.py:243-277,406-407).
        Returns the time-integrated objective (FP64), or the states."""
        torch = self._torch()
        if self.mesh.caseDir and os.path.isdir(foam.time_dir(self.mesh.caseDir, startTime)):
            self.readFields(startTime)          # solver.py:217, for cases on disk
        if self.u is None:
            raise RuntimeError('no fields: call readFields()/setFields() first')
        if self.ctx is None:
            self.compile()
        mesh = self.mesh
        t = startTime
        dts = dt
        timeIndex = startIndex
        if isinstance(dts, np.ndarray):
            dt = float(dts[timeIndex])          # absolute step index, solver.py:226-227
        fixed = self.fixedTimeStep or isinstance(dts, np.ndarray)
        timeSeries, timeSteps, solutions = [], [], []
        if source is not None:
            self.updateSource(source(self.getFields(), mesh, t))
        if perturbation:
            self._perturb(perturbation, t, revert=False)

        dev = 'cuda'
        obj = torch.zeros(1, dtype=torch.float64, device=dev)
        dtc = torch.zeros(1, dtype=torch.float64, device=dev)
        acc = torch.zeros(1, dtype=torch.float64, device=dev)
        series = []
        start = time.time()

        def iterate(t, i):
            return t < endTime and i < nSteps

        def updateTime(t0, dt):
            return round(t0 + dt, 12)

        while iterate(t, timeIndex):
            report = ((timeIndex + 1) % reportInterval == 0)
            write = ((timeIndex + 1) % writeInterval == 0) or not iterate(updateTime(t, dt), timeIndex + 1)
            if mode == 'forward':
                solutions.append(self.u.clone())
            tape = tapes[timeIndex - startIndex] if tapes is not None else None
            self.step(self.u, dt, self.source, dtc, obj, tape)
            timeSteps.append([t, dt])
            timeIndex += 1
            t = updateTime(t, dt)
            if timeIndex > avgStart:
                acc += obj
            series.append(obj.clone())
            if report:
                if not bool(torch.isfinite(self.u).all()):
                    raise FloatingPointError('NaN/Inf in the state at step %d (t=%g)' % (timeIndex, t))
            if isinstance(dts, np.ndarray):
                dt = float(dts[timeIndex]) if timeIndex < len(dts) else dt
            elif not fixed:
                dt = min(2 * self.CFL / dtc.item(), dt * self.stepFactor, endTime - t)
            if write and mode != 'forward':
                if mode in ('orig', 'simulation'):
                    self.writeFields(t)
                if mode == 'orig':
                    with open(self.timeStepFile, 'ab') as f:
                        np.savetxt(f, np.array(timeSteps))
                if mode in ('orig', 'perturb'):
                    with open(self.timeSeriesFile, 'ab') as f:
                        np.savetxt(f, torch.cat(series).cpu().numpy())
                series, timeSteps = [], []
                self.writeStatusFile([timeIndex, t, dt, result + acc.item()])
        if perturbation:
            self._perturb(perturbation, t, revert=True)
        self.lastTime, self.lastDt, self.wallTime = t, dt, time.time() - start
        if mode == 'forward':
            solutions.append(self.u.clone())     # nSteps+1 states, as solver.py:234-241,376-382
            return solutions
        return result + acc.item()

    def step(self, u, dt, source, dtc_out, obj_out, tape=None):
        """One time step of the device state u in place, dtc and the
        objective (FP64, on the first intermediate stage) into the device
        scalars dtc_out / obj_out.  A case-file objective is evaluated by
        torch on the objective stage, which the step's tape (a scratch one
        if none is given) or u_n holds."""
        from .objectives import PythonObjective
        pyobj = self.objective if isinstance(getattr(self, 'objective', None), PythonObjective) else None
        if pyobj is None:
            self.ctx.primal_step(u, dt, source=source, dtc_out=dtc_out, obj_out=obj_out, tape=tape)
            return
        tape = tape if tape is not None else self._scratch_tape()
        u_n = u.clone() if self.ctx.n_stages == 1 else None
        self.ctx.primal_step(u, dt, source=source, dtc_out=dtc_out, tape=tape)
        if obj_out is not None:
            obj_out.copy_(pyobj.value(self.ctx, self, self.ctx.stage_state(u_n, tape)))

    def _scratch_tape(self):
        t = getattr(self, '_objTape', None)
        if t is None or t.dtype != self.ctx.tdtype:
            t = self._objTape = self.ctx.new_tapes(1)[0]
        return t

    def _perturb(self, perturbation, t, revert):
        parameters, perturb = perturbation
        values = perturb(self.getFields(), self.mesh, t)
        if not isinstance(values, list) or (len(parameters) == 1 and len(values) > 1):
            values = [values]
        sign = -1. if revert else 1.
        for param, value in zip(parameters, values):
            if param == 'source':
                self.updateSource([sign * np.asarray(v, np.float64) for v in value], perturb=True)
            elif isinstance(param, tuple) and param[0] == 'BCs':
                # patch.inputs[k][1][:] += value (solver.py:262-270), on the
                # host spec and the device copy
                _, phi, patchID, key = param
                self._perturb_bc(phi, patchID, key, sign * np.asarray(value, np.float64))
            elif param == 'mesh':
                # field += delta for each of Mesh.gradFields (solver.py:254-261)
                self._perturb_mesh([sign * np.asarray(v, np.float64) for v in value])
            else:
                raise NotImplementedError('perturbation of %r (source, mesh and BC values are supported)'
                                          % (param,))

    def _perturb_mesh(self, deltas):
        """Adds deltas (the arrays of Mesh.gradFields, in order) to the mesh
        and rebuilds the device context from it.  The kernels divide each
        cell's face sum by its volume, so volumesL/volumesR perturbations
        must be those of volumes at the faces' owners/neighbours."""
        from .mesh import Mesh
        m = self.mesh
        if len(deltas) != len(Mesh.gradFields):
            raise ValueError('a mesh perturbation has %d arrays (Mesh.gradFields)' % len(Mesh.gradFields))
        d = dict(zip(Mesh.gradFields, deltas))
        nIF = m.nInternalFaces
        dV = d['volumes'].reshape(-1)
        for name, want in (('volumesL', dV[m.owner]), ('volumesR', dV[m.neighbour[:nIF]])):
            got = d[name].reshape(-1)
            if got.shape != want.shape or not np.allclose(got, want, rtol=1e-12, atol=0):
                raise NotImplementedError('%s perturbed apart from volumes: the step divides by cell volumes'
                                          % name)
        for name in Mesh.gradFields:
            a = getattr(m, name)
            a += d[name].reshape(a.shape)
        self.compile()

    def _perturb_bc(self, phi, patchID, key, delta):
        from . import BCs
        spec = self.boundaries[phi][patchID]
        n = self.mesh.boundary[patchID]['nFaces']
        k = 3 if key in ('U0', 'direction') or (key == 'value' and phi == 'U') else 1
        spec[key] = BCs.parse_value(spec[key], n, k) + np.broadcast_to(delta.reshape(-1, k), (n, k))
        self.ctx.set_bc_value(phi, patchID, key, spec[key])

    def _torch(self):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('adfvm_amd needs a HIP device')
        return torch

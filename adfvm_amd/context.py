"""Device context: one libadfvm context per (mesh, physics, boundary set).

Replaces adpy's `Function.initialize(localRank, mesh)` -> `external_init`
(adFVM/solver.py:70, adFVM/cpp/external.cpp:21-42): the mesh and boundary
values are uploaded once; each step then passes only device state pointers.

Tensors passed in are torch tensors on the current HIP device, contiguous,
in the context precision; state tensors have shape (5, nInternalCells).
"""
import ctypes as C

import numpy as np

from . import _lib

BC_CODES = {
    'cyclic': _lib.BC_CYCLIC, 'slidingPeriodic1D': _lib.BC_CYCLIC,
    'zeroGradient': _lib.BC_ZERO_GRADIENT, 'empty': _lib.BC_ZERO_GRADIENT, 'inletOutlet': _lib.BC_ZERO_GRADIENT,
    'symmetryPlane': _lib.BC_SYMMETRY, 'slip': _lib.BC_SYMMETRY,
    'fixedValue': _lib.BC_FIXED_VALUE, 'calculated': _lib.BC_CALCULATED,
    # characteristic conditions on U/T leave the field alone; p carries the rule
    'CBC_UPT': _lib.BC_CALCULATED, 'CBC_TOTAL_PT': _lib.BC_CALCULATED,
}


def _need_torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('adfvm_amd kernels need a HIP device (torch.cuda.is_available() is False)')
    return torch


def _dptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(None)


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream)


def dot_noalloc(a, b, out, stream=None):
    """sum_j a[j] b[j] in FP64 into out[0] through adfvm_dot_noalloc: no
    context, and neither the library nor this wrapper allocates device memory
    (out is the caller's).  a, b: contiguous HIP-device float32 or float64
    tensors of one dtype and element count (at most 2^27); out: a contiguous
    float64 HIP-device tensor.  Calls on one GPU share the library's partial
    buffer, so calls overlapping in time on one GPU race."""
    import torch
    for name, t in (('a', a), ('b', b)):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous() or \
                t.dtype not in (torch.float32, torch.float64):
            raise ValueError('%s: expected a contiguous float32/float64 HIP-device tensor' % name)
    if a.dtype != b.dtype or a.numel() != b.numel():
        raise ValueError('a, b: dtypes %s/%s, %d/%d elements' % (a.dtype, b.dtype, a.numel(), b.numel()))
    if not isinstance(out, torch.Tensor) or out.dtype != torch.float64 or not out.is_cuda or \
            not out.is_contiguous() or out.numel() < 1:
        raise ValueError('out: needs a contiguous float64 HIP-device tensor')
    lib = _lib.load()
    rc = lib.adfvm_dot_noalloc(4 if a.dtype == torch.float32 else 8, _dptr(a), _dptr(b), a.numel(), _dptr(out),
                               _stream(stream))
    if rc != 0:
        raise _lib.LibraryError('libadfvm error %d: %s' % (rc, lib.adfvm_noalloc_last_error().decode()))
    return out


class Context(object):
    """Owns an adfvm_ctx.

    mesh: adfvm_amd.mesh.Mesh (built).
    patch_bcs: dict patch -> {'U': spec, 'T': spec, 'p': spec} with spec a dict
      {'type': ..., 'value': array} in adFVM's boundaryField form; missing
      entries use the mesh's default boundary (mesh.py:702-711).
    physics: dict(gamma, Cp, Pr, mu=(model, mu0, muA, muTs), riemann,
      boundaryRiemann, small, integrator, objective, objectivePatch).
    """

    def __init__(self, mesh, patch_bcs, physics, precision):
        self.torch = _need_torch()
        lib = _lib.load()
        self.lib = lib
        self.mesh = mesh
        self.precision = np.dtype(precision)
        self.nIC = mesh.nInternalCells
        self.nBF = mesh.nFaces - mesh.nInternalFaces
        self._keep = []
        names = list(mesh.boundary.keys())
        self._bcinfo = {}
        descs = (_lib.PatchDesc * max(1, len(names)))()
        for i, name in enumerate(names):
            b = mesh.boundary[name]
            spec = patch_bcs.get(name, {})
            default = mesh.defaultBoundary[name]['type']
            fspec = {f: spec.get(f, {'type': default}) for f in ('U', 'T', 'p')}
            ptype = b['type']
            cbc = _lib.CBC_NONE
            if fspec['p']['type'] == 'CBC_UPT':
                cbc = _lib.CBC_UPT
            elif fspec['p']['type'] == 'CBC_TOTAL_PT':
                cbc = _lib.CBC_TOTAL_PT
            if ptype in ('cyclic', 'slidingPeriodic1D'):
                flux = _lib.FLUX_COUPLED
            elif cbc != _lib.CBC_NONE:
                flux = _lib.FLUX_CHARACTERISTIC
            else:
                flux = _lib.FLUX_BOUNDARY
            n = b['nFaces']
            d = descs[i]
            d.startFace, d.nFaces, d.fluxKind = b['startFace'], n, flux
            d.neighbourPatch = names.index(b['neighbourPatch']) if flux == _lib.FLUX_COUPLED else -1
            d.bcU, d.bcT, d.bcp = [BC_CODES[fspec[f]['type']] for f in ('U', 'T', 'p')]
            d.cbc = cbc
            self._bcinfo[name] = dict(index=i, cbc=cbc, types={f: fspec[f]['type'] for f in ('U', 'T', 'p')},
                                      start=b['startFace'] - mesh.nInternalFaces, n=n,
                                      direction=cbc == _lib.CBC_TOTAL_PT and fspec['p'].get('direction') is not None)

            def arr(v, k):
                if v is None:
                    return None
                a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64).reshape(-1, k) if np.size(v) != k
                                                         else np.asarray(v, np.float64).reshape(1, k), (n, k)))
                self._keep.append(a)
                return a
            if cbc == _lib.CBC_UPT:
                ps = fspec['p']
                vU, vT, vp = arr(ps['U0'], 3), arr(ps['T0'], 1), arr(ps['p0'], 1)
            else:
                vU = arr(fspec['U'].get('value'), 3)
                vT = arr(fspec['T'].get('value'), 1)
                vp = arr(fspec['p'].get('value'), 1)
            d.valueU, d.valueT, d.valuep = [_lib.ptr(a, C.c_double) for a in (vU, vT, vp)]
            if cbc == _lib.CBC_TOTAL_PT:
                ps = fspec['p']
                d.Tt = _lib.ptr(arr(ps['Tt'], 1), C.c_double)
                d.pt = _lib.ptr(arr(ps['pt'], 1), C.c_double)
                d.direction = _lib.ptr(arr(ps.get('direction'), 3), C.c_double)
        self._descs = descs

        sz = _lib.MeshSizes(mesh.nInternalFaces, mesh.nFaces, self.nBF, mesh.nInternalCells,
                            mesh.nGhostCells, mesh.nCells)
        c64 = lambda a: _lib.ptr(np.ascontiguousarray(a, np.float64), C.c_double)
        c32 = lambda a: _lib.ptr(np.ascontiguousarray(a, np.int32), C.c_int32)
        hold = [np.ascontiguousarray(getattr(mesh, k)) for k in
                ('owner', 'neighbour', 'cellFaces', 'cellOwner', 'cellNeighbours', 'areas', 'normals',
                 'weights', 'deltas', 'deltasUnit', 'linearWeights', 'quadraticWeights', 'volumes')]
        desc = _lib.SolverDesc(sz, c32(hold[0]), c32(hold[1]), c32(hold[2]), c32(hold[3]), c32(hold[4]),
                               c64(hold[5]), c64(hold[6]), c64(hold[7]), c64(hold[8]), c64(hold[9]),
                               c64(hold[10]), c64(hold[11]), c64(hold[12]), len(names), descs)
        ph = _lib.Physics()
        ph.precision = self.precision.itemsize
        ph.gamma, ph.Cp, ph.Pr = physics['gamma'], physics['Cp'], physics['Pr']
        ph.muModel, ph.mu0, ph.muA, ph.muTs = physics['mu']
        ph.riemann = physics.get('riemann', _lib.RIEMANN_ROE)
        ph.boundaryRiemann = physics.get('boundaryRiemann', _lib.RIEMANN_ROE)
        ph.small = physics['small']
        ph.integrator = physics.get('integrator', _lib.RK_SSPRK3)
        ph.objective = physics.get('objective', _lib.OBJ_NONE)
        obj_patch = physics.get('objectivePatch')
        ph.objectivePatch = names.index(obj_patch) if obj_patch is not None else -1
        ph.limiter = int(physics.get('limiter', 0))
        self.integrator = ph.integrator
        handle = C.c_void_p()
        rc = lib.adfvm_ctx_create(C.byref(desc), C.byref(ph), C.byref(handle))
        self.handle = handle
        if rc != _lib.ADFVM_OK:
            msg = lib.adfvm_last_error(handle).decode() if handle.value else ''
            lib.adfvm_ctx_destroy(handle)
            self.handle = None
            raise _lib.LibraryError('adfvm_ctx_create failed (%d): %s' % (rc, msg))

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value:
            self.lib.adfvm_ctx_destroy(h)
            self.handle = None

    # ------------------------------------------------------------ helpers --
    @property
    def tdtype(self):
        return self.torch.float32 if self.precision == np.float32 else self.torch.float64

    def _chk(self, rc):
        _lib.check(rc, self.handle)

    def _check(self, t, shape, name='array', dtype=None):
        """Raises ValueError unless t is a contiguous HIP-device tensor of
        the given shape and dtype (the context precision by default): the
        kernels index it with the strides that shape implies."""
        dtype = self.tdtype if dtype is None else dtype
        if not hasattr(t, 'dtype') or not hasattr(t, 'shape'):
            raise ValueError('%s: expected a torch tensor, got %r' % (name, type(t)))
        if t.dtype != dtype:
            raise ValueError('%s: dtype %s, the context computes in %s' % (name, t.dtype, dtype))
        if tuple(t.shape) != tuple(shape):
            raise ValueError('%s: shape %s, expected %s' % (name, tuple(t.shape), tuple(shape)))
        if not t.is_contiguous():
            raise ValueError('%s: not contiguous' % name)
        if not t.is_cuda:
            raise ValueError('%s: on %s, expected a HIP device tensor' % (name, t.device))

    def state(self, fill=None):
        t = self.torch.empty((5, self.nIC), dtype=self.tdtype, device='cuda')
        if fill is not None:
            t.fill_(fill)
        return t

    # ------------------------------------------------------------ entries --
    def tape_bytes(self):
        return int(self.lib.adfvm_tape_bytes(self.handle))

    def new_tapes(self, n):
        """n step tapes as one device tensor of shape (n, tape scalars)."""
        k = self.tape_bytes() // self.precision.itemsize
        return self.torch.empty((n, k), dtype=self.tdtype, device='cuda')

    def set_segments(self, cell_starts):
        """Ensemble of meshes in this context: segment k owns cells
        [cell_starts[k], cell_starts[k+1]); objectives then come per segment
        (obj_out needs len(cell_starts)-1 doubles)."""
        a = np.ascontiguousarray(cell_starts, np.int64)
        self._chk(self.lib.adfvm_set_segments(self.handle, len(a) - 1, _lib.ptr(a, C.c_int64)))
        self.nSeg = len(a) - 1

    OBJECTIVE_STATES = {'stage1': _lib.OBJSTATE_STAGE1, 'start': _lib.OBJSTATE_START}

    @property
    def objective_state(self):
        """The state each step's objective is evaluated on: 'stage1' (the
        reference's hook, the first intermediate RK stage, density.py:412-413;
        the default) or 'start' (the step's start state u_n)."""
        code = self.lib.adfvm_objective_state(self.handle)
        return {v: k for k, v in self.OBJECTIVE_STATES.items()}[code]

    @objective_state.setter
    def objective_state(self, which):
        if which not in self.OBJECTIVE_STATES:
            raise ValueError('objective state %r (stage1, start)' % (which,))
        self._chk(self.lib.adfvm_set_objective_state(self.handle, self.OBJECTIVE_STATES[which]))

    def _check_out64(self, t, n, name):
        if t is None:
            return
        if t.dtype != self.torch.float64 or not t.is_contiguous() or t.numel() < n or not t.is_cuda:
            raise ValueError('%s: needs a contiguous float64 HIP-device tensor of at least %d entries' % (name, n))

    def _check_tape(self, tape):
        if tape is None:
            return
        if not tape.is_contiguous() or not tape.is_cuda or tape.dtype != self.tdtype or \
                tape.numel() * tape.element_size() < self.tape_bytes():
            raise ValueError('tape: needs a contiguous %s HIP-device tensor of %d bytes' % (self.tdtype, self.tape_bytes()))

    def primal_step(self, u, dt, source=None, dtc_out=None, obj_out=None, tape=None, stream=None):
        self._check(u, (5, self.nIC), 'u')
        self._check_out64(obj_out, getattr(self, 'nSeg', 1), 'obj_out')
        self._check_out64(dtc_out, 1, 'dtc_out')
        if source is not None:
            self._check(source, (5, self.nIC), 'source')
        self._check_tape(tape)
        self._chk(self.lib.adfvm_primal_step(self.handle, _dptr(u), float(dt), _dptr(source), _dptr(dtc_out),
                                             _dptr(obj_out), _dptr(tape), _stream(stream)))

    def adjoint_step(self, u_n, dt, ua, obja=1.0, source=None, source_grad=None, bc_grad=None, tape=None,
                     stream=None, obj_grad=None):
        """bc_grad (device (5, nBoundaryFaces), from bc_state()): the
        gradient with respect to the boundary-condition values is added to
        it; bc_param() locates one ('BCs', field, patch, key) parameter in it.
        obj_grad (device (5, nIC)): a caller objective's dJ/du at the
        objective stage (stage_state()), added with weight obja."""
        if obj_grad is not None:
            self._check(obj_grad, (5, self.nIC), 'obj_grad')
        self._check(u_n, (5, self.nIC), 'u_n')
        self._check(ua, (5, self.nIC), 'ua')
        for a, nm in ((source, 'source'), (source_grad, 'source_grad')):
            if a is not None:
                self._check(a, (5, self.nIC), nm)
        if bc_grad is not None:
            self._check(bc_grad, (5, self.nBF), 'bc_grad')
        self._check_tape(tape)
        if obj_grad is not None:
            self._chk(self.lib.adfvm_adjoint_step_cotangent(self.handle, _dptr(u_n), float(dt), _dptr(source),
                                                            _dptr(ua), float(obja), _dptr(obj_grad),
                                                            _dptr(source_grad), _dptr(bc_grad), _dptr(tape),
                                                            _stream(stream)))
            return
        self._chk(self.lib.adfvm_adjoint_step(self.handle, _dptr(u_n), float(dt), _dptr(source), _dptr(ua),
                                              float(obja), _dptr(source_grad), _dptr(bc_grad), _dptr(tape),
                                              _stream(stream)))

    # the reference's differentiable mesh arrays, in order (Mesh.gradFields, mesh.py:28-32)
    MESH_GRAD_FIELDS = ('areas', 'volumesL', 'volumesR', 'weights', 'deltas', 'normals', 'deltasUnit',
                        'linearWeights', 'quadraticWeights', 'volumes')

    def mesh_grad_state(self):
        """A zeroed device array for the mesh-geometry gradient (parameter
        'mesh', apps/adjoint.py:107-110) and its registration: every later
        adjoint step adds d(obja*J + <ua, u_{n+1}>)/d(mesh) to it, until
        set_mesh_grad(None).  mesh_grad_arrays() splits it into the
        reference's arrays."""
        n = int(self.lib.adfvm_mesh_grad_size(self.handle))
        g = self.torch.zeros(n, dtype=self.tdtype, device='cuda')
        self.set_mesh_grad(g)
        return g

    def set_mesh_grad(self, g):
        if g is not None:
            self._check(g, (int(self.lib.adfvm_mesh_grad_size(self.handle)),), 'mesh_grad')
        self._mesh_grad = g                  # keeps the array alive while the library holds it
        self._chk(self.lib.adfvm_set_mesh_grad(self.handle, _dptr(g)))

    def mesh_grad_arrays(self, g):
        """Views of a mesh-gradient array as the reference's arrays: areas
        (nF,1), volumesL (nF,1), volumesR (nIF,1), weights (nF,1), deltas
        (nF,1), normals (nF,3), deltasUnit (nF,3), linearWeights (nF,2),
        quadraticWeights (nF,2,3), volumes (nIC,1)."""
        nF, nIF, nIC = self.mesh.nFaces, self.mesh.nInternalFaces, self.nIC
        shapes = [(nF, 1), (nF, 1), (nIF, 1), (nF, 1), (nF, 1), (nF, 3), (nF, 3), (nF, 2), (nF, 2, 3), (nIC, 1)]
        out, o = [], 0
        for sh in shapes:
            n = int(np.prod(sh))
            out.append(g[o:o + n].reshape(sh))
            o += n
        assert o == g.numel()
        return out

    @property
    def n_stages(self):
        return 1 if self.integrator == _lib.RK_EULER else 3

    def stage_state(self, u_n, tape):
        """The state the objective is evaluated on in the step from u_n
        (density.py:412-413): the first intermediate stage, which the step's
        tape holds first (a view into it), or u_n itself for a one-stage
        integrator or with objective_state 'start'."""
        if self.n_stages == 1 or self.objective_state == 'start':
            return u_n
        if tape is None:
            raise ValueError('the objective stage of a multi-stage step is read from its tape')
        self._check_tape(tape)
        return tape.reshape(-1)[:5 * self.nIC].view(5, self.nIC)

    def adjoint_viscosity(self, u_n, ua, dt, vtype, scaling, iterations=1000, m2norm_out=None, stream=None):
        """Adjoint-viscosity smoothing of ua in place after the adjoint step
        from u_n (apps/adjoint.py:127-140, postpro.py:491-719): vtype
        'uniform', 'abarbanel' or 'turkel'; scaling the adjoint's scaling
        (adjParams[0]); `iterations` Jacobi sweeps (the reference: 1000).
        m2norm_out (device (nIC,)): the scaled viscosity per cell."""
        self._check(u_n, (5, self.nIC), 'u_n')
        self._check(ua, (5, self.nIC), 'ua')
        if m2norm_out is not None:
            self._check(m2norm_out, (self.nIC,), 'm2norm_out')
        codes = {'uniform': _lib.VISC_UNIFORM, 'abarbanel': _lib.VISC_ABARBANEL, 'turkel': _lib.VISC_TURKEL}
        if vtype not in codes:
            raise ValueError('adjoint viscosity type %r (uniform, abarbanel, turkel)' % (vtype,))
        self._chk(self.lib.adfvm_adjoint_viscosity(self.handle, _dptr(u_n), _dptr(ua), float(dt), codes[vtype],
                                                   float(scaling), int(iterations), _dptr(m2norm_out),
                                                   _stream(stream)))

    def bc_state(self):
        """Zeroed (5, nBoundaryFaces) device array for adjoint_step's bc_grad."""
        return self.torch.zeros((5, self.nBF), dtype=self.tdtype, device='cuda')

    def _bc_slot(self, field, patch, key):
        """(patch index, ADFVM_BCV_* slot, components) of the value array
        ('BCs', field, patch, key): 'value' of a fixedValue / calculated
        U, T or p; U0, T0, p0 of a CBC_UPT and Tt, pt, direction (when
        given) of a CBC_TOTAL_PT, both on p (BCs.py:111-184)."""
        if patch not in self._bcinfo:
            raise KeyError('no patch %r' % patch)
        info = self._bcinfo[patch]
        slot = None
        if info['cbc'] == _lib.CBC_UPT:
            if field == 'p':
                slot = {'U0': _lib.BCV_U, 'T0': _lib.BCV_T, 'p0': _lib.BCV_P}.get(key)
        elif info['cbc'] == _lib.CBC_TOTAL_PT:
            if field == 'p':
                slot = {'Tt': _lib.BCV_TT, 'pt': _lib.BCV_PT}.get(key)
                if key == 'direction' and info['direction']:
                    slot = _lib.BCV_DIRECTION
        elif info['types'].get(field) in ('fixedValue', 'calculated') and key == 'value':
            slot = {'U': _lib.BCV_U, 'T': _lib.BCV_T, 'p': _lib.BCV_P}[field]
        if slot is None:
            raise ValueError('(%r, %r, %r) is not a boundary-condition value of this case' % (field, patch, key))
        return info['index'], slot, 3 if slot in (_lib.BCV_U, _lib.BCV_DIRECTION) else 1

    def bc_param(self, field, patch, key):
        """(rows, faces) slices of a bc_grad array that hold the parameter
        ('BCs', field, patch, key) of apps/adjoint.py:111-118: the gradient
        with respect to that value array is bc_grad[rows, faces].T, shaped
        (nFaces, components) like the value.  A 'calculated' value sets no
        ghost after the first step and has no gradient rows."""
        _, slot, _ = self._bc_slot(field, patch, key)
        info = self._bcinfo[patch]
        if info['cbc'] == _lib.CBC_NONE and info['types'][field] != 'fixedValue':
            raise ValueError('(%r, %r, %r) is not a fixedValue' % (field, patch, key))
        rows = {_lib.BCV_U: slice(0, 3), _lib.BCV_T: slice(3, 4), _lib.BCV_P: slice(4, 5), _lib.BCV_TT: slice(3, 4),
                _lib.BCV_PT: slice(4, 5), _lib.BCV_DIRECTION: slice(0, 3)}[slot]
        return rows, slice(info['start'], info['start'] + info['n'])

    def set_bc_value(self, field, patch, key, value, stream=None):
        """Replace the value array ('BCs', field, patch, key) on the device
        (value: (nFaces, components) or broadcastable to it)."""
        index, slot, k = self._bc_slot(field, patch, key)
        n = self._bcinfo[patch]['n']
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, np.float64).reshape(-1, k), (n, k)))
        self._chk(self.lib.adfvm_set_bc_values(self.handle, index, slot, C.c_void_p(a.ctypes.data), a.size,
                                               _stream(stream)))

    def gradients(self, u, out=None, stream=None):
        self._check(u, (5, self.nIC), 'u')
        out = out if out is not None else self.torch.empty((15, self.nIC), dtype=self.tdtype, device='cuda')
        self._check(out, (15, self.nIC), 'out')
        self._chk(self.lib.adfvm_gradients(self.handle, _dptr(u), _dptr(out), _stream(stream)))
        return out

    def residual(self, u, out=None, dtc_cell=None, stream=None):
        self._check(u, (5, self.nIC), 'u')
        out = out if out is not None else self.torch.empty((5, self.nIC), dtype=self.tdtype, device='cuda')
        self._check(out, (5, self.nIC), 'out')
        if dtc_cell is not None:
            self._check(dtc_cell, (self.nIC,), 'dtc_cell')
        self._chk(self.lib.adfvm_residual(self.handle, _dptr(u), _dptr(out), _dptr(dtc_cell), _stream(stream)))
        return out

    def residual_vjp(self, u, rbar, out=None, stream=None):
        self._check(u, (5, self.nIC), 'u')
        self._check(rbar, (5, self.nIC), 'rbar')
        out = out if out is not None else self.torch.empty((5, self.nIC), dtype=self.tdtype, device='cuda')
        self._check(out, (5, self.nIC), 'out')
        self._chk(self.lib.adfvm_residual_vjp(self.handle, _dptr(u), _dptr(rbar), _dptr(out), _stream(stream)))
        return out

    def objective(self, u, out=None, stream=None):
        self._check(u, (5, self.nIC), 'u')
        n = getattr(self, 'nSeg', 1)
        out = out if out is not None else self.torch.zeros(n, dtype=self.torch.float64, device='cuda')
        self._check_out64(out, n, 'out')
        self._chk(self.lib.adfvm_objective(self.handle, _dptr(u), _dptr(out), _stream(stream)))
        return out

    def primitive_ghosts(self, u, out=None, stream=None):
        self._check(u, (5, self.nIC), 'u')
        nC = self.mesh.nCells
        out = out if out is not None else self.torch.empty((5, nC), dtype=self.tdtype, device='cuda')
        self._check(out, (5, nC), 'out')
        self._chk(self.lib.adfvm_primitive_ghosts(self.handle, _dptr(u), _dptr(out), _stream(stream)))
        return out

    def primitive_ghosts_vjp(self, u, prim_bar, out=None, bc_grad=None, stream=None):
        """(d prim/d u)^T prim_bar for prim = primitive_ghosts(u): prim_bar
        (5, nCells) -> (5, nIC); the boundary-value part is added to bc_grad
        (5, nBoundaryFaces) when given."""
        self._check(u, (5, self.nIC), 'u')
        self._check(prim_bar, (5, self.mesh.nCells), 'prim_bar')
        out = out if out is not None else self.torch.empty((5, self.nIC), dtype=self.tdtype, device='cuda')
        self._check(out, (5, self.nIC), 'out')
        if bc_grad is not None:
            self._check(bc_grad, (5, self.nBF), 'bc_grad')
        self._chk(self.lib.adfvm_primitive_ghosts_vjp(self.handle, _dptr(u), _dptr(prim_bar), _dptr(out),
                                                      _dptr(bc_grad), _stream(stream)))
        return out

    def grad_cell(self, phi, stream=None):
        """Gauss gradient of phi (k, nCells: ghost values included) ->
        (3k, nInternalCells), rows d(phi_i)/dx_j at 3i+j."""
        if getattr(phi, 'ndim', 0) != 2:
            raise ValueError('phi: expected (k, nCells)')
        k = phi.shape[0]
        self._check(phi, (k, self.mesh.nCells), 'phi')
        out = self.torch.empty((3 * k, self.nIC), dtype=self.tdtype, device='cuda')
        self._chk(self.lib.adfvm_grad_cell(self.handle, _dptr(phi), k, _dptr(out), _stream(stream)))
        return out

    def interp_second_order(self, phi, grad, swap, stream=None):
        """interp.secondOrder on every face: phi (k, nCells) and grad
        (3k, nCells) -- gradients on every cell, ghosts included (grad_cell
        returns internal rows only; the caller supplies the ghost rows) ->
        (k, nFaces)."""
        if getattr(phi, 'ndim', 0) != 2:
            raise ValueError('phi: expected (k, nCells)')
        k = phi.shape[0]
        self._check(phi, (k, self.mesh.nCells), 'phi')
        self._check(grad, (3 * k, self.mesh.nCells), 'grad')
        if swap not in (0, 1):
            raise ValueError('swap must be 0 or 1')
        out = self.torch.empty((k, self.mesh.nFaces), dtype=self.tdtype, device='cuda')
        self._chk(self.lib.adfvm_interp_second_order(self.handle, _dptr(phi), _dptr(grad), k, swap, _dptr(out),
                                                     _stream(stream)))
        return out

    def dot(self, a, b, out=None, stream=None):
        """sum_j a[j] b[j] in FP64 (a, b: contiguous device tensors of the
        context precision with the same number of elements)."""
        if not hasattr(a, 'numel') or not hasattr(b, 'numel'):
            raise ValueError('a, b: expected torch tensors')
        self._check(a, tuple(a.shape), 'a')
        self._check(b, tuple(b.shape), 'b')
        if b.numel() != a.numel():
            raise ValueError('a, b: %d and %d elements' % (a.numel(), b.numel()))
        out = out if out is not None else self.torch.zeros(1, dtype=self.torch.float64, device='cuda')
        self._check_out64(out, 1, 'out')
        self._chk(self.lib.adfvm_dot(self.handle, _dptr(a), _dptr(b), a.numel(), _dptr(out), _stream(stream)))
        return out

    def adjoint_energy(self, ua, u, out=None, stream=None):
        self._check(ua, (5, self.nIC), 'ua')
        self._check(u, (5, self.nIC), 'u')
        out = out if out is not None else self.torch.zeros(1, dtype=self.torch.float64, device='cuda')
        self._check_out64(out, 1, 'out')
        self._chk(self.lib.adfvm_adjoint_energy(self.handle, _dptr(ua), _dptr(u), _dptr(out), _stream(stream)))
        return out

    def profile_start(self):
        self._chk(self.lib.adfvm_profile_reset(self.handle))
        self._chk(self.lib.adfvm_profile(self.handle, 1, _stream(None)))

    def profile_stop(self):
        """Returns {kernel name: (total ms, launches)} recorded since
        profile_start (HIP events on the launch stream)."""
        self._chk(self.lib.adfvm_profile(self.handle, 0, _stream(None)))
        out = {}
        for i in range(self.lib.adfvm_profile_count(self.handle)):
            buf = C.create_string_buffer(64)
            ms, n = C.c_double(), C.c_int64()
            self._chk(self.lib.adfvm_profile_entry(self.handle, i, buf, 64, C.byref(ms), C.byref(n)))
            out[buf.value.decode()] = (ms.value, n.value)
        return out

    def autotune(self, u, reps=5, stream=None):
        """Pick each stencil kernel's block size by timing it on state u;
        returns {kernel: (block, ms per launch)}."""
        self._check(u, (5, self.nIC))
        self._chk(self.lib.adfvm_autotune(self.handle, _dptr(u), int(reps), _stream(stream)))
        return self.kernel_config()

    def kernel_config(self):
        out = {}
        for i in range(6):
            buf = C.create_string_buffer(64)
            b, ms = C.c_int32(), C.c_double()
            self._chk(self.lib.adfvm_kernel_config(self.handle, i, buf, 64, C.byref(b), C.byref(ms)))
            out[buf.value.decode()] = (b.value, ms.value)
        return out

    def set_kernel_block(self, name, block):
        self._chk(self.lib.adfvm_set_kernel_block(self.handle, name.encode(), int(block)))

    def set_streaming(self, mode):
        """Non-temporal loads of single-use rows: True, False, or None for
        the automatic choice (on when a stage's working set exceeds the
        256 MB Infinity Cache).  Results do not depend on it."""
        self._chk(self.lib.adfvm_set_streaming(self.handle, -1 if mode is None else int(bool(mode))))

    @property
    def streaming(self):
        return bool(self.lib.adfvm_streaming(self.handle))

    def pack(self, rho, rhoU, rhoE, out=None, stream=None):
        self._check(rho, (self.nIC, 1), 'rho')
        self._check(rhoU, (self.nIC, 3), 'rhoU')
        self._check(rhoE, (self.nIC, 1), 'rhoE')
        out = out if out is not None else self.state()
        self._check(out, (5, self.nIC), 'out')
        self._chk(self.lib.adfvm_pack_state(self.handle, _dptr(rho), _dptr(rhoU), _dptr(rhoE), _dptr(out),
                                            _stream(stream)))
        return out

    def unpack(self, u, stream=None):
        self._check(u, (5, self.nIC), 'u')
        t = self.torch
        rho = t.empty((self.nIC, 1), dtype=self.tdtype, device='cuda')
        rhoU = t.empty((self.nIC, 3), dtype=self.tdtype, device='cuda')
        rhoE = t.empty((self.nIC, 1), dtype=self.tdtype, device='cuda')
        self._chk(self.lib.adfvm_unpack_state(self.handle, _dptr(u), _dptr(rho), _dptr(rhoU), _dptr(rhoE),
                                              _stream(stream)))
        return rho, rhoU, rhoE

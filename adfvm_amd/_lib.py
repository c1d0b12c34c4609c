"""ctypes binding of libadfvm.so (declarations in include/adfvm.h).

The library is built in-tree by `adfvm_amd.build` (hipcc for gfx950).  There
is no fallback: if the library is missing or fails to load, every entry
point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libadfvm.so')

ADFVM_OK = 0
PATCH_OTHER, PATCH_CYCLIC, PATCH_PROCESSOR = 0, 1, 2
BC_CYCLIC, BC_ZERO_GRADIENT, BC_SYMMETRY, BC_FIXED_VALUE, BC_CALCULATED = 0, 1, 2, 3, 4
FLUX_COUPLED, FLUX_CHARACTERISTIC, FLUX_BOUNDARY = 0, 1, 2
CBC_NONE, CBC_UPT, CBC_TOTAL_PT = 0, 1, 2
BCV_U, BCV_T, BCV_P, BCV_TT, BCV_PT, BCV_DIRECTION = 0, 1, 2, 3, 4, 5
VISC_UNIFORM, VISC_ABARBANEL, VISC_TURKEL = 0, 1, 2
RIEMANN_ROE, RIEMANN_LF = 0, 1
MU_CONSTANT, MU_SUTHERLAND = 0, 1
RK_SSPRK3, RK_EULER = 0, 1
OBJ_NONE, OBJ_KINETIC_ENERGY, OBJ_DRAG = 0, 1, 2
OBJSTATE_STAGE1, OBJSTATE_START = 0, 1

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class MeshInput(C.Structure):
    _fields_ = [('nPoints', C.c_int32), ('nFaces', C.c_int32), ('nInternalFaces', C.c_int32),
                ('nPatches', C.c_int32), ('points', f64p), ('faces', i32p), ('owner', i32p),
                ('neighbour', i32p), ('patchStartFace', i32p), ('patchNFaces', i32p),
                ('patchType', i32p), ('patchNeighbour', i32p)]


class MeshSizes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ('nInternalFaces', 'nFaces', 'nBoundaryFaces',
                                         'nInternalCells', 'nGhostCells', 'nCells')]


class MeshOutput(C.Structure):
    _fields_ = [('cellFaces', i32p), ('cells', i32p), ('neighbour', i32p), ('cellOwner', i32p),
                ('cellNeighbours', i32p), ('cellNeighboursMatOp', i32p),
                ('normals', f64p), ('faceCentres', f64p), ('areas', f64p), ('cellCentres', f64p),
                ('volumes', f64p), ('deltas', f64p), ('deltasUnit', f64p), ('weights', f64p),
                ('linearWeights', f64p), ('quadraticWeights', f64p), ('volumesL', f64p),
                ('volumesR', f64p), ('nLocalCells', i32p)]


class PatchDesc(C.Structure):
    _fields_ = [('startFace', C.c_int32), ('nFaces', C.c_int32), ('fluxKind', C.c_int32),
                ('neighbourPatch', C.c_int32), ('bcU', C.c_int32), ('bcT', C.c_int32),
                ('bcp', C.c_int32), ('cbc', C.c_int32),
                ('valueU', f64p), ('valueT', f64p), ('valuep', f64p),
                ('Tt', f64p), ('pt', f64p), ('direction', f64p)]


class Physics(C.Structure):
    _fields_ = [('precision', C.c_int32), ('gamma', C.c_double), ('Cp', C.c_double), ('Pr', C.c_double),
                ('muModel', C.c_int32), ('mu0', C.c_double), ('muA', C.c_double), ('muTs', C.c_double),
                ('riemann', C.c_int32), ('boundaryRiemann', C.c_int32), ('small', C.c_double),
                ('integrator', C.c_int32), ('objective', C.c_int32), ('objectivePatch', C.c_int32),
                ('limiter', C.c_int32)]


class SolverDesc(C.Structure):
    _fields_ = [('sizes', MeshSizes), ('owner', i32p), ('neighbour', i32p), ('cellFaces', i32p),
                ('cellOwner', i32p), ('cellNeighbours', i32p), ('areas', f64p), ('normals', f64p),
                ('weights', f64p), ('deltas', f64p), ('deltasUnit', f64p), ('linearWeights', f64p),
                ('quadraticWeights', f64p), ('volumes', f64p), ('nPatches', C.c_int32),
                ('patches', C.POINTER(PatchDesc))]


vp = C.c_void_p
# name -> (restype, argtypes); every symbol include/adfvm.h declares
SIGNATURES = {
    'adfvm_mesh_sizes': (C.c_int, [C.POINTER(MeshInput), C.POINTER(MeshSizes)]),
    'adfvm_mesh_build': (C.c_int, [C.POINTER(MeshInput), C.POINTER(MeshOutput)]),
    'adfvm_mesh_last_error': (C.c_char_p, []),
    'adfvm_ctx_create': (C.c_int, [C.POINTER(SolverDesc), C.POINTER(Physics), C.POINTER(vp)]),
    'adfvm_ctx_destroy': (None, [vp]),
    'adfvm_last_error': (C.c_char_p, [vp]),
    'adfvm_ctx_precision': (C.c_int, [vp]),
    'adfvm_set_segments': (C.c_int, [vp, C.c_int32, C.POINTER(C.c_int64)]),
    'adfvm_segments': (C.c_int, [vp]),
    'adfvm_set_objective_state': (C.c_int, [vp, C.c_int32]),
    'adfvm_objective_state': (C.c_int, [vp]),
    'adfvm_autotune': (C.c_int, [vp, vp, C.c_int32, vp]),
    'adfvm_kernel_config': (C.c_int, [vp, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_double)]),
    'adfvm_set_kernel_block': (C.c_int, [vp, C.c_char_p, C.c_int32]),
    'adfvm_set_streaming': (C.c_int, [vp, C.c_int32]),
    'adfvm_streaming': (C.c_int, [vp]),
    'adfvm_primal_step': (C.c_int, [vp, vp, C.c_double, vp, vp, vp, vp, vp]),
    'adfvm_tape_bytes': (C.c_int64, [vp]),
    'adfvm_adjoint_step': (C.c_int, [vp, vp, C.c_double, vp, vp, C.c_double, vp, vp, vp, vp]),
    'adfvm_adjoint_step_cotangent': (C.c_int, [vp, vp, C.c_double, vp, vp, C.c_double, vp, vp, vp, vp, vp]),
    'adfvm_primitive_ghosts_vjp': (C.c_int, [vp, vp, vp, vp, vp, vp]),
    'adfvm_set_bc_values': (C.c_int, [vp, C.c_int32, C.c_int32, vp, C.c_int64, vp]),
    'adfvm_adjoint_viscosity': (C.c_int, [vp, vp, vp, C.c_double, C.c_int32, C.c_double, C.c_int32, vp, vp]),
    'adfvm_gradients': (C.c_int, [vp, vp, vp, vp]),
    'adfvm_residual': (C.c_int, [vp, vp, vp, vp, vp]),
    'adfvm_residual_vjp': (C.c_int, [vp, vp, vp, vp, vp]),
    'adfvm_objective': (C.c_int, [vp, vp, vp, vp]),
    'adfvm_primitive_ghosts': (C.c_int, [vp, vp, vp, vp]),
    'adfvm_grad_cell': (C.c_int, [vp, vp, C.c_int32, vp, vp]),
    'adfvm_interp_second_order': (C.c_int, [vp, vp, vp, C.c_int32, C.c_int32, vp, vp]),
    'adfvm_dot': (C.c_int, [vp, vp, vp, C.c_int64, vp, vp]),
    'adfvm_adjoint_energy': (C.c_int, [vp, vp, vp, vp, vp]),
    'adfvm_dot_noalloc': (C.c_int, [C.c_int32, vp, vp, C.c_int64, vp, vp]),
    'adfvm_noalloc_last_error': (C.c_char_p, []),
    'adfvm_mesh_grad_size': (C.c_int64, [vp]),
    'adfvm_set_mesh_grad': (C.c_int, [vp, vp]),
    'adfvm_profile': (C.c_int, [vp, C.c_int32, vp]),
    'adfvm_profile_count': (C.c_int, [vp]),
    'adfvm_profile_entry': (C.c_int, [vp, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int64)]),
    'adfvm_profile_reset': (C.c_int, [vp]),
    'adfvm_source_hash': (C.c_char_p, []),
    'adfvm_pack_state': (C.c_int, [vp, vp, vp, vp, vp, vp]),
    'adfvm_unpack_state': (C.c_int, [vp, vp, vp, vp, vp, vp]),
}

_lib = None


class LibraryError(RuntimeError):
    pass


def load():
    """Load libadfvm.so (once).  Raises if it is absent, or if it was built
    from sources other than the csrc/ and include/ files next to it (the
    SHA-256 the build embeds): there is no non-native path, and no stale
    one."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryError('libadfvm.so not built: run `python -c "import __graft_entry__ as g; g.build()"` '
                           'or `python -m adfvm_amd.build` (expected at %s)' % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    from .build import source_hash
    if not hasattr(lib, 'adfvm_source_hash'):
        raise LibraryError('%s predates the source check: rebuild it (python -m adfvm_amd.build)' % LIB_PATH)
    lib.adfvm_source_hash.restype = C.c_char_p
    built, now = lib.adfvm_source_hash().decode(), source_hash()
    if built != now:
        raise LibraryError('%s was built from other sources (hash %s..., the sources give %s...): rebuild it '
                           '(python -m adfvm_amd.build)' % (LIB_PATH, built[:12], now[:12]))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a, ctype):
    """ctypes pointer into a C-contiguous numpy array (or NULL for None)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(C.POINTER(ctype))


def check(rc, ctx=None):
    if rc != ADFVM_OK:
        lib = load()
        if ctx is not None:
            msg = lib.adfvm_last_error(ctx)
        else:
            msg = lib.adfvm_mesh_last_error()
        raise LibraryError('libadfvm error %d: %s' % (rc, msg.decode() if msg else ''))

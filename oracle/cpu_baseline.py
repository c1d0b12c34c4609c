"""ORACLE-SIDE CPU BASELINE (test and measurement infrastructure only):
build and ctypes binding of oracle/cpu_step.cpp, the C++/OpenMP FP64 step
and adjoint step timed as the CPU path beside the GPU.  Only tests/ and
bench.py's cpu_baseline leg may use it; the product never does.

    python -m oracle.cpu_baseline        # builds oracle/_cpu/libadfvm_cpu.so
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'cpu_step.cpp')
OUT_DIR = os.path.join(HERE, '_cpu')
LIB = os.path.join(OUT_DIR, 'libadfvm_cpu.so')
# x86-64-v3 (AVX2, FMA), not -march=native: the library is built on one host
# and timed on another (the MI355X machine's host CPU)
FLAGS = ['-O3', '-std=c++17', '-fPIC', '-shared', '-fopenmp', '-march=x86-64-v3', '-fno-math-errno',
         '-I', os.path.join(ROOT, 'adfvm_amd', 'csrc')]


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + '.tmp'
    cmd = ['g++'] + FLAGS + [SRC, '-o', tmp]
    print(' '.join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError('%s not built (python -m oracle.cpu_baseline)' % LIB)
        lib = C.CDLL(LIB)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        lib.adfvm_cpu_create.restype = vp
        lib.adfvm_cpu_create.argtypes = [C.c_long, C.c_long, C.c_long, ip, ip, ip, ip, ip] + [dp] * 8 + \
            [C.c_double] * 5 + [C.c_int]
        lib.adfvm_cpu_destroy.argtypes = [vp]
        lib.adfvm_cpu_step.argtypes = [vp, dp, C.c_double, dp]
        lib.adfvm_cpu_adjoint_step.argtypes = [vp, dp, C.c_double, dp, C.c_double]
        lib.adfvm_cpu_threads.restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class CpuStep(object):
    """FP64 step / adjoint step on a periodic mesh (adfvm_amd.mesh.Mesh whose
    patches are all cyclic), constant viscosity mu0, the kinetic-energy
    objective on the stage-1 state."""

    def __init__(self, mesh, mu0, gamma=1.4, Cp=1004.5, Pr=0.7, small=1e-30, riemann=0):
        lib = load()
        m = mesh
        nIF, nF = m.nInternalFaces, m.nFaces
        pair = np.full(nF, -1, np.int32)
        for name, b in m.boundary.items():
            if b['type'] != 'cyclic':
                raise ValueError('the CPU baseline takes periodic meshes (patch %r is %s)' % (name, b['type']))
            nb = m.boundary[b['neighbourPatch']]
            pair[b['startFace']:b['startFace'] + b['nFaces']] = np.arange(nb['startFace'], nb['startFace'] + b['nFaces'])
        i32 = lambda a: np.ascontiguousarray(a, np.int32)
        f64 = lambda a: np.ascontiguousarray(a, np.float64)
        self._keep = [i32(m.owner), i32(m.neighbour), i32(m.cellFaces), i32(m.cellOwner), pair,
                      f64(m.areas), f64(m.normals), f64(m.weights), f64(m.deltas), f64(m.deltasUnit),
                      f64(m.linearWeights), f64(m.quadraticWeights), f64(m.volumes)]
        k = self._keep
        self.h = lib.adfvm_cpu_create(m.nInternalCells, nIF, nF, *[_p(a, C.c_int) for a in k[:5]],
                                      *[_p(a, C.c_double) for a in k[5:]], gamma, Cp, Pr, float(mu0), small,
                                      int(riemann))
        if not self.h:
            raise ValueError('adfvm_cpu_create failed (a boundary face without a cyclic pair)')
        self.lib = lib
        self.nIC = m.nInternalCells

    def __del__(self):
        if getattr(self, 'h', None):
            self.lib.adfvm_cpu_destroy(self.h)
            self.h = None

    @staticmethod
    def threads():
        return int(load().adfvm_cpu_threads())

    def step(self, u, dt):
        """u (5, nIC) float64 C-contiguous, updated in place; returns the
        stage-1 kinetic energy."""
        assert u.dtype == np.float64 and u.flags['C_CONTIGUOUS'] and u.shape == (5, self.nIC)
        obj = C.c_double()
        self.lib.adfvm_cpu_step(self.h, _p(u, C.c_double), float(dt), C.byref(obj))
        return obj.value

    def adjoint_step(self, u_n, dt, ua, obja=1.0):
        assert u_n.dtype == np.float64 and u_n.flags['C_CONTIGUOUS'] and u_n.shape == (5, self.nIC)
        assert ua.dtype == np.float64 and ua.flags['C_CONTIGUOUS'] and ua.shape == (5, self.nIC)
        self.lib.adfvm_cpu_adjoint_step(self.h, _p(u_n, C.c_double), float(dt), _p(ua, C.c_double), float(obja))


if __name__ == '__main__':
    build()

"""Build libadfvm.so in-tree: hipcc for the gfx950 kernels, g++ for the
host mesh builder (strict FP evaluation order), linked into one shared
library next to this file.

    python -m adfvm_amd.build
"""
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
BUILD_DIR = os.path.join(ROOT, '_build')
LIB = os.path.join(HERE, 'libadfvm.so')
ARCH = 'gfx950'

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
# packed FP32 instructions are switched off for all device code; multiply-add
# contraction is decided per source expression (-ffp-contract=on), not by the
# backend, so results do not depend on launch bounds or instruction scheduling.
# FP32 division and square root use the hardware reciprocal and square root
# (about 1 ulp) instead of the correctly rounded expansions, which are a fifth
# of the face kernels' instructions; FP64 keeps IEEE division.
HIP_FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH, '-ffp-contract=on',
             '-fno-hip-fp32-correctly-rounded-divide-sqrt',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-munsafe-fp-atomics', '-I', INCLUDE, '-I', CSRC]
CXX_FLAGS = ['-O2', '-std=c++17', '-fPIC', '-ffp-contract=off', '-I', INCLUDE]


def source_hash():
    """SHA-256 over the library's sources (every file of csrc/ and the C
    header), embedded in the build and checked by the loader, so that a
    library built from other sources is never loaded."""
    h = hashlib.sha256()
    files = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC))] + [os.path.join(INCLUDE, 'adfvm.h')]
    for path in files:
        if os.path.isfile(path):
            h.update(os.path.basename(path).encode() + b'\0')
            with open(path, 'rb') as f:
                h.update(f.read())
            h.update(b'\0')
    return h.hexdigest()


def run(cmd):
    print(' '.join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    mesh_obj = os.path.join(BUILD_DIR, 'mesh_build.o')
    run(['g++'] + CXX_FLAGS + ['-c', os.path.join(CSRC, 'mesh_build.cpp'), '-o', mesh_obj])
    objs.append(mesh_obj)
    hash_def = '-DADFVM_SOURCE_HASH="%s"' % source_hash()
    for src in sorted(os.listdir(CSRC)):
        if src.endswith('.hip'):
            obj = os.path.join(BUILD_DIR, src[:-4] + '.o')
            run([HIPCC] + HIP_FLAGS + [hash_def, '-c', os.path.join(CSRC, src), '-o', obj])
            objs.append(obj)
    tmp = LIB + '.tmp'
    run([HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)

"""Precision, constants, patch groups and the command-line flags of adFVM's
config module (adFVM/config.py), without importing anything at argument-parse
time so that library use (tests, bench) is unaffected by sys.argv.

Precision follows the reference's GPU rule (adFVM/config.py:52-67): kernels
run in float32 unless `--gpu_double` is given, in which case float64.
"""
import argparse
import os
import sys
import time

import numpy as np

runtime = time.time()

# patch groups, adFVM/config.py:387-391
processorPatches = ['processor', 'processorCyclic']
cyclicPatches = ['cyclic', 'slidingPeriodic1D']
coupledPatches = cyclicPatches + processorPatches
defaultPatches = coupledPatches + ['symmetryPlane', 'empty']

gpu = True
precision = np.float32
SMALL = VSMALL = LARGE = None
fileFormat = 'binary'


def set_precision(p):
    """float32 or float64; sets SMALL/VSMALL/LARGE as adFVM/config.py:170-178."""
    global precision, SMALL, VSMALL, LARGE
    precision = np.dtype(p).type
    if precision == np.float64:
        SMALL, VSMALL, LARGE = 1e-30, 1e-300, 1e300
    elif precision == np.float32:
        SMALL, VSMALL, LARGE = 1e-9, 1e-30, 1e30
    else:
        raise ValueError('precision must be float32 or float64')


set_precision(np.float32)


def small_for(p):
    return 1e-30 if np.dtype(p) == np.float64 else 1e-9


def parse_user_args(argv=None):
    """The reference's global flags (adFVM/config.py:17-32); returns
    (user, remaining).  `--gpu_double` selects float64 kernels."""
    parser = argparse.ArgumentParser(add_help=False)
    parser.add_argument('-g', '--gpu', nargs='?', const=0, type=int, default=0, dest='use_gpu')
    parser.add_argument('--gpu_double', action='store_true', dest='use_gpu_double')
    parser.add_argument('-c', '--compile', action='store_true')
    parser.add_argument('-e', '--compile_exit', action='store_true')
    parser.add_argument('-l', '--no_compile', action='store_true')
    user, rest = parser.parse_known_args(sys.argv[1:] if argv is None else argv)
    set_precision(np.float64 if user.use_gpu_double else np.float32)
    return user, rest


def importModule(localsDict, module):
    """Case files are Python modules whose public globals become the app's
    (adFVM/config.py:148-156).  The file is executed afresh on every call, so
    that several runs in one process each get their own solver objects."""
    import importlib.util
    caseDir, caseFile = os.path.split(module)
    sys.path.append(os.path.abspath(caseDir))
    path = module if module.endswith('.py') else module + '.py'
    spec = importlib.util.spec_from_file_location(caseFile.split('.')[0], path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    for attr in dir(mod):
        if not attr.startswith('_'):
            localsDict[attr] = getattr(mod, attr)


def timeFunction(string):
    """Wall-time decorator (adFVM/config.py:158-168), single rank."""
    def decorator(function):
        def wrapper(*args, **kwargs):
            start = time.time()
            out = function(*args, **kwargs)
            print(string + ':', time.time() - start)
            return out
        return wrapper
    return decorator

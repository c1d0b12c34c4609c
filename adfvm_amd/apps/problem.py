"""Primal runs of a case file: apps/problem.py of the reference.

    python -m adfvm_amd.apps.problem <case file> [orig|perturb|source] [--gpu_double]

orig     run nSteps from startTime, write fields every writeInterval, log
         the time steps to <nSteps>.<writeInterval>.txt, append
         'orig <sim> <info> <objective average> <noise>' to objective.txt
perturb  replay the logged time steps with each perturbation added to the
         parameter (the source term) and append the change of the averaged
         objective ('perturb ...' line, problem.py:47-66)
source   as orig with the perturbation applied (problem.py:98-118)

A status file allows a killed orig run to continue from its last write.
"""
import argparse
import os
import sys

import numpy as np

from .. import config

DEFAULTS = dict(
    source=lambda *args: [0.] * len(args[0]),
    perturb=[],
    reportInterval=1,
    sampleInterval=1,
    viscousInterval=1,
    parameters=[],
    adjParams=[0.0, None, None],
    avgStart=0,
    runCheckpoints=2 ** 30,
)


def load(argv=None):
    """Parse the global flags, import the case file, check its intervals
    (problem.py:21-45).  Returns (namespace dict, remaining args)."""
    user, rest = config.parse_user_args(argv)
    parser = argparse.ArgumentParser()
    parser.add_argument('caseFile')
    args, rest = parser.parse_known_args(rest)
    ns = dict(DEFAULTS)
    sys.path.append(os.getcwd())
    config.importModule(ns, args.caseFile)
    nSteps, writeInterval = ns['nSteps'], ns['writeInterval']
    for name in ('reportInterval', 'sampleInterval', 'viscousInterval'):
        k = ns[name]
        if not (nSteps >= writeInterval >= k and writeInterval % k == 0):
            raise ValueError('need nSteps >= writeInterval >= %s and writeInterval %% %s == 0' % (name, name))
    if not isinstance(ns['parameters'], list):
        ns['parameters'] = [ns['parameters']]
    if not isinstance(ns['perturb'], list):
        ns['perturb'] = [ns['perturb']]
    ns['nPerturb'] = len(ns['perturb'])
    primal = ns['primal']
    primal.timeStepFile = primal.mesh.case + '{0}.{1}.txt'.format(nSteps, writeInterval)
    return ns, rest


def writeResult(ns, option, result, info='-', timeSeriesFile=None):
    """Append '<option> <index> <info> <average> <noise>' lines to
    objective.txt (problem.py:47-66).  The average divides by
    nSteps - avgStart; 'perturb' subtracts the first line's value.  The noise
    column (an autoregressive estimate of the average's variance from the
    time series in the reference, through the external `ar` package) is
    written as 0."""
    nSteps, avgStart = ns['nSteps'], ns['avgStart']
    primal = ns['primal']
    globalResult = [res / (nSteps - avgStart) for res in result]
    if option == 'perturb':
        with open(primal.resultFile) as f:
            previous = float(f.readline().split(' ')[3])
        globalResult[0] -= previous
    noise = [0. for _ in result]
    with open(primal.resultFile, 'a') as handle:
        for index in range(len(result)):
            handle.write('{} {} {} {} {}\n'.format(option, index, info, repr(float(globalResult[index])),
                                                    noise[index]))


def main(argv=None):
    ns, rest = load(argv)
    parser = argparse.ArgumentParser()
    parser.add_argument('option', nargs='?', default='orig')
    user = parser.parse_args(rest)
    primal = ns['primal']
    nSteps, writeInterval = ns['nSteps'], ns['writeInterval']
    startTime, dt = ns['startTime'], ns['dt']

    if os.path.exists(primal.statusFile):
        startIndex, startTime, dt, initResult = primal.readStatusFile()
        print('Read status file, index =', startIndex)
    else:
        startIndex, initResult = 0, 0.

    if user.option in ('orig', 'source'):
        dts, nSims, mode = dt, 1, 'orig'
    elif user.option == 'perturb':
        timeSteps = np.loadtxt(primal.timeStepFile, ndmin=2)
        timeSteps = np.concatenate([timeSteps, [[0, 0]]])
        dts, nSims, mode = timeSteps[:, 1], ns['nPerturb'], 'perturb'
    else:
        raise SystemExit('unknown option %r (orig, perturb, source)' % user.option)

    primal.readFields(startTime)
    primal.compile()
    for sim in range(nSims):
        if user.option in ('perturb', 'source'):
            if not ns['parameters']:
                raise SystemExit('parameter not given')
            perturbation = (ns['parameters'], ns['perturb'][sim])
            primal.timeSeriesFile = primal.mesh.case + 'timeSeries_{}.txt'.format(sim)
        else:
            perturbation = None
        result = primal.run(result=initResult, startTime=startTime, dt=dts, nSteps=nSteps,
                            writeInterval=writeInterval, reportInterval=ns['reportInterval'], mode=mode,
                            startIndex=startIndex, source=ns['source'], perturbation=perturbation,
                            avgStart=ns['avgStart'])
        writeResult(ns, user.option, [result], '{}'.format(sim), primal.timeSeriesFile)
        primal.removeStatusFile()
        startIndex, initResult = 0, 0.


if __name__ == '__main__':
    main()

"""Adjoint run of a case file: apps/adjoint.py of the reference.

    python -m adfvm_amd.apps.adjoint <case file> [--readFields] [--homogeneous] [--gpu_double]

Needs the orig run's written checkpoints and time-step log (run
`python -m adfvm_amd.apps.problem <case file>` first).  Appends
'adjoint <index> <scaling> <sensitivity average> <noise>' to objective.txt
when the sweep reaches the first step; a status file lets an interrupted
sweep continue (runCheckpoints segments per invocation).
"""
import argparse

from .problem import load, writeResult
from ..adjoint import Adjoint


def main(argv=None):
    ns, rest = load(argv)
    parser = argparse.ArgumentParser()
    parser.add_argument('--readFields', action='store_true')
    parser.add_argument('--homogeneous', action='store_true')
    user = parser.parse_args(rest)
    primal = ns['primal']
    adjParams = ns['adjParams']
    adjoint = Adjoint(primal, scaling=adjParams[0], viscosityType=adjParams[1], viscosityScaler=adjParams[2])
    primal.readFields(ns['startTime'])
    primal.compile()
    adjoint.createFields()
    data = adjoint.initPrimalData(ns['nSteps'], ns['startTime'], ns['dt'], ns['nPerturb'])
    adjoint.forceReadFields = user.readFields
    adjoint.homogeneousAdjoint = user.homogeneous
    result = adjoint.run(*data, nSteps=ns['nSteps'], writeInterval=ns['writeInterval'], perturb=ns['perturb'],
                         parameters=ns['parameters'], source=ns['source'], sampleInterval=ns['sampleInterval'],
                         reportInterval=ns['reportInterval'], avgStart=ns['avgStart'],
                         runCheckpoints=ns['runCheckpoints'], viscousInterval=ns['viscousInterval'])
    if adjoint.finished:
        writeResult(ns, 'adjoint', result, str(adjoint.scaling), adjoint.sensTimeSeriesFile)


if __name__ == '__main__':
    main()

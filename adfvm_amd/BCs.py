"""Boundary-condition types of adFVM (adFVM/BCs.py) as specifications for the
ghost-cell kernels.

A field's boundary is a dict patch -> {'type': name, ...} in the form of an
OpenFOAM boundaryField entry (adFVM/field.py:123-146).  Values may be given
as numbers / arrays or as OpenFOAM strings ('uniform 300', 'uniform (0 0 0)',
'nonuniform List<scalar> n (...)').

  type            keys                 rule (reference lines)
  calculated      value (initial)      ghost keeps its value         BCs.py:74-84
  cyclic                               paired owner's value          BCs.py:86-95
  zeroGradient                         owner's value                 BCs.py:97-104
  empty / inletOutlet                  as zeroGradient               BCs.py:136-137
  fixedValue      value                value                         BCs.py:111-118
  symmetryPlane / slip                 owner's value, vectors lose   BCs.py:120-135
                                       the normal component
  CBC_UPT         U0, T0, p0 (on p)    U0, T0, p0                    BCs.py:152-161
  CBC_TOTAL_PT    Tt, pt[, direction]  from the owner's U, T         BCs.py:164-184
                  (on p)
The characteristic types are read from the p field's entry, as
RCF.characteristicBoundary does (adFVM/density.py:333-337).
"""
import re

import numpy as np

TYPES = {
    'calculated': (), 'cyclic': (), 'slidingPeriodic1D': (), 'zeroGradient': (), 'empty': (), 'inletOutlet': (),
    'fixedValue': ('value',), 'symmetryPlane': (), 'slip': (),
    'CBC_UPT': ('U0', 'T0', 'p0'), 'CBC_TOTAL_PT': ('Tt', 'pt'),
}
CHARACTERISTIC = ('CBC_UPT', 'CBC_TOTAL_PT')


def parse_value(v, nFaces, ncomp):
    """A boundary value as an (nFaces, ncomp) float64 array.  Accepts numbers,
    arrays, or OpenFOAM 'uniform x' / 'uniform (x y z)' /
    'nonuniform List<...> n (...)' strings (adFVM/mesh.py extractField)."""
    if isinstance(v, str):
        s = v.strip()
        if s.startswith('uniform'):
            nums = [float(x) for x in re.findall(r'[-+0-9.eE]+', s[len('uniform'):])]
            a = np.array(nums, np.float64)
        elif s.startswith('nonuniform'):
            body = s[s.index('(') + 1:s.rindex(')')]
            nums = [float(x) for x in re.findall(r'[-+]?[0-9.]+(?:[eE][-+]?[0-9]+)?', body)]
            a = np.array(nums, np.float64)
        else:
            raise ValueError('cannot parse boundary value %r' % v)
    else:
        a = np.asarray(v, np.float64)
    if a.size == ncomp:
        return np.tile(a.reshape(1, ncomp), (nFaces, 1))
    if a.size == nFaces * ncomp:
        return a.reshape(nFaces, ncomp)
    raise ValueError('boundary value of size %d for %d faces x %d components' % (a.size, nFaces, ncomp))


def validate(mesh, boundaries):
    """Check a {field: {patch: spec}} set against the mesh and the known
    types; returns {patch: {field: spec}} with values parsed to arrays."""
    out = {}
    for patch in mesh.boundary:
        n = mesh.boundary[patch]['nFaces']
        ptype = mesh.boundary[patch]['type']
        out[patch] = {}
        for field, k in (('U', 3), ('T', 1), ('p', 1)):
            spec = dict(boundaries.get(field, {}).get(patch, mesh.defaultBoundary[patch]))
            t = spec.get('type')
            if t not in TYPES:
                raise ValueError('unsupported boundary type %r on patch %r field %s' % (t, patch, field))
            coupled = ptype in ('cyclic', 'slidingPeriodic1D')
            if coupled != (t in ('cyclic', 'slidingPeriodic1D')):
                raise ValueError('patch %r of type %s needs a %s condition for %s, got %s'
                                 % (patch, ptype, 'cyclic' if coupled else 'non-cyclic', field, t))
            if t == 'fixedValue' or (t == 'calculated' and 'value' in spec):
                spec['value'] = parse_value(spec['value'], n, k)
            if field == 'p' and t in CHARACTERISTIC:
                for key in TYPES[t]:
                    if key not in spec:
                        raise ValueError('%s on patch %r needs %r' % (t, patch, key))
                kk = {'U0': 3, 'T0': 1, 'p0': 1, 'Tt': 1, 'pt': 1, 'direction': 3}
                for key in list(spec):
                    if key in kk:
                        spec[key] = parse_value(spec[key], n, kk[key])
            out[patch][field] = spec
        if out[patch]['U']['type'] == 'calculated' and 'value' not in out[patch]['U'] \
                and out[patch]['p']['type'] not in CHARACTERISTIC:
            out[patch]['U']['value'] = np.zeros((n, 3))
        for field in ('T', 'p'):
            s = out[patch][field]
            if s['type'] == 'calculated' and 'value' not in s and out[patch]['p']['type'] not in CHARACTERISTIC:
                raise ValueError('calculated %s on patch %r needs a value (the ghost keeps it)' % (field, patch))
    return out

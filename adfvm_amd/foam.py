"""OpenFOAM case I/O: constant/polyMesh and volScalar/volVector fields.

Reads ASCII and binary files (the reference reads with adFVM/mesh.py:206-235
and adFVM/field.py:198-264); writes ASCII.  Only what the solver path needs:
points, faces (quads), owner, neighbour, boundary; internalField
uniform/nonuniform and boundaryField dictionaries.
"""
import os
import re

import numpy as np

from .meshgen import MeshData

HEADER = '''/*--------------------------------*- C++ -*----------------------------------*\\
| =========                 |                                                 |
| \\\\      /  F ield         | OpenFOAM: The Open Source CFD Toolbox           |
|  \\\\    /   O peration     | Version:  2.2.2                                 |
|   \\\\  /    A nd           |                                                 |
|    \\\\/     M anipulation  |                                                 |
\\*---------------------------------------------------------------------------*/
FoamFile
{
    version     2.0;
    format      ascii;
    class       %s;
    location    "%s";
    object      %s;
}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

'''


def _strip_comments(b):
    b = re.sub(rb'/\*.*?\*/', b'', b, flags=re.S)
    return re.sub(rb'//[^\n]*', b'', b)


def _header(content):
    m = re.search(rb'FoamFile\s*\{(.*?)\}', content, re.S)
    if not m:
        raise ValueError('not an OpenFOAM file (no FoamFile header)')
    d = dict(re.findall(rb'(\w+)\s+([^;]+);', m.group(1)))
    return {k.decode(): v.decode().strip().strip('"') for k, v in d.items()}, content[m.end():]


def _list_body(data, start):
    """Locate 'N (' after start; returns (N, index after '(')."""
    m = re.compile(rb'(\d+)\s*\(').search(data, start)
    return int(m.group(1)), m.end()


def read_list(path, kind):
    """kind: 'label', 'scalar', 'vector', 'face'."""
    raw = open(path, 'rb').read()
    hdr, body = _header(raw)
    binary = hdr.get('format') == 'binary'
    if hdr.get('class') == 'faceCompactList':
        n1, s = _list_body(body, 0)
        offs = np.frombuffer(body[s:s + 4 * n1], np.int32) if binary else None
        if binary:
            n2, s2 = _list_body(body, s + 4 * n1)
            pts = np.frombuffer(body[s2:s2 + 4 * n2], np.int32)
        else:
            text = _strip_comments(body)
            n1, s = _list_body(text, 0)
            e = text.index(b')', s)
            offs = np.array(text[s:e].split(), np.int64)
            n2, s2 = _list_body(text, e)
            e2 = text.index(b')', s2)
            pts = np.array(text[s2:e2].split(), np.int64)
        nsz = np.diff(offs)
        if not (nsz == 4).all():
            raise ValueError('only quadrilateral faces are supported (hex meshes)')
        return pts.reshape(-1, 4).astype(np.int32)
    if binary:
        n, s = _list_body(body, 0)
        if kind == 'label':
            return np.frombuffer(body[s:s + 4 * n], np.int32).copy()
        if kind == 'scalar':
            return np.frombuffer(body[s:s + 8 * n], np.float64).copy()
        if kind == 'vector':
            return np.frombuffer(body[s:s + 24 * n], np.float64).reshape(n, 3).copy()
        raise ValueError('binary %s lists are not supported' % kind)
    text = _strip_comments(body)
    n, s = _list_body(text, 0)
    e = text.rindex(b')')
    inner = text[s:e]
    if kind == 'face':
        faces = re.findall(rb'(\d+)\s*\(([^)]*)\)', inner)
        if any(int(k) != 4 for k, _ in faces):
            raise ValueError('only quadrilateral faces are supported (hex meshes)')
        return np.array([v.split() for _, v in faces], np.int32)
    vals = np.array(inner.replace(b'(', b' ').replace(b')', b' ').split(), np.float64)
    if kind == 'label':
        return vals.astype(np.int32)
    if kind == 'vector':
        return vals.reshape(-1, 3)
    return vals


def read_boundary(path):
    text = _strip_comments(open(path, 'rb').read())
    _, body = _header(text)
    body = body.decode()
    patches = re.findall(r'([A-Za-z0-9_]+)\s*\{([^{}]*)\}', body)
    out = {}
    for name, inner in patches:
        d = dict(re.findall(r'(\w+)\s+([^;]+);', inner))
        d = {k: v.strip() for k, v in d.items()}
        d['nFaces'] = int(d['nFaces'])
        d['startFace'] = int(d['startFace'])
        if d.get('type') == 'cyclicAMI':               # mesh.py:271-273
            d['type'] = 'cyclic'
        out[name] = d
    return out


def read_polymesh(caseDir, currTime='constant'):
    """(points, faces, owner, neighbour, boundary) of a case, as
    adFVM/mesh.py:176-204 reads them."""
    pm = os.path.join(caseDir, 'constant', 'polyMesh')
    points = read_list(os.path.join(pm, 'points'), 'vector')
    faces = read_list(os.path.join(pm, 'faces'), 'face')
    owner = read_list(os.path.join(pm, 'owner'), 'label')
    neighbour = read_list(os.path.join(pm, 'neighbour'), 'label')
    boundary = read_boundary(os.path.join(pm, 'boundary'))
    from collections import OrderedDict
    b = OrderedDict(sorted(boundary.items(), key=lambda kv: kv[1]['startFace']))
    return MeshData(points, faces, owner, neighbour, b)


def write_polymesh(md, caseDir):
    pm = os.path.join(caseDir, 'constant', 'polyMesh')
    os.makedirs(pm, exist_ok=True)

    def w(name, cls, body):
        with open(os.path.join(pm, name), 'w') as f:
            f.write(HEADER % (cls, 'constant/polyMesh', name))
            f.write(body)

    pts = '\n'.join('(%.17g %.17g %.17g)' % tuple(p) for p in md.points)
    w('points', 'vectorField', '%d\n(\n%s\n)\n' % (len(md.points), pts))
    fcs = '\n'.join('4(%d %d %d %d)' % tuple(f) for f in md.faces)
    w('faces', 'faceList', '%d\n(\n%s\n)\n' % (len(md.faces), fcs))
    w('owner', 'labelList', '%d\n(\n%s\n)\n' % (len(md.owner), '\n'.join(map(str, md.owner))))
    w('neighbour', 'labelList', '%d\n(\n%s\n)\n' % (len(md.neighbour), '\n'.join(map(str, md.neighbour))))
    lines = ['%d\n(' % len(md.boundary)]
    for name, b in md.boundary.items():
        lines.append('    %s\n    {' % name)
        lines.append('        type            %s;' % b['type'])
        if 'neighbourPatch' in b:
            lines.append('        neighbourPatch  %s;' % b['neighbourPatch'])
        lines.append('        nFaces          %d;' % b['nFaces'])
        lines.append('        startFace       %d;' % b['startFace'])
        lines.append('    }')
    lines.append(')\n')
    w('boundary', 'polyBoundaryMesh', '\n'.join(lines))


def time_string(t):
    """Time directory name: integers plain, others with 11 decimals
    (adFVM/mesh.py:129-136)."""
    t = float(t)
    return '%d' % int(t) if t.is_integer() else '{0:.11f}'.format(t)


def time_dir(caseDir, t):
    return os.path.join(caseDir, time_string(t))


def _field_values(text, n, ncomp, binary):
    m = re.search(rb'internalField\s+uniform\s+(\([^)]*\)|[^;]+);', text)
    if m:
        v = np.array(m.group(1).replace(b'(', b' ').replace(b')', b' ').split(), np.float64)
        return np.tile(v.reshape(1, ncomp), (n, 1)), text[m.end():]
    m = re.search(rb'internalField\s+nonuniform\s+List<\w+>\s*(\d+)\s*\(', text)
    if not m:
        raise ValueError('internalField not found')
    cnt = int(m.group(1))
    if cnt != n:
        raise ValueError('internalField has %d values for %d cells' % (cnt, n))
    s = m.end()
    if binary:
        vals = np.frombuffer(text[s:s + cnt * ncomp * 8], np.float64).reshape(cnt, ncomp).copy()
        rest = text[s + cnt * ncomp * 8:]
    else:
        e = text.index(b';', s)
        vals = np.array(text[s:e].replace(b'(', b' ').replace(b')', b' ').split(), np.float64)
        vals = vals[:cnt * ncomp].reshape(cnt, ncomp)
        rest = text[e:]
    return vals, rest


def read_field(caseDir, t, name, n, ncomp):
    """(internal values (n, ncomp), boundaryField {patch: {key: value-str}})."""
    raw = open(os.path.join(time_dir(caseDir, t), name), 'rb').read()
    hdr, body = _header(raw)
    vals, rest = _field_values(body, n, ncomp, hdr.get('format') == 'binary')
    m = re.search(rb'boundaryField\s*\{', rest)
    bnd = {}
    if m:
        text = _strip_comments(rest[m.end():]).decode('latin-1')
        for pname, inner in re.findall(r'([A-Za-z0-9_]+)\s*\{([^{}]*)\}', text):
            d = {}
            for k, v in re.findall(r'(\w+)\s+([^;]+);', inner):
                d[k] = v.strip()
            bnd[pname] = d
    return vals, bnd


def read_fields(mesh, t, suffix=''):
    nIC = mesh.nInternalCells
    U, bU = read_field(mesh.caseDir, t, 'U' + suffix, nIC, 3)
    T, bT = read_field(mesh.caseDir, t, 'T' + suffix, nIC, 1)
    p, bp = read_field(mesh.caseDir, t, 'p' + suffix, nIC, 1)
    return U, T, p, {'U': bU, 'T': bT, 'p': bp}


def _fmt_value(v, ncomp):
    a = np.asarray(v, np.float64).reshape(-1, ncomp)
    if (a == a[0]).all():
        return 'uniform ' + ('%.17g' % a[0, 0] if ncomp == 1 else '(%s)' % ' '.join('%.17g' % x for x in a[0]))
    cls = 'scalar' if ncomp == 1 else 'vector'
    body = '\n'.join(('%.17g' % r[0]) if ncomp == 1 else '(%s)' % ' '.join('%.17g' % x for x in r) for r in a)
    return 'nonuniform List<%s> %d\n(\n%s\n)\n' % (cls, len(a), body)


def write_fields(mesh, t, fields, boundaries, default=None):
    """Write each field of {name: (nIC, k) array} to the time directory;
    a patch's boundaryField entry is boundaries[name][patch], else
    default[patch] (the mesh's calculated boundary by default)."""
    d = time_dir(mesh.caseDir, t)
    os.makedirs(d, exist_ok=True)
    for name, arr in fields.items():
        arr = np.asarray(arr, np.float64)
        ncomp = arr.shape[1] if arr.ndim == 2 else 1
        cls = 'volScalarField' if ncomp == 1 else 'volVectorField'
        lines = ['dimensions      [0 0 0 0 0 0 0];', '', 'internalField   %s;' % _fmt_value(arr, ncomp).rstrip(),
                 '', 'boundaryField', '{']
        specs = boundaries.get(name, {})
        for patch in mesh.boundary:
            spec = specs.get(patch, (default or mesh.calculatedBoundary)[patch])
            lines.append('    %s\n    {' % patch)
            lines.append('        type            %s;' % spec['type'])
            for k, v in spec.items():
                if k == 'type':
                    continue
                if isinstance(v, str):
                    lines.append('        %-15s %s;' % (k, v))
                else:
                    kk = 3 if k in ('U0', 'direction') or (k == 'value' and name == 'U') else 1
                    lines.append('        %-15s %s;' % (k, _fmt_value(v, kk).rstrip()))
            lines.append('    }')
        lines.append('}')
        with open(os.path.join(d, name), 'w') as f:
            f.write(HEADER % (cls, os.path.basename(d), name))
            f.write('\n'.join(lines) + '\n')

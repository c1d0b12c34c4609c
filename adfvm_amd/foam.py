"""OpenFOAM case I/O: constant/polyMesh and volScalar/volVector fields.

Reads ASCII and binary files (the reference reads with adFVM/mesh.py:206-235
and adFVM/field.py:198-264).  Writes the polyMesh in ASCII and fields in
config.fileFormat, binary by default as the reference's
(adFVM/config.py:198, IOField.writeFoamField field.py:436-476 with
writeField mesh.py:1048-1059: a nonuniform list is its count, '(', the
float64 values' bytes and ')').  Only what the solver path needs: points,
faces (quads), owner, neighbour, boundary; internalField uniform/nonuniform
and boundaryField dictionaries.
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
    format      %s;
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
            f.write(HEADER % ('ascii', cls, 'constant/polyMesh', name))
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


_TOKEN = re.compile(rb'\s*(?://[^\n]*\s*|/\*.*?\*/\s*)*([A-Za-z0-9_.\-+<>]+|[{}])', re.S)
_BIN_LIST = re.compile(rb'\s*nonuniform\s+List<(\w+)>\s*(\d+)\s*\(')


def _boundary_field(text, start, binary):
    """{patch: {key: value}} of the boundaryField dictionary opening at
    text[start] ('{').  Values are stripped strings, except a binary file's
    nonuniform lists, read by their byte count (8 per scalar component: the
    bytes may hold any character) as (count, components) float64 arrays."""
    out = {}
    i = start + 1
    while True:
        m = _TOKEN.match(text, i)
        if m is None or m.group(1) == b'}':
            return out
        name, i = m.group(1).decode(), m.end()
        m = _TOKEN.match(text, i)
        if m is None or m.group(1) != b'{':
            raise ValueError('boundaryField: no dictionary for patch %s' % name)
        i = m.end()
        d = {}
        while True:
            m = _TOKEN.match(text, i)
            if m is None:
                raise ValueError('boundaryField: patch %s not closed' % name)
            i = m.end()
            if m.group(1) == b'}':
                break
            key = m.group(1).decode()
            mb = _BIN_LIST.match(text, i) if binary else None
            if mb:
                k = 3 if mb.group(1) == b'vector' else 1
                cnt = int(mb.group(2))
                s = mb.end()
                d[key] = np.frombuffer(text[s:s + 8 * k * cnt], np.float64).reshape(cnt, k).copy()
                e = re.compile(rb'\s*\)\s*;').match(text, s + 8 * k * cnt)
                if e is None:
                    raise ValueError('boundaryField: %s.%s holds fewer than %d values' % (name, key, cnt))
                i = e.end()
            else:
                e = text.index(b';', i)
                d[key] = text[i:e].decode('latin-1').strip()
                i = e + 1
        out[name] = d


def read_field(caseDir, t, name, n, ncomp):
    """(internal values (n, ncomp), boundaryField {patch: {key: value}}):
    values as strings ('uniform 300', 'zeroGradient', an ASCII nonuniform
    list), a binary file's nonuniform lists as float64 arrays."""
    raw = open(os.path.join(time_dir(caseDir, t), name), 'rb').read()
    hdr, body = _header(raw)
    binary = hdr.get('format') == 'binary'
    vals, rest = _field_values(body, n, ncomp, binary)
    m = re.search(rb'boundaryField\s*\{', rest)
    bnd = _boundary_field(rest, m.end() - 1, binary) if m else {}
    return vals, bnd


def read_fields(mesh, t, suffix=''):
    nIC = mesh.nInternalCells
    U, bU = read_field(mesh.caseDir, t, 'U' + suffix, nIC, 3)
    T, bT = read_field(mesh.caseDir, t, 'T' + suffix, nIC, 1)
    p, bp = read_field(mesh.caseDir, t, 'p' + suffix, nIC, 1)
    return U, T, p, {'U': bU, 'T': bT, 'p': bp}


def _fmt_value(v, ncomp, binary=False):
    """A field or boundary value as it follows its keyword, as bytes: uniform
    when every row is equal, else a nonuniform list -- in a binary file the
    count, '(', the values' float64 bytes and ')' (adFVM/mesh.py:1048-1059)."""
    a = np.asarray(v, np.float64).reshape(-1, ncomp)
    if len(a) and (a == a[0]).all():
        return ('uniform ' + ('%.17g' % a[0, 0] if ncomp == 1 else
                              '(%s)' % ' '.join('%.17g' % x for x in a[0]))).encode()
    cls = 'scalar' if ncomp == 1 else 'vector'
    if binary:
        return b'nonuniform List<%s>\n%d\n(' % (cls.encode(), len(a)) + np.ascontiguousarray(a).tobytes() + b')\n'
    # ASCII: the list closes on its own line; the caller adds ';'
    body = '\n'.join(('%.17g' % r[0]) if ncomp == 1 else '(%s)' % ' '.join('%.17g' % x for x in r) for r in a)
    return ('nonuniform List<%s> %d\n(\n%s\n)\n' % (cls, len(a), body)).encode()


def write_fields(mesh, t, fields, boundaries, default=None, fmt=None):
    """Write each field of {name: (nIC, k) array} to the time directory;
    a patch's boundaryField entry is boundaries[name][patch], else
    default[patch] (the mesh's calculated boundary by default).  fmt:
    'binary' or 'ascii' (config.fileFormat when None)."""
    from . import config
    fmt = fmt or config.fileFormat
    if fmt not in ('binary', 'ascii'):
        raise ValueError('file format %r (binary, ascii)' % (fmt,))
    binary = fmt == 'binary'
    d = time_dir(mesh.caseDir, t)
    os.makedirs(d, exist_ok=True)
    for name, arr in fields.items():
        arr = np.asarray(arr, np.float64)
        ncomp = arr.shape[1] if arr.ndim == 2 else 1
        cls = 'volScalarField' if ncomp == 1 else 'volVectorField'
        def val(a, k):
            # a binary list keeps ')\n' before ';', as the reference writes it
            # (and its bytes are never stripped)
            b = _fmt_value(a, k, binary)
            return b if binary and b.endswith(b')\n') else b.rstrip()

        out = [b'dimensions      [0 0 0 0 0 0 0];\n\n',
               b'internalField   ' + val(arr, ncomp) + b';\n\n', b'boundaryField\n{\n']
        specs = boundaries.get(name, {})
        for patch in mesh.boundary:
            spec = specs.get(patch, (default or mesh.calculatedBoundary)[patch])
            out.append(b'    %s\n    {\n' % patch.encode())
            out.append(b'        type            %s;\n' % spec['type'].encode())
            for k, v in spec.items():
                if k == 'type':
                    continue
                if isinstance(v, str):
                    out.append(b'        %-15s %s;\n' % (k.encode(), v.encode()))
                else:
                    kk = 3 if k in ('U0', 'direction') or (k == 'value' and name == 'U') else 1
                    out.append(b'        %-15s ' % k.encode() + val(v, kk) + b';\n')
            out.append(b'    }\n')
        out.append(b'}\n')
        with open(os.path.join(d, name), 'wb') as f:
            f.write((HEADER % (fmt, cls, os.path.basename(d), name)).encode())
            f.write(b''.join(out))

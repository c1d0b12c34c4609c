"""Hex-mesh generators that emit OpenFOAM-ordered face lists.

No mesh files ship with adFVM's cases (its `.gitignore` drops
`constant/polyMesh/{points,faces,owner,neighbour}` and the tests run OpenFOAM
`blockMesh`, `tests/setup_tests.sh:7-8`), so this module plays blockMesh's
part.  A generated mesh is the tuple adFVM's `Mesh.readFoam` returns
(`adFVM/mesh.py:176-204`): points, quad faces, owner, neighbour, boundary.

Conventions (the same ones `Mesh.populateSizes`, `cmesh.buildBeforeWrite` and
`createGhostCells` rely on, `mesh.py:517-523`, `cmesh.cpp:3-63`,
`mesh.py:758-819`):

* internal faces come first, upper-triangular: owner < neighbour, sorted by
  (owner, neighbour);
* a face's points run so that (p0-p1) x (p1-p2) points from owner to
  neighbour (outward on boundary faces), the normal rule of `cmesh.cpp:72-87`;
* boundary faces follow, grouped by patch in patch order; the k-th face of a
  cyclic patch pairs with the k-th face of its `neighbourPatch`.

`renumber_cells` applies a cell permutation (a seeded random one, then reverse
Cuthill-McKee) and restores these conventions.
"""
from collections import OrderedDict

import numpy as np


class MeshData(object):
    """points (nP,3) f64, faces (nF,4) i32, owner (nF,) i32,
    neighbour (nIF,) i32, boundary: OrderedDict name -> dict(type, startFace,
    nFaces[, neighbourPatch])."""

    def __init__(self, points, faces, owner, neighbour, boundary, nCells=None):
        self._nCells = nCells
        self.points = np.ascontiguousarray(points, np.float64)
        self.faces = np.ascontiguousarray(faces, np.int32)
        self.owner = np.ascontiguousarray(owner, np.int32)
        self.neighbour = np.ascontiguousarray(neighbour, np.int32)
        self.boundary = boundary

    @property
    def nCells(self):
        if self._nCells is not None:
            return self._nCells
        return int(self.owner.max()) + 1

    def copy(self):
        b = OrderedDict((k, dict(v)) for k, v in self.boundary.items())
        return MeshData(self.points.copy(), self.faces.copy(), self.owner.copy(),
                        self.neighbour.copy(), b, self._nCells)


def _structured(coords, shape, wrap_i, side_patches):
    """Single structured hex block.

    coords: (ni+?, nj+1, nk+1, 3) point coordinates indexed [i, j, k]; along i
    there are ni+1 point planes, or ni when wrap_i (the last cell plane joins
    plane 0 through internal faces, an O-grid seam).
    side_patches: dict side -> (patchName, patchType[, neighbourPatch]) for
    sides 'x0','x1','y0','y1','z0','z1' (x0 is absent when wrap_i).
    """
    ni, nj, nk = shape
    npi = ni if wrap_i else ni + 1
    assert coords.shape == (npi, nj + 1, nk + 1, 3)
    points = coords.transpose(2, 1, 0, 3).reshape(-1, 3)   # i fastest

    def P(i, j, k):
        return (np.asarray(i) % npi) + npi * (np.asarray(j) + (nj + 1) * np.asarray(k))

    def C(i, j, k):
        return np.asarray(i) + ni * (np.asarray(j) + nj * np.asarray(k))

    I, J, K = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing='ij')
    I, J, K = I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()

    # oriented quads, normal along +x / +y / +z at the low corner (i,j,k)
    def quad_x(i, j, k):
        return np.stack([P(i, j, k), P(i, j + 1, k), P(i, j + 1, k + 1), P(i, j, k + 1)], -1)

    def quad_y(i, j, k):
        return np.stack([P(i, j, k), P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j, k)], -1)

    def quad_z(i, j, k):
        return np.stack([P(i, j, k), P(i + 1, j, k), P(i + 1, j + 1, k), P(i, j + 1, k)], -1)

    fa, fb, fq = [], [], []   # cell on the -side, cell on the +side, quad (normal -> +side)
    m = I < ni - 1 if not wrap_i else np.ones_like(I, bool)
    fa.append(C(I[m], J[m], K[m])); fb.append(C((I[m] + 1) % ni, J[m], K[m])); fq.append(quad_x(I[m] + 1, J[m], K[m]))
    m = J < nj - 1
    fa.append(C(I[m], J[m], K[m])); fb.append(C(I[m], J[m] + 1, K[m])); fq.append(quad_y(I[m], J[m] + 1, K[m]))
    m = K < nk - 1
    fa.append(C(I[m], J[m], K[m])); fb.append(C(I[m], J[m], K[m] + 1)); fq.append(quad_z(I[m], J[m], K[m] + 1))
    fa, fb, fq = np.concatenate(fa), np.concatenate(fb), np.concatenate(fq)

    # boundary sides, faces in (j,k)/(i,k)/(i,j) order with the first index fastest
    jj, kk = [a.T.ravel() for a in np.meshgrid(np.arange(nj), np.arange(nk), indexing='ij')]
    ii_y, kk_y = [a.T.ravel() for a in np.meshgrid(np.arange(ni), np.arange(nk), indexing='ij')]
    ii_z, jj_z = [a.T.ravel() for a in np.meshgrid(np.arange(ni), np.arange(nj), indexing='ij')]
    sides = OrderedDict()
    if not wrap_i:
        sides['x0'] = (C(0, jj, kk), quad_x(0, jj, kk)[:, ::-1])
        sides['x1'] = (C(ni - 1, jj, kk), quad_x(ni, jj, kk))
    sides['y0'] = (C(ii_y, 0, kk_y), quad_y(ii_y, 0, kk_y)[:, ::-1])
    sides['y1'] = (C(ii_y, nj - 1, kk_y), quad_y(ii_y, nj, kk_y))
    sides['z0'] = (C(ii_z, jj_z, 0), quad_z(ii_z, jj_z, 0)[:, ::-1])
    sides['z1'] = (C(ii_z, jj_z, nk - 1), quad_z(ii_z, jj_z, nk))

    patch_faces = OrderedDict()
    patch_info = OrderedDict()
    for side, (cells, quads) in sides.items():
        spec = side_patches[side]
        name, ptype = spec[0], spec[1]
        if name not in patch_faces:
            patch_faces[name] = ([], [])
            patch_info[name] = {'type': ptype}
            if len(spec) > 2:
                patch_info[name]['neighbourPatch'] = spec[2]
        patch_faces[name][0].append(cells)
        patch_faces[name][1].append(quads)

    bowner, bquads, boundary = [], [], OrderedDict()
    start = len(fa)
    for name, (cells, quads) in patch_faces.items():
        cells, quads = np.concatenate(cells), np.concatenate(quads)
        boundary[name] = dict(patch_info[name], startFace=start, nFaces=len(cells))
        start += len(cells)
        bowner.append(cells)
        bquads.append(quads)
    return _assemble(points, fa, fb, fq, np.concatenate(bowner), np.concatenate(bquads), boundary, ni * nj * nk)


def _assemble(points, ca, cb, quads, bowner, bquads, boundary, nCells):
    """Orient internal faces owner->neighbour with owner<neighbour, sort them
    upper-triangular, append boundary faces."""
    ca = np.asarray(ca, np.int64); cb = np.asarray(cb, np.int64)
    flip = ca > cb
    owner = np.where(flip, cb, ca)
    neigh = np.where(flip, ca, cb)
    quads = np.where(flip[:, None], quads[:, ::-1], quads)
    order = np.lexsort((neigh, owner))
    faces = np.concatenate([quads[order], bquads]).astype(np.int32)
    owner = np.concatenate([owner[order], bowner]).astype(np.int32)
    return MeshData(points, faces, owner, neigh[order].astype(np.int32), boundary, nCells)


def renumber_cells(mesh, new_index):
    """Relabel cell c as new_index[c] and restore the face-order conventions.

    Boundary faces keep their patch and in-patch position (so cyclic pairing
    by position survives); internal faces are re-oriented and re-sorted."""
    new_index = np.asarray(new_index, np.int64)
    nIF = len(mesh.neighbour)
    ca = new_index[mesh.owner[:nIF]]
    cb = new_index[mesh.neighbour]
    return _assemble(mesh.points, ca, cb, mesh.faces[:nIF], new_index[mesh.owner[nIF:]],
                     mesh.faces[nIF:], OrderedDict((k, dict(v)) for k, v in mesh.boundary.items()),
                     mesh.nCells)


def cell_adjacency(mesh):
    import scipy.sparse as sp
    nIF = len(mesh.neighbour)
    n = mesh.nCells
    o = mesh.owner[:nIF].astype(np.int64)
    nb = mesh.neighbour.astype(np.int64)
    data = np.ones(2 * nIF, np.int8)
    return sp.csr_matrix((data, (np.concatenate([o, nb]), np.concatenate([nb, o]))), shape=(n, n))


def rcm_permutation(mesh):
    """new_index such that cells are in SciPy reverse Cuthill-McKee order over
    the internal-face adjacency graph."""
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    order = reverse_cuthill_mckee(cell_adjacency(mesh), symmetric_mode=True)
    new_index = np.empty(len(order), np.int64)
    new_index[order] = np.arange(len(order))
    return new_index


def random_permutation(n, seed):
    return np.random.default_rng(seed).permutation(n)


def permute_and_rcm(mesh, seed=None, rcm=True):
    """Seeded random cell permutation, then RCM reorder (either step optional)."""
    if seed is not None:
        mesh = renumber_cells(mesh, random_permutation(mesh.nCells, seed))
    if rcm:
        mesh = renumber_cells(mesh, rcm_permutation(mesh))
    # adFVM takes nInternalCells = max(owner)+1 (mesh.py:521): the last cell
    # must own a face
    if int(mesh.owner.max()) + 1 != mesh.nCells:
        raise ValueError('cell numbering leaves the last cell without an owned face')
    return mesh


def concatenate(meshes, prefixes=None):
    """Disjoint union of meshes as one mesh (an ensemble in one context):
    mesh k's cells follow mesh k-1's; the internal faces of all meshes come
    first, in mesh order, then every mesh's patches, renamed prefix + name
    (cyclic partners renamed alike).  Each mesh's own cell and face order is
    kept, so per-mesh geometry, slot tables and kernel arithmetic are those
    of the mesh alone.  Returns (MeshData, cell offsets (len(meshes)+1,))."""
    prefixes = prefixes or ['m%d_' % k for k in range(len(meshes))]
    pts, fint, oint, nint, fb, ob = [], [], [], [], [], []
    boundary = OrderedDict()
    cur = sum(len(m.neighbour) for m in meshes)
    p_off = c_off = 0
    offsets = [0]
    for m, pre in zip(meshes, prefixes):
        nIF = len(m.neighbour)
        pts.append(m.points)
        fint.append(m.faces[:nIF] + p_off)
        oint.append(m.owner[:nIF] + c_off)
        nint.append(m.neighbour + c_off)
        for name, b in sorted(m.boundary.items(), key=lambda kv: kv[1]['startFace']):
            st, n = b['startFace'], b['nFaces']
            fb.append(m.faces[st:st + n] + p_off)
            ob.append(m.owner[st:st + n] + c_off)
            nb = dict(b)
            nb['startFace'] = cur
            if 'neighbourPatch' in b:
                nb['neighbourPatch'] = pre + b['neighbourPatch']
            boundary[pre + name] = nb
            cur += n
        p_off += len(m.points)
        c_off += m.nCells
        offsets.append(c_off)
    md = MeshData(np.concatenate(pts), np.concatenate(fint + fb), np.concatenate(oint + ob),
                  np.concatenate(nint), boundary, c_off)
    return md, np.array(offsets, np.int64)


def box(n, lengths=(1., 1., 1.), origin=(0., 0., 0.), patches=None, x_coords=None):
    """Cartesian box of n=(nx,ny,nz) cells.  patches maps sides
    'x0','x1','y0','y1','z0','z1' -> (name, type[, neighbourPatch])."""
    nx, ny, nz = n
    if patches is None:
        patches = {s: ('walls', 'patch') for s in ['x0', 'x1', 'y0', 'y1', 'z0', 'z1']}
    xs = np.linspace(origin[0], origin[0] + lengths[0], nx + 1) if x_coords is None else np.asarray(x_coords, np.float64)
    ys = np.linspace(origin[1], origin[1] + lengths[1], ny + 1)
    zs = np.linspace(origin[2], origin[2] + lengths[2], nz + 1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
    coords = np.stack([X, Y, Z], -1)
    return _structured(coords, (nx, ny, nz), False, patches)


def transform(mesh, fn):
    """The same mesh with every point moved by fn((nP,3) -> (nP,3)): an
    affine map keeps faces planar and turns a Cartesian box into
    parallelepiped cells whose centre-to-centre directions are not the face
    normals (a non-orthogonal mesh)."""
    out = mesh.copy()
    out.points = np.ascontiguousarray(fn(out.points), np.float64)
    return out


def sheared_box(n, lengths=(1., 1., 1.), shear=((1., 0.3, 0.2), (0., 1., 0.25), (0., 0., 1.)),
                patches=None, x_coords=None):
    """Cartesian box mapped by the matrix `shear` (points x -> shear @ x)."""
    S = np.asarray(shear, np.float64)
    return transform(box(n, lengths, patches=patches, x_coords=x_coords), lambda p: p @ S.T)


def periodic_box(n, L=2 * np.pi, origin=(0., 0., 0.)):
    """Fully periodic cube: three cyclic patch pairs (like the convection
    case's blockMeshDict, `cases/convection/constant/polyMesh/blockMeshDict:41-90`)."""
    if np.isscalar(n):
        n = (n, n, n)
    if np.isscalar(L):
        L = (L, L, L)
    patches = {'x0': ('patch0_half0', 'cyclic', 'patch0_half1'), 'x1': ('patch0_half1', 'cyclic', 'patch0_half0'),
               'y0': ('patch1_half0', 'cyclic', 'patch1_half1'), 'y1': ('patch1_half1', 'cyclic', 'patch1_half0'),
               'z0': ('patch2_half0', 'cyclic', 'patch2_half1'), 'z1': ('patch2_half1', 'cyclic', 'patch2_half0')}
    return box(n, L, origin, patches)


def shock_tube(nx=1000, length=1.0, dy=None, origin=0.0):
    """1-D tube: nx cells along x, one cell in y and z; 'sides' at both x ends
    and 'empty' on y/z (the shockTube case's patches, cases/shockTube/
    constant/polyMesh/blockMeshDict)."""
    dx = length / nx
    dy = dx if dy is None else dy
    patches = {'x0': ('sides', 'patch'), 'x1': ('sides', 'patch'),
               'y0': ('empty', 'empty'), 'y1': ('empty', 'empty'),
               'z0': ('empty', 'empty'), 'z1': ('empty', 'empty')}
    return box((nx, 1, 1), (length, dy, dy), (origin, -dy / 2, -dy / 2), patches)


def o_grid(n_theta=512, n_r=256, r0=0.5, r1=20.0, ratio=1.02, dz=None, wall='cylinder', farfield='farfield'):
    """O-grid around a cylinder of radius r0: n_theta x n_r x 1 hex cells, the
    theta seam joined by internal faces, radial spacing growing geometrically
    by `ratio` from r0 to r1, 'empty' front/back."""
    g = ratio ** np.arange(n_r)
    dr = (r1 - r0) / g.sum()
    r = r0 + np.concatenate([[0.], np.cumsum(dr * g)])
    r[-1] = r1
    th = -2 * np.pi * np.arange(n_theta) / n_theta     # decreasing theta: j=r outward, k=z, right-handed
    dz = (2 * np.pi * r0 / n_theta) if dz is None else dz
    z = np.array([-dz / 2, dz / 2])
    TH, R, Z = np.meshgrid(th, r, z, indexing='ij')
    coords = np.stack([R * np.cos(TH), R * np.sin(TH), Z], -1)
    patches = {'y0': (wall, 'patch'), 'y1': (farfield, 'patch'),
               'z0': ('empty', 'empty'), 'z1': ('empty', 'empty')}
    return _structured(coords, (n_theta, n_r, 1), True, patches)

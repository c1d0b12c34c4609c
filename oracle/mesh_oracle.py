"""ORACLE (test infrastructure only): numpy restatement of adFVM's mesh index
construction and geometry.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.  The product builds meshes in
C++ (adfvm_amd/csrc/mesh_build.cpp).

Every function cites the reference lines it restates.  Integer outputs are
meant to be bit-identical to the product's; floating-point outputs follow
cmesh.cpp's operation order (evaluated with numpy in float64).
"""
import numpy as np

CYCLIC_TYPES = ('cyclic', 'slidingPeriodic1D')          # adFVM/config.py:389
PROCESSOR_TYPES = ('processor', 'processorCyclic')      # adFVM/config.py:388


class OracleMesh(object):
    pass


def populate_sizes(m):
    """adFVM/mesh.py:517-523."""
    m.nInternalFaces = len(m.neighbour)
    m.nFaces = len(m.owner)
    m.nBoundaryFaces = m.nFaces - m.nInternalFaces
    m.nInternalCells = int(np.max(m.owner)) + 1
    m.nGhostCells = m.nBoundaryFaces
    m.nCells = m.nInternalCells + m.nGhostCells


def build_before_write(m):
    """cellFaces fill order and hex point lists, adFVM/cpp/cmesh.cpp:3-63.

    cellFaces[c] = faces owned by c in face order (internal then boundary),
    then faces whose neighbour is c in face order."""
    m.cellFaces = cell_faces(m.owner, m.neighbour, m.nInternalCells)
    m.cells = build_cells(m.faces, m.cellFaces) if m.with_cells else None


def cell_faces_loop(owner, neighbour, nIC):
    """Literal restatement of the two fill loops, cmesh.cpp:9-19."""
    cellFaces = np.zeros((nIC, 6), np.int32)
    fill = np.zeros(nIC, np.int64)
    for f in range(len(owner)):
        p = owner[f]
        cellFaces[p, fill[p]] = f
        fill[p] += 1
    for f in range(len(neighbour)):
        p = neighbour[f]
        cellFaces[p, fill[p]] = f
        fill[p] += 1
    return cellFaces


def cell_faces(owner, neighbour, nIC):
    """Same order as cell_faces_loop: a stable sort by cell of the owner list
    followed by the neighbour list."""
    cells = np.concatenate([owner, neighbour]).astype(np.int64)
    ids = np.concatenate([np.arange(len(owner)), np.arange(len(neighbour))])
    order = np.argsort(cells, kind='stable')
    return ids[order].reshape(nIC, 6).astype(np.int32)


def build_cells(faces, cellFaces):
    """cmesh.cpp:20-60.  For each point of the cell's first face, find the
    adjacent face containing it and take that face's point across the edge
    leaving the first face.  The reference writes `l = (k - 1) % 4`, which is
    -1 for k == 0 in C; the intended cyclic predecessor (k + 3) % 4 is used."""
    nIC = cellFaces.shape[0]
    cells = np.zeros((nIC, 8), np.int32)
    for i in range(nIC):
        first = faces[cellFaces[i, 0]]
        for j in range(4):
            point = first[j]
            found = False
            for n in range(1, 6):
                nxt = faces[cellFaces[i, n]]
                for k in range(4):
                    if nxt[k] == point:
                        l = (k + 1) % 4
                        for mm in range(4):
                            if first[mm] == nxt[l]:
                                l = (k + 3) % 4
                                break
                        cells[i, 4 + j] = nxt[l]
                        found = True
                        break
                if found:
                    break
        cells[i, :4] = first
    return cells


def split_patches(m):
    """adFVM/mesh.py:238-249 and :100-101 (sortedPatches)."""
    ids = sorted(m.boundary.keys(), key=lambda x: (m.boundary[x]['startFace'], m.boundary[x]['nFaces']))
    local, remote = [], []
    for pid in ids:
        b = m.boundary[pid]
        b['cellStartFace'] = b['startFace'] - m.nInternalFaces + m.nInternalCells
        (remote if b['type'] in PROCESSOR_TYPES else local).append(pid)
    m.localPatches, m.remotePatches = local, remote
    m.sortedPatches = sorted(local)


def face_geometry(points, faces):
    """Unit normals, areas and area-weighted centres, cmesh.cpp:70-121."""
    a, b, c = points[faces[:, 0]], points[faces[:, 1]], points[faces[:, 2]]
    v1 = a - b
    v2 = b - c
    normal = np.stack([v1[:, 1] * v2[:, 2] - v1[:, 2] * v2[:, 1],
                       v1[:, 2] * v2[:, 0] - v1[:, 0] * v2[:, 2],
                       v1[:, 0] * v2[:, 1] - v1[:, 1] * v2[:, 0]], 1)
    Ns = np.sqrt(normal[:, 0] * normal[:, 0] + normal[:, 1] * normal[:, 1] + normal[:, 2] * normal[:, 2])
    normal = normal / Ns[:, None]
    pts = points[faces]                                   # (nF,4,3)
    fc = (((pts[:, 0] + pts[:, 1]) + pts[:, 2]) + pts[:, 3]) / 4
    area = np.zeros(len(faces))
    sumc = np.zeros((len(faces), 3))
    for j in range(4):
        p = pts[:, j]
        q = pts[:, (j + 1) % 4]
        avg = (p + q + fc) / 3
        v1 = q - p
        v2 = fc - p
        N = np.stack([v1[:, 1] * v2[:, 2] - v1[:, 2] * v2[:, 1],
                      v1[:, 2] * v2[:, 0] - v1[:, 0] * v2[:, 2],
                      v1[:, 0] * v2[:, 1] - v1[:, 1] * v2[:, 0]], 1)
        ns = np.sqrt(N[:, 0] * N[:, 0] + N[:, 1] * N[:, 1] + N[:, 2] * N[:, 2])
        area = area + ns / 2
        sumc = sumc + ns[:, None] * avg / 2
    return normal, sumc / area[:, None], area


def cell_geometry(cellFaces, faceCentres, areas, normals):
    """Pyramid-decomposition centres and volumes, cmesh.cpp:126-163."""
    fcs = faceCentres[cellFaces]                          # (nIC,6,3)
    cc = fcs[:, 0]
    for j in range(1, 6):
        cc = cc + fcs[:, j]
    cc = cc / 6
    vol = np.zeros(len(cellFaces))
    sumc = np.zeros((len(cellFaces), 3))
    for j in range(6):
        f = cellFaces[:, j]
        fcj = faceCentres[f]
        h = cc - fcj
        aN = areas[f][:, None] * normals[f]
        v = aN[:, 0] * h[:, 0] + aN[:, 1] * h[:, 1] + aN[:, 2] * h[:, 2]
        v = np.abs(v / 3)
        vol = vol + v
        sumc = sumc + v[:, None] * (3. / 4 * fcj + 1. / 4 * cc)
    return sumc / vol[:, None], vol


def create_ghost_cells(m):
    """Ghost-extended neighbour and ghost centres, adFVM/mesh.py:758-819
    (single rank: no processor patches)."""
    nIF, nIC = m.nInternalFaces, m.nInternalCells
    m.neighbour = np.concatenate([m.neighbour, np.zeros(m.nBoundaryFaces, np.int32)]).astype(np.int32)
    m.cellCentres = np.concatenate([m.cellCentres, np.zeros((m.nBoundaryFaces, 3))])
    nLocal = nIC
    for pid, patch in m.boundary.items():
        start, nF = patch['startFace'], patch['nFaces']
        cstart = start - nIF + nIC
        if nF == 0:
            continue
        if patch['type'] not in PROCESSOR_TYPES:
            nLocal += nF
        m.neighbour[start:start + nF] = np.arange(cstart, cstart + nF)
        if patch['type'] in CYCLIC_TYPES:
            nb = m.boundary[patch['neighbourPatch']]
            nstart = nb['startFace']
            idx = m.owner[nstart:nstart + nF]
            transform = m.faceCentres[start] - m.faceCentres[nstart]
            patch['loc_neighbourIndices'] = idx
            m.cellCentres[cstart:cstart + nF] = transform + m.cellCentres[idx]
        else:
            m.cellCentres[cstart:cstart + nF] = m.faceCentres[start:start + nF]
    m.nLocalCells = nLocal


def face_weights(m):
    """deltas, deltasUnit, weights, linearWeights, quadraticWeights,
    cmesh.cpp:173-238."""
    P = m.cellCentres[m.owner]
    N = m.cellCentres[m.neighbour]
    delta = P - N
    d = np.sqrt(delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1] + delta[:, 2] * delta[:, 2])
    m.deltas = d
    m.deltasUnit = -delta / d[:, None]
    F = m.faceCentres
    nF = F - N
    pF = F - P
    n = m.normals
    nD = np.abs(nF[:, 0] * n[:, 0] + nF[:, 1] * n[:, 1] + nF[:, 2] * n[:, 2])
    pD = np.abs(pF[:, 0] * n[:, 0] + pF[:, 1] * n[:, 1] + pF[:, 2] * n[:, 2])
    m.weights = nD / (nD + pD)
    w1 = (-delta[:, 0] * pF[:, 0] + -delta[:, 1] * pF[:, 1]) + -delta[:, 2] * pF[:, 2]
    w2 = (delta[:, 0] * nF[:, 0] + delta[:, 1] * nF[:, 1]) + delta[:, 2] * nF[:, 2]
    dd = (delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1]) + delta[:, 2] * delta[:, 2]
    w1 = w1 / dd
    w2 = w2 / dd
    m.linearWeights = np.stack([w1 / 3, w2 / 3], 1)
    q0 = 2. / 3 * pF + 1. / 3 * (pF + w1[:, None] * delta)
    q1 = 2. / 3 * nF + 1. / 3 * (nF - w2[:, None] * delta)
    m.quadraticWeights = np.stack([q0, q1], 1)


def cell_neighbours(m):
    """cellOwner, cellNeighbours (full, ghost indices) and the matop variant
    (-1 for ghosts), cmesh.cpp:240-269."""
    c = np.arange(m.nInternalCells)[:, None]
    p = m.owner[m.cellFaces]
    n = m.neighbour[m.cellFaces]
    own = (p == c)
    m.cellOwner = own.astype(np.int32)
    m.cellNeighbours = np.where(own, n, p).astype(np.int32)
    m.cellNeighboursMatOp = np.where(own, np.where(n < m.nInternalCells, n, -1), p).astype(np.int32)


def build(meshdata, with_cells=False):
    """Mesh.build order, adFVM/mesh.py:83-127 with cmesh.build, cmesh.cpp:65-270."""
    m = OracleMesh()
    m.with_cells = with_cells
    m.points = np.asarray(meshdata.points, np.float64)
    m.faces = np.asarray(meshdata.faces, np.int32)
    m.owner = np.asarray(meshdata.owner, np.int32)
    m.neighbour = np.asarray(meshdata.neighbour, np.int32)
    m.boundary = {k: dict(v) for k, v in meshdata.boundary.items()}
    populate_sizes(m)
    build_before_write(m)
    split_patches(m)
    m.normals, m.faceCentres, m.areas = face_geometry(m.points, m.faces)
    m.cellCentres, m.volumes = cell_geometry(m.cellFaces, m.faceCentres, m.areas, m.normals)
    create_ghost_cells(m)
    face_weights(m)
    cell_neighbours(m)
    m.volumesL = m.volumes[m.owner]                               # mesh.py:108
    m.volumesR = m.volumes[m.neighbour[:m.nInternalFaces]]        # mesh.py:109
    m.nRemoteCells = m.nCells - m.nLocalCells                     # mesh.py:110
    m.nLocalFaces = m.nLocalCells - m.nInternalCells + m.nInternalFaces
    return m

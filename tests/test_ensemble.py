"""The batched ensemble: the concatenated mesh reproduces every instance's
own mesh arrays (CPU), and on the GPU every instance of the ensemble run
matches the same instance run alone, bit for bit, with and without HIP
graphs."""
import numpy as np
import pytest

from adfvm_amd import meshgen
from adfvm_amd.mesh import Mesh


def test_concatenated_mesh_reproduces_instances():
    ms = [meshgen.permute_and_rcm(meshgen.periodic_box(n), seed=n) for n in (4, 5, 3)]
    ms.append(meshgen.permute_and_rcm(meshgen.box((3, 2, 2)), seed=1))
    md, off = meshgen.concatenate(ms, ['a_', 'b_', 'c_', 'd_'])
    U = Mesh.from_data(md)
    assert list(off) == [0, 64, 189, 216, 228]
    nIF_all = U.nInternalFaces
    fi = 0
    fb = nIF_all
    for m, o, pre in zip(ms, off, ['a_', 'b_', 'c_', 'd_']):
        M = Mesh.from_data(m)
        nIC, nIF, nF = M.nInternalCells, M.nInternalFaces, M.nFaces
        nB = nF - nIF
        cells = slice(o, o + nIC)
        fint = slice(fi, fi + nIF)
        fbnd = slice(fb, fb + nB)
        for name in ('areas', 'normals', 'weights', 'deltas', 'deltasUnit', 'linearWeights', 'quadraticWeights'):
            a, b = getattr(U, name), getattr(M, name)
            assert np.array_equal(a[fint], b[:nIF]) and np.array_equal(a[fbnd], b[nIF:]), name
        assert np.array_equal(U.volumes[cells], M.volumes)
        assert np.array_equal(U.cellCentres[cells], M.cellCentres[:nIC])
        # slot tables: same faces in the same order, renumbered
        fmap = np.concatenate([np.arange(fi, fi + nIF), np.arange(fb, fb + nB)])
        assert np.array_equal(U.cellFaces[cells], fmap[M.cellFaces])
        assert np.array_equal(U.cellOwner[cells], M.cellOwner)
        for name, b in M.boundary.items():
            ub = U.boundary[pre + name]
            assert ub['nFaces'] == b['nFaces'] and ub['type'] == b['type']
            assert ub['startFace'] - fb == b['startFace'] - nIF
            if 'neighbourPatch' in b:
                assert ub['neighbourPatch'] == pre + b['neighbourPatch']
        fi += nIF
        fb += nB


@pytest.mark.gpu
@pytest.mark.parametrize('graphs', [False, True])
def test_ensemble_matches_instances_bitwise(graphs):
    torch = pytest.importorskip('torch')
    from adfvm_amd.ensemble import TaylorGreenEnsemble
    from adfvm_amd.context import Context
    from adfvm_amd import adjoint, _lib, config
    ens = TaylorGreenEnsemble(n_instances=5, sizes=(6, 8, 10), seed=3, precision=np.float32)
    print('ensemble sizes', ens.sizes, 'cells', ens.n_cells, flush=True)
    n_steps, interval = 9, 3                     # three replays of each graph
    J, ua = ens.run(n_steps=n_steps, interval=interval, graphs=graphs)
    assert J.shape == (5,) and np.all(J > 0)
    ua = ua.cpu().numpy()
    for k, n in enumerate(ens.sizes):
        md = meshgen.permute_and_rcm(meshgen.periodic_box(n), seed=3 * 100003 + k)
        ph = dict(gamma=1.4, Cp=1004.5, Pr=0.7, mu=(_lib.MU_CONSTANT, ens.mu, 0., 0.), riemann=_lib.RIEMANN_ROE,
                  boundaryRiemann=_lib.RIEMANN_ROE, small=config.small_for(np.float32), integrator=_lib.RK_SSPRK3,
                  objective=_lib.OBJ_KINETIC_ENERGY)
        ctx = Context(Mesh.from_data(md), {}, ph, np.float32)
        o0, o1 = ens.offsets[k], ens.offsets[k + 1]
        u0 = ens.u0[:, o0:o1].contiguous()
        u = u0.clone()
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        Jk = torch.zeros(1, dtype=torch.float64, device='cuda')
        for _ in range(n_steps):
            ctx.primal_step(u, ens.dt, obj_out=obj)
            Jk += obj
        uak = adjoint.sweep(ctx, u0, [ens.dt] * n_steps, schedule='uniform', interval=interval)
        assert J[k] == Jk.item(), (k, J[k], Jk.item())
        assert np.array_equal(ua[:, o0:o1], uak.cpu().numpy()), k

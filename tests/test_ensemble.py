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
@pytest.mark.parametrize('mode', ['eager', 'segment_graphs', 'whole_graph'])
def test_ensemble_matches_instances_bitwise(mode):
    torch = pytest.importorskip('torch')
    from adfvm_amd.ensemble import TaylorGreenEnsemble
    from adfvm_amd.context import Context
    from adfvm_amd import adjoint, _lib, config
    ens = TaylorGreenEnsemble(n_instances=5, sizes=(6, 8, 10), seed=3, precision=np.float32)
    print('ensemble sizes', ens.sizes, 'cells', ens.n_cells, flush=True)
    n_steps, interval = 9, 3                     # three replays of each segment graph
    graphs, whole = mode != 'eager', mode == 'whole_graph'
    J, ua = ens.run(n_steps=n_steps, interval=interval, graphs=graphs, whole=whole)
    if whole:
        # a second replay of the captured sweep gives the same bits
        ua = ua.clone()
        J2, ua2 = ens.run(n_steps=n_steps, interval=interval, graphs=graphs, whole=whole)
        assert np.array_equal(J, J2) and torch.equal(ua, ua2)
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


@pytest.mark.gpu
def test_segment_objectives_zero_without_objective():
    """A context of concatenated meshes with cell segments and no objective
    writes one zero per segment into obj_out (every entry, not only the
    first)."""
    torch = pytest.importorskip('torch')
    from adfvm_amd.context import Context
    from adfvm_amd import _lib, config
    ms = [meshgen.permute_and_rcm(meshgen.periodic_box(n), seed=n) for n in (4, 5, 3)]
    md, off = meshgen.concatenate(ms)
    ph = dict(gamma=1.4, Cp=1004.5, Pr=0.7, mu=(_lib.MU_CONSTANT, 1e-3, 0., 0.), riemann=_lib.RIEMANN_ROE,
              boundaryRiemann=_lib.RIEMANN_ROE, small=config.small_for(np.float64), integrator=_lib.RK_SSPRK3,
              objective=_lib.OBJ_NONE)
    ctx = Context(Mesh.from_data(md), {}, ph, np.float64)
    ctx.set_segments(off)
    from tests import casefactory as cf
    m = ctx.mesh
    U = np.zeros((ctx.nIC, 3)) + [10., 2., 0.]
    u = torch.as_tensor(cf.state_from_prims(U, np.full((ctx.nIC, 1), 300.), np.full((ctx.nIC, 1), 1e5)),
                        device='cuda').contiguous()
    obj = torch.full((3,), 7.0, dtype=torch.float64, device='cuda')
    ctx.primal_step(u, 1e-6, obj_out=obj)
    assert obj.cpu().tolist() == [0.0, 0.0, 0.0]
    obj.fill_(7.0)
    ctx.objective(u, out=obj)
    assert obj.cpu().tolist() == [0.0, 0.0, 0.0]


@pytest.mark.gpu
def test_ensemble_full_size_workload():
    """The batched-ensemble workload at its full size: 64 Taylor-Green
    instances, sizes drawn by seed 0 from {32, 48, 64, 96}^3, each with a
    seeded +-1% velocity perturbation, 200 steps primal + adjoint of each
    instance's summed kinetic energy in FP32 (FP64 objectives), checkpoints
    every 20 steps, the whole sweep one captured HIP graph.  Checks:
    (1) every instance conserves sum V rho, sum V rhoU, sum V rhoE over the
        200 steps to FP32 round-off (relative <= 1e-6);
    (2) the first 32^3 instance's state after 200 steps and its objective
        against the oracle (torch FP64 on the GPU, same FP32-rounded initial
        data): relative L2 <= 1e-5 per conserved variable, summed objective
        <= 2e-5;
    (3) the smallest and the largest drawn instance, each in an FP64 context
        of its own: the adjoint along a seeded random momentum perturbation
        against central differences (eps = 1e-6) of the summed objective,
        <= 1e-4; and the ensemble's FP32 gradient of that instance against
        that FP64 adjoint, relative L2 <= 1e-4 per conserved variable."""
    torch = pytest.importorskip('torch')
    from adfvm_amd.ensemble import TaylorGreenEnsemble
    from adfvm_amd.graphs import GraphSweep
    from adfvm_amd.context import Context
    from adfvm_amd import adjoint, _lib, config
    from tests import casefactory as cf
    n_steps, interval = 200, 20
    ens = TaylorGreenEnsemble(n_instances=64, seed=0, precision=np.float32)
    sizes = ens.sizes
    print('ensemble: %d cells, sizes %s' % (ens.n_cells, {s: sizes.count(s) for s in sorted(set(sizes))}), flush=True)
    sweep = GraphSweep(ens.ctx, ens.dt, interval, 1.0, True, True, whole=True)
    sweep.keep_final = True
    J, ua = sweep.run(ens.u0, n_steps)
    assert np.isfinite(J).all() and bool(torch.isfinite(ua).all())
    u0 = ens.u0.double().cpu().numpy()
    uN = sweep.u_final.double().cpu().numpy()
    ua = ua.double().cpu().numpy()
    del sweep
    torch.cuda.empty_cache()
    V = ens.mesh.volumes[:, 0]
    off = ens.offsets
    # (1) conservation per instance
    worst = [0.0, 0.0, 0.0]
    for k in range(len(sizes)):
        c = slice(off[k], off[k + 1])
        for j, rows in enumerate((slice(0, 1), slice(1, 4), slice(4, 5))):
            a = (u0[rows, c] * V[c]).sum(1)
            b = (uN[rows, c] * V[c]).sum(1)
            scale = (np.abs(u0[rows, c]) * V[c]).sum()
            worst[j] = max(worst[j], float(np.abs(b - a).max() / scale))
    print('conservation: worst relative change rho/rhoU/rhoE %s' % ' '.join('%.2e' % w for w in worst), flush=True)
    assert max(worst) <= 1e-6
    # (2) one 32^3 instance against the oracle
    k32 = sizes.index(32) if 32 in sizes else sizes.index(min(sizes))
    n = sizes[k32]
    md = meshgen.permute_and_rcm(meshgen.periodic_box(n), seed=0 * 100003 + k32)
    oc = cf.CaseDef(md, {}, ens.mu, objective=('kinetic_energy',)).oracle().to('cuda')
    c = slice(off[k32], off[k32 + 1])
    x = torch.as_tensor(u0[:, c], device='cuda')
    Jo = 0.0
    with torch.no_grad():
        for _ in range(n_steps):
            r, rU, rE, _, obj = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, ens.dt)
            x = torch.cat([r.T, rU.T, rE.T])
            Jo += float(obj)
    xo = x.cpu().numpy()
    del oc, x
    torch.cuda.empty_cache()
    errs = [float(np.linalg.norm(uN[s, c] - xo[s]) / np.linalg.norm(xo[s])) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
    print('instance %d (%d^3) vs oracle: rel-L2 %s objective %.2e' % (k32, n, errs, abs(J[k32] - Jo) / Jo), flush=True)
    assert max(errs) <= 1e-5 and abs(J[k32] - Jo) <= 2e-5 * abs(Jo)
    # (3) FP64 adjoint vs central differences on the smallest and largest instance
    ph = dict(gamma=1.4, Cp=1004.5, Pr=0.7, mu=(_lib.MU_CONSTANT, ens.mu, 0., 0.), riemann=_lib.RIEMANN_ROE,
              boundaryRiemann=_lib.RIEMANN_ROE, small=config.small_for(np.float64), integrator=_lib.RK_SSPRK3,
              objective=_lib.OBJ_KINETIC_ENERGY)
    for k in sorted({sizes.index(min(sizes)), sizes.index(max(sizes))}):
        n = sizes[k]
        md = meshgen.permute_and_rcm(meshgen.periodic_box(n), seed=0 * 100003 + k)
        ctx = Context(Mesh.from_data(md), {}, ph, np.float64)
        c = slice(off[k], off[k + 1])
        u0k = torch.as_tensor(u0[:, c], device='cuda').contiguous()
        ua64 = adjoint.sweep(ctx, u0k, [ens.dt] * n_steps, schedule='uniform', interval=interval).cpu().numpy()
        rng = np.random.default_rng(100 + k)
        v = np.zeros_like(u0[:, c])
        v[1:4] = u0[0:1, c] * rng.normal(size=(3, c.stop - c.start))
        vt = torch.as_tensor(v, device='cuda')

        def Jof(e):
            u = (u0k + e * vt).contiguous()
            obj = torch.zeros(1, dtype=torch.float64, device='cuda')
            tot = torch.zeros(1, dtype=torch.float64, device='cuda')
            for _ in range(n_steps):
                ctx.primal_step(u, ens.dt, obj_out=obj)
                tot += obj
            return tot.item()
        eps = 1e-6
        fd = (Jof(eps) - Jof(-eps)) / (2 * eps)
        ad = float((ua64 * v).sum())
        g32 = ua[:, c]
        e32 = [float(np.linalg.norm(g32[s] - ua64[s]) / np.linalg.norm(ua64[s])) for s in (slice(0, 1), slice(1, 4), slice(4, 5))]
        print('instance %d (%d^3): adjoint %.12e fd %.12e rel %.2e; ensemble FP32 gradient vs FP64 rel-L2 %s'
              % (k, n, ad, fd, abs(ad - fd) / abs(fd), e32), flush=True)
        assert abs(ad - fd) <= 1e-4 * abs(fd)
        assert max(e32) <= 1e-4
        del ctx

"""The drop-in boundary at the reference's positional contract (the primal
Function, density.py:97-105 called at solver.py:312-323; primal_grad,
apps/adjoint.py:94-126 called at :268-291) and case-file objectives
(objective(fields, solver), density.py:354-357, templates/cylinder_test.py:
9-37) on the GPU, each against the Context entry points and the oracle."""
import os

import numpy as np
import pytest

from adfvm_amd import config
from tests import casefactory as cf
from tests.test_gpu_primal import rel_l2

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(autouse=True)
def _restore_precision():
    old = config.precision
    yield
    config.set_precision(old)


def positional_inputs(s, u, dt, source=None, overrides=None):
    """The reference's primal input list built from an RCF: state arrays,
    dt, mesh tensors and scalars, the 3 source arrays, then the boundary
    values in the reference's order (field, sorted patch, BC input)."""
    from adfvm_amd.density import bc_inputs
    from adfvm_amd import BCs
    m = s.mesh
    nIC = m.nInternalCells
    src = source if source is not None else np.zeros((5, nIC))
    specs = BCs.validate(m, s.boundaries)
    bcs = []
    for field, patch, key, k in bc_inputs(s, specs):
        v = BCs.parse_value(specs[patch][field][key], m.boundary[patch]['nFaces'], k)
        if overrides and (field, patch, key) in overrides:
            v = overrides[(field, patch, key)]
        bcs.append(v)
    return ([u[0:1].T.copy(), u[1:4].T.copy(), u[4:5].T.copy(), np.array([[dt]])] + m.getTensor() + m.getScalar()
            + [src[0:1].T.copy(), src[1:4].T.copy(), src[4:5].T.copy()] + bcs)


def test_bc_inputs_follow_the_reference_order():
    """U, T, p; patches sorted by name; BC inputs in creation order."""
    case, u, _ = cf.bc_param_cases()['upt_drag']()
    s = cf.rcf(case, u, np.float64)
    from adfvm_amd.density import bc_inputs
    got = [(f, p, k) for f, p, k, _ in bc_inputs(s)]
    assert got == [('U', 'inlet', 'value'), ('U', 'top', 'value'), ('T', 'inlet', 'value'),
                   ('p', 'outlet', 'U0'), ('p', 'outlet', 'T0'), ('p', 'outlet', 'p0')]


def test_primal_map_uses_positional_source_and_bc_values():
    """RCF.map with the reference's positional list: with the solver's own
    values it equals Context.primal_step bitwise; a boundary value changed
    in the list (the inlet velocity) and a source passed in the list are
    the ones used -- bitwise the step after set_bc_value / with that source,
    and within 1e-12 of the oracle with the same values; a changed mesh
    array raises (mesh perturbations are not supported), so does a wrong
    argument count."""
    case, u, _ = cf.bc_param_cases()['upt_drag']()
    s = cf.rcf(case, u, np.float64)
    ctx = s.ctx
    dt = 1e-6
    rng = np.random.default_rng(3)
    ud = torch.as_tensor(u, device='cuda').contiguous()
    # 1. own values
    out = s.map(*positional_inputs(s, u, dt))
    ref = ud.clone()
    dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
    obj = torch.zeros(1, dtype=torch.float64, device='cuda')
    ctx.primal_step(ref, dt, dtc_out=dtc, obj_out=obj)
    got = np.concatenate([out[0].T, out[1].T, out[2].T])
    assert np.array_equal(got, ref.cpu().numpy())
    assert out[3][0, 0] == dtc.item() and out[4][0, 0] == obj.item()
    # 2. a perturbed inlet velocity and a source, through the list
    n_in = s.mesh.boundary['inlet']['nFaces']
    Uin = np.array([30., 2., 0.]) + rng.normal(size=(n_in, 3))
    src = rng.normal(size=u.shape) * np.abs(u).mean(1, keepdims=True) * 10
    out = s.map(*positional_inputs(s, u, dt, source=src, overrides={('U', 'inlet', 'value'): Uin}))
    got = np.concatenate([out[0].T, out[1].T, out[2].T])
    ref = ud.clone()
    from adfvm_amd.context import Context
    ctx2 = case.context(np.float64)
    ctx2.set_bc_value('U', 'inlet', 'value', Uin)
    ctx2.primal_step(ref, dt, source=torch.as_tensor(src, device='cuda').contiguous())
    assert np.array_equal(got, ref.cpu().numpy())
    bcs = {k: dict(v) for k, v in case.bcs.items()}
    bcs['inlet'] = dict(bcs['inlet'])
    bcs['inlet']['U'] = {'type': 'fixedValue', 'value': Uin}
    oc = case.oracle(bcs=bcs)
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    gi = [t(a) for a in cf.ghost_init(case)]
    r, rU, rE, _, _ = oc.step(t(u[0:1].T), t(u[1:4].T), t(u[4:5].T), dt, source=[t(src[0:1].T), t(src[1:4].T),
                                                                                  t(src[4:5].T)], ghostInit=gi)
    want = np.concatenate([r.T.numpy(), rU.T.numpy(), rE.T.numpy()])
    for sl in (slice(0, 1), slice(1, 4), slice(4, 5)):
        assert rel_l2(got[sl], want[sl]) < 1e-12
    # the context keeps the last values passed; the solver's spec follows
    assert np.array_equal(np.asarray(s.boundaries['U']['inlet']['value']), Uin)
    # 3. errors
    args = positional_inputs(s, u, dt)
    args[4 + 5] = args[4 + 5] * 1.01                    # 'normals' (gradFields order)
    with pytest.raises(NotImplementedError):
        s.map(*args)
    with pytest.raises(ValueError):
        s.map(*positional_inputs(s, u, dt)[:-1])


@pytest.mark.parametrize('param', ['source', 'mesh', ('BCs', 'U', 'inlet', 'value'), ('BCs', 'p', 'outlet', 'p0')])
def test_adjoint_map_positional_contract(param):
    """Adjoint.map (primal_grad) with the reference's positional list:
    adjoint fields and the parameter gradient (the three source arrays, the
    ten Mesh.gradFields arrays, or the BC value array) equal
    Context.adjoint_step's bitwise; the gradient accumulates over calls
    until zero_static; a non-zero dtc seed raises."""
    from adfvm_amd.adjoint import Adjoint
    case, u, _ = cf.bc_param_cases()['upt_drag']()
    s = cf.rcf(case, u, np.float64)
    ctx = s.ctx
    adj = Adjoint(s).compileSolver((param,))
    dt = 1e-6
    rng = np.random.default_rng(4)
    ua = rng.normal(size=u.shape) * np.abs(u).mean(1, keepdims=True)
    src = rng.normal(size=u.shape) * np.abs(u).mean(1, keepdims=True) * 10 if param == 'source' else None
    pin = positional_inputs(s, u, dt, source=src)
    tail = [ua[0:1].T.copy(), ua[1:4].T.copy(), ua[4:5].T.copy(), np.zeros((1, 1)), np.ones((1, 1)), np.zeros((1, 1))]
    out1 = adj.map(*(pin + tail), return_static=False, zero_static=False)
    out2 = adj.map(*(pin + tail))
    ud = torch.as_tensor(u, device='cuda').contiguous()
    uad = torch.as_tensor(ua, device='cuda').contiguous()
    sg, bg = ctx.state(0.), ctx.bc_state()
    mg = ctx.mesh_grad_state() if param == 'mesh' else None
    sd = torch.as_tensor(src, device='cuda').contiguous() if src is not None else None
    for _ in range(2):
        ua_i = uad.clone()
        ctx.adjoint_step(ud, dt, ua_i, obja=1.0, source=sd, source_grad=sg, bc_grad=bg)
    if mg is not None:
        ctx.set_mesh_grad(None)
    got = np.concatenate([out2[0].T, out2[1].T, out2[2].T])
    assert np.array_equal(got, ua_i.cpu().numpy())
    assert out1[3] is None
    if param == 'mesh':
        assert len(out2) == 3 + 10 and all(a is None for a in out1[3:])
        for a, b in zip(out2[3:], ctx.mesh_grad_arrays(mg)):
            assert a.shape == tuple(b.shape) and np.array_equal(a, b.cpu().numpy())
        assert max(np.abs(a).max() for a in out2[3:]) > 0
    elif param == 'source':
        g = sg.cpu().numpy()
        assert np.array_equal(out2[3], g[0:1].T) and np.array_equal(out2[4], g[1:4].T) and \
            np.array_equal(out2[5], g[4:5].T)
    else:
        rows, faces = ctx.bc_param(*param[1:])
        assert np.array_equal(out2[3], bg[rows, faces].cpu().numpy().T)
        assert np.abs(out2[3]).max() > 0
    bad = tail[:3] + [np.ones((1, 1))] + tail[4:]
    with pytest.raises(NotImplementedError):
        adj.map(*(pin + bad))


def test_primitive_ghosts_vjp_against_autograd():
    """(d prim/d u)^T pbar of adfvm_primitive_ghosts_vjp against the
    oracle's autograd of its primitive + ghost map (every ghost rule on the
    mixed boxes, cyclic on the Taylor-Green box), and the boundary-value
    part against autograd with respect to the values: FP64 <= 1e-12."""
    cases_ = {'upt_drag': cf.bc_param_cases()['upt_drag'](), 'totalpt_direction': cf.bc_param_cases()['totalpt_direction']()}
    tg, pr = cf.tgv(5, seed=2)
    cases_['tgv'] = (tg, cf.tgv_state(tg, pr), [])
    for name, (case, u, params) in sorted(cases_.items()):
        nC = case.mesh.nCells
        pbar = np.random.default_rng(9).normal(size=(5, nC))
        bcs, ts = cf.bc_value_tensors(case, params) if params else (case.bcs, {})
        oc = case.oracle(bcs=bcs)
        x = torch.tensor(u, dtype=torch.float64, requires_grad=True)
        gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
        U, T, p = oc.primitive(x[0:1].T, x[1:4].T, x[4:5].T)
        U, T, p = oc.ghosts(U, T, p, gi)
        prim = torch.cat([U.T, T.T, p.T])
        L = (prim * torch.as_tensor(pbar)).sum()
        leaves = [x] + [ts[q[:3]] for q in params]
        grads = torch.autograd.grad(L, leaves)
        ctx = case.context(np.float64)
        ud = torch.as_tensor(u, device='cuda').contiguous()
        pb = torch.as_tensor(pbar, device='cuda').contiguous()
        bg = ctx.bc_state()
        ub = ctx.primitive_ghosts_vjp(ud, pb, bc_grad=bg).cpu().numpy()
        e = rel_l2(ub, grads[0].numpy())
        print('%s prim-ghost VJP rel-L2 %.2e' % (name, e))
        assert e < 1e-12
        bgn = bg.cpu().numpy()
        for q, g in zip(params, grads[1:]):
            rows, faces = ctx.bc_param(*q[:3])
            e = rel_l2(bgn[rows, faces].T, g.numpy())
            print('%s %s rel-L2 %.2e' % (name, q[:3], e))
            assert e < 1e-12


def _cylinder_solvers(objective_fn, n_theta=64, n_r=32):
    from adfvm_amd import cases, objectives
    from adfvm_amd.density import RCF
    config.set_precision(np.float64)
    st = cases.cylinder(n_theta, n_r, ratio=1.1)
    out = []
    for obj in (objectives.drag('cylinder'), objective_fn):
        s = RCF(st.md, mu=st.mu, boundaryRiemannSolver='eulerLaxFriedrichs', objective=obj, fixedTimeStep=True)
        s.setFields(st.U, st.T, st.p, boundary=st.boundary)
        s.compile()
        out.append(s)
    return st, out


def test_case_file_drag_objective_matches_builtin_kernel(tmp_path, monkeypatch):
    """The reference's objectiveDrag (templates/cylinder_test.py:9-19),
    written as a torch function of the primitive fields
    (adfvm_amd/templates/cylinder_objective_test.py), against the library's
    drag kernel on the 64 x 32 cylinder: the objective on every step's
    stage-1 state, and the adjoint sweep (adjoint state and the wall
    velocity's boundary-value gradient), FP64 relative <= 1e-12."""
    import importlib.util
    from adfvm_amd import adjoint
    monkeypatch.setenv('ADFVM_CASE_DIR', str(tmp_path / 'cylinder'))
    monkeypatch.chdir(tmp_path)
    spec = importlib.util.spec_from_file_location(
        'cyl_obj', os.path.join(REPO, 'adfvm_amd', 'templates', 'cylinder_objective_test.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    st, (sb, sp) = _cylinder_solvers(mod.objectiveDrag)
    dt, n = 2e-6, 6
    Js = []
    for s in (sb, sp):
        u = s.u.clone()
        dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
        obj = torch.zeros(1, dtype=torch.float64, device='cuda')
        J = []
        for _ in range(n):
            s.step(u, dt, None, dtc, obj)
            J.append(obj.item())
        Js.append(np.array(J))
    print('drag per step', Js[0], Js[1])
    assert np.abs(Js[0] - Js[1]).max() <= 1e-12 * np.abs(Js[0]).max()
    res = []
    for s in (sb, sp):
        bg = s.ctx.bc_state()
        ua = adjoint.sweep(s.ctx, s.u, [dt] * n, schedule='store_all', bc_grad=bg, solver=s)
        rows, faces = s.ctx.bc_param('U', 'cylinder', 'value')
        res.append((ua.cpu().numpy(), bg[rows, faces].cpu().numpy()))
    for k in range(2):
        e = rel_l2(res[1][k], res[0][k])
        print('adjoint %d rel-L2 %.2e' % (k, e))
        assert e <= 1e-12 and np.abs(res[0][k]).max() > 0
    # tapes off: the objective stage is recomputed
    ua = adjoint.sweep(sp.ctx, sp.u, [dt] * n, schedule='store_all', tapes=False, solver=sp)
    assert rel_l2(ua.cpu().numpy(), res[0][0]) <= 1e-12


def test_apps_case_file_objective_fd_vs_adjoint(tmp_path, monkeypatch):
    """The reference's test_adjoint workflow (orig, perturb, adjoint through
    the apps) with an objective that has no library kernel: the area-mean
    wall pressure on the cylinder, written in the case file as a torch
    function (adfvm_amd/templates/cylinder_objective_test.py with
    ADFVM_CYL_OBJECTIVE=pressure); |fd - adj| / |fd| < 1e-3."""
    from adfvm_amd.apps import problem, adjoint as adjoint_app
    case = str(tmp_path / 'cylinder')
    monkeypatch.setenv('ADFVM_CASE_DIR', case)
    monkeypatch.setenv('ADFVM_CYL_OBJECTIVE', 'pressure')
    monkeypatch.chdir(tmp_path)
    tpl = os.path.join(REPO, 'adfvm_amd', 'templates', 'cylinder_objective_test.py')
    for app, args in ((problem, [tpl, 'orig']), (problem, [tpl, 'perturb']), (adjoint_app, [tpl])):
        print('running', app.__name__, ' '.join(args[1:]), flush=True)
        app.main(args + ['--gpu_double'])
    with open(os.path.join(case, 'objective.txt')) as f:
        data = f.readlines()
    print(''.join(data))
    fdSens = float(data[1].split(' ')[-2])
    adjSens = float(data[2].split(' ')[-2])
    assert fdSens != 0
    assert abs(fdSens - adjSens) / abs(fdSens) < 1e-3

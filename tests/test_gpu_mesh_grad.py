"""Mesh-geometry gradients of the adjoint step (the parameter 'mesh' of
apps/adjoint.py:107-110: the step differentiated with respect to each array
of Mesh.gradFields, mesh.py:28-32) through the C ABI, against the oracle's
autograd with the same arrays as independent leaves (oracle
Case.mesh_leaves, itself checked against central differences in
tests/test_oracle.py).  FP64 relative L2 per array <= 1e-9; FP32 <= 2e-3
(the FP32 geometry is the FP64 geometry rounded)."""
import numpy as np
import pytest

from tests import casefactory as cf
from tests.test_gpu_primal import CASES, rel_l2
from tests.test_gpu_adjoint import scaled_rbar, adjoint_case

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def oracle_mesh_vjp(case, u, dt, ua, obja):
    """d(<ua, step(u)> + obja J)/d(mesh array) for each of the ten arrays."""
    oc = case.oracle()
    leaves = oc.mesh_leaves()
    x = torch.tensor(u, dtype=torch.float64)
    gi = [torch.as_tensor(a, dtype=torch.float64) for a in cf.ghost_init(case)]
    r, rU, rE, dtc, obj = oc.step(x[0:1].T, x[1:4].T, x[4:5].T, dt, ghostInit=gi)
    L = (torch.cat([r.T, rU.T, rE.T]) * torch.as_tensor(ua)).sum() + obja * obj
    grads = torch.autograd.grad(L, leaves, allow_unused=True)
    return [g.numpy() if g is not None else np.zeros(tuple(l.shape)) for g, l in zip(grads, leaves)]


@pytest.mark.parametrize('name', sorted(CASES))
def test_mesh_gradient_matches_autograd(name):
    from adfvm_amd.context import Context
    case, u = adjoint_case(name)
    ua = scaled_rbar(case, u, 3) * 1e3
    dt = 1e-6 if name.startswith('mixed') else 2e-5
    obja = 1.0 if case.objective else 0.0
    ref = oracle_mesh_vjp(case, u, dt, ua, obja)
    for prec, tol in ((np.float64, 1e-9), (np.float32, 2e-3)):
        ctx = case.context(prec)
        ud = torch.as_tensor(u.astype(prec), device='cuda').contiguous()
        uad = torch.as_tensor(ua.astype(prec), device='cuda').contiguous()
        g = ctx.mesh_grad_state()
        ctx.adjoint_step(ud, dt, uad, obja=obja)
        arrays = [a.cpu().numpy().astype(np.float64) for a in ctx.mesh_grad_arrays(g)]
        ctx.set_mesh_grad(None)
        errs = {}
        for fname, got, want in zip(Context.MESH_GRAD_FIELDS, arrays, ref):
            assert got.shape == want.shape, (fname, got.shape, want.shape)
            nw = np.linalg.norm(want)
            errs[fname] = rel_l2(got, want) if nw > 0 else float(np.linalg.norm(got))
        print('%s %s mesh-gradient rel-L2: %s' % (name, np.dtype(prec).name,
              ' '.join('%s %.1e' % (k, v) for k, v in errs.items())), flush=True)
        assert max(errs.values()) < tol, errs


def test_mesh_gradient_off_leaves_the_adjoint_unchanged():
    """Registering a mesh-gradient array changes nothing else: the adjoint
    state of the step is bitwise the same with and without it, and a second
    step adds to the array."""
    case, u = adjoint_case('mixed_drag')
    ua = scaled_rbar(case, u, 4) * 1e3
    ctx = case.context(np.float64)
    ud = torch.as_tensor(u, device='cuda').contiguous()
    a1 = torch.as_tensor(ua, device='cuda').contiguous()
    ctx.adjoint_step(ud, 1e-6, a1, obja=1.0)
    g = ctx.mesh_grad_state()
    a2 = torch.as_tensor(ua, device='cuda').contiguous()
    ctx.adjoint_step(ud, 1e-6, a2, obja=1.0)
    assert torch.equal(a1, a2)
    g1 = g.clone()
    a3 = torch.as_tensor(ua, device='cuda').contiguous()
    ctx.adjoint_step(ud, 1e-6, a3, obja=1.0)
    assert torch.allclose(g, 2 * g1, rtol=1e-12, atol=0)
    ctx.set_mesh_grad(None)

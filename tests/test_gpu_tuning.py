"""Kernel block sizes (adfvm_autotune / adfvm_set_kernel_block) and the
streaming policy (adfvm_set_streaming) change launch shapes and cache hints
only: the step and its adjoint are bitwise identical under every block size
and either policy, and the autotuner picks one of the offered sizes."""
import numpy as np
import pytest

from tests import casefactory as cf
from tests.test_gpu_primal import CASES

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

KERNELS = ['grad_kernel', 'face_flux_kernel', 'stage_kernel', 'face_vjp_kernel', 'vjp_gather_kernel',
           'grad_vjp_kernel']


def step_and_adjoint(ctx, u, dt):
    a = u.clone()
    dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
    obj = torch.zeros(1, dtype=torch.float64, device='cuda')
    ctx.primal_step(a, dt, dtc_out=dtc, obj_out=obj)
    ua = torch.ones_like(u) * 1e-3
    ctx.adjoint_step(u, dt, ua, obja=1.0)
    return a.cpu().numpy(), ua.cpu().numpy(), dtc.item(), obj.item()


@pytest.mark.parametrize('name', ['tgv6', 'mixed_upt'])
@pytest.mark.parametrize('prec', [np.float32, np.float64])
def test_block_sizes_bitwise(name, prec):
    case, u0 = CASES[name]()
    ctx = case.context(prec)
    u = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
    dt = 1e-6
    ref = step_and_adjoint(ctx, u, dt)
    for bs in (64, 128, 512):
        for k in KERNELS:
            ctx.set_kernel_block(k, bs)
        got = step_and_adjoint(ctx, u, dt)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), bs
        assert got[2] == ref[2] and got[3] == ref[3]
    cfg = ctx.autotune(u, reps=2)
    assert set(cfg) == set(KERNELS)
    for k, (bs, ms) in cfg.items():
        assert bs in (64, 128, 256, 512) and ms > 0
    got = step_and_adjoint(ctx, u, dt)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    with pytest.raises(Exception):
        ctx.set_kernel_block('grad_kernel', 96)


@pytest.mark.parametrize('name', ['tgv6', 'mixed_upt', 'mixed_totalpt_lf', 'mixed_drag'])
@pytest.mark.parametrize('prec', [np.float32, np.float64])
def test_streaming_policy_bitwise(name, prec):
    """Non-temporal loads of single-use rows (on or off) leave every result
    bit for bit the same: primal step, dtc, objective and adjoint step
    (cyclic, ghost and characteristic faces, the drag objective).  Small meshes default
    to plain loads (their stage working set fits the Infinity Cache)."""
    case, u0 = CASES[name]()
    ctx = case.context(prec)
    assert ctx.streaming is False
    u = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
    dt = 1e-6
    ref = step_and_adjoint(ctx, u, dt)
    ctx.set_streaming(True)
    assert ctx.streaming is True
    got = step_and_adjoint(ctx, u, dt)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert got[2] == ref[2] and got[3] == ref[3]
    ctx.set_streaming(None)
    assert ctx.streaming is False
    assert ctx.lib.adfvm_set_streaming(ctx.handle, 2) != 0      # modes are -1, 0, 1


def test_streaming_policy_automatic_choice():
    """The automatic policy streams when one stage's working set (face
    geometry, face tables and flux records; cell state, gradient record,
    slot geometry, volume, stage states and slot tables) exceeds 256 MB:
    TGV 64^3 is about 159 MB in FP32 (plain loads) and 303 MB in FP64
    (streamed).  Either policy gives the same step bit for bit there too."""
    case, params = cf.tgv(64, seed=0)
    u0 = cf.tgv_state(case, params)
    for prec, want in ((np.float32, False), (np.float64, True)):
        ctx = case.context(prec)
        assert ctx.streaming is want
        u = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
        ref = step_and_adjoint(ctx, u, 1e-5)
        ctx.set_streaming(not want)
        got = step_and_adjoint(ctx, u, 1e-5)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        del ctx

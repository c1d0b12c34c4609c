"""adfvm_dot_noalloc, the allocation-free variant of adfvm_dot, called
through the C ABI: its result equals adfvm_dot's bit for bit (same kernels,
same partition into block partials, same fixed folding order), on two calls
in a row whose second input is smaller, in FP32 and FP64; sizes beyond its
maximum are refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from tests import casefactory as cf

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('prec', [np.float32, np.float64])
def test_dot_noalloc_matches_dot_on_two_calls(prec):
    from adfvm_amd.context import dot_noalloc
    case, params = cf.tgv(12, seed=7)
    ctx = case.context(prec)
    g = torch.Generator(device='cpu').manual_seed(11)
    a = torch.randn(5, ctx.nIC, generator=g, dtype=torch.float64).to(ctx.tdtype).cuda().contiguous()
    b = torch.randn(5, ctx.nIC, generator=g, dtype=torch.float64).to(ctx.tdtype).cuda().contiguous()
    out = torch.zeros(1, dtype=torch.float64, device='cuda')
    ref = torch.zeros(1, dtype=torch.float64, device='cuda')
    exact = float((a.double() * b.double()).sum())
    # first call: the whole state-shaped arrays
    before = torch.cuda.memory_allocated()
    dot_noalloc(a, b, out)
    torch.cuda.synchronize()
    assert torch.cuda.memory_allocated() == before
    ctx.dot(a, b, out=ref)
    print('%s n=%d: noalloc %.17g dot %.17g' % (np.dtype(prec).name, a.numel(), out.item(), ref.item()), flush=True)
    assert out.item() == ref.item()
    assert abs(out.item() - exact) <= 1e-12 * (a.double().abs() * b.double().abs()).sum().item()
    # second call, smaller input: the partials of the first call are still in
    # the buffer; the result depends only on this call's data
    m = 777
    a2, b2 = a.reshape(-1)[:m].contiguous(), b.reshape(-1)[:m].contiguous()
    dot_noalloc(a2, b2, out)
    ctx.dot(a2, b2, out=ref)
    print('%s n=%d: noalloc %.17g dot %.17g' % (np.dtype(prec).name, m, out.item(), ref.item()), flush=True)
    assert out.item() == ref.item()
    # an empty input gives 0
    dot_noalloc(a2[:0], b2[:0], out)
    assert out.item() == 0.0


def test_dot_noalloc_refuses_sizes_beyond_its_maximum():
    """n = 2^27 + 1 is refused by the host-side size check, which returns
    before any kernel is launched (the pointers are never dereferenced)."""
    from adfvm_amd import _lib
    lib = _lib.load()
    a = torch.zeros(4, dtype=torch.float32, device='cuda')
    out = torch.zeros(1, dtype=torch.float64, device='cuda')
    p = C.c_void_p(a.data_ptr())
    rc = lib.adfvm_dot_noalloc(4, p, p, (1 << 27) + 1, C.c_void_p(out.data_ptr()), C.c_void_p(None))
    assert rc == 1
    assert b'exceeds the maximum' in lib.adfvm_noalloc_last_error()
    rc = lib.adfvm_dot_noalloc(2, p, p, 4, C.c_void_p(out.data_ptr()), C.c_void_p(None))
    assert rc == 1

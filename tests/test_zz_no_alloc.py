"""vgc_mlp_grad_noalloc (static scratch, no allocation) against vgc_mlp_grad
through the C ABI: equal bit for bit on two calls in a row, the second one
smaller, and a clear error beyond the scratch's row limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VGC_ERR_CAPACITY = -3


def test_mlp_grad_noalloc_matches_mlp_grad(gpu_ctx):
    import torch
    from tests.test_warmstart import _rand_problem

    lib = gpu_ctx.lib.cdll
    f = lib.vgc_mlp_grad_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    for n, seed in ((20000, 1), (3001, 2)):
        p, X, T = _rand_problem(n, seed)
        P = torch.from_numpy(p.astype(np.float32)).cuda()
        Xd = torch.from_numpy(X.astype(np.float32)).cuda()
        Td = torch.from_numpy(T.astype(np.float32)).cuda()
        g_ref, l_ref = gpu_ctx.mlp_grad(P, Xd, Td)
        g = torch.full((P.numel(),), float("nan"), dtype=torch.float32, device="cuda")
        loss = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        rc = f(gpu_ctx.ctx, P.data_ptr(), Xd.data_ptr(), Td.data_ptr(), n, g.data_ptr(),
               loss.data_ptr())
        assert rc == 0, gpu_ctx.lib.last_error(gpu_ctx.ctx)
        assert torch.equal(g, g_ref)
        assert loss.item() == l_ref
    rc = f(gpu_ctx.ctx, P.data_ptr(), Xd.data_ptr(), Td.data_ptr(), 32769, g.data_ptr(), None)
    assert rc == VGC_ERR_CAPACITY
    assert "32768" in gpu_ctx.lib.last_error(gpu_ctx.ctx)

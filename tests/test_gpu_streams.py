"""Stream contract of the C ABI: a call waits for the work the caller queued
before it on its stream (here torch's current stream), so device inputs
written by asynchronous kernels are complete when the library reads them."""
import numpy as np
import pytest

import geogram_amd as ga

pytestmark = pytest.mark.gpu


def _busy(torch, n=8192, reps=12):
    # about 100 ms of matrix products queued on the current stream
    a = torch.randn(n, n, device="cuda") / n ** 0.5
    for _ in range(reps):
        a = torch.tanh(a @ a)
    return a


@pytest.mark.parametrize("side_stream", [False, True])
def test_call_waits_for_callers_stream(gpu_ctx, side_stream):
    import torch

    x = ga.sample_points(1, 3, 20000)
    decoy = ga.sample_points(0, 4, 20000)
    ref = gpu_ctx.cells(x, k0=0)
    src = torch.from_numpy(x).cuda()
    s = torch.cuda.Stream() if side_stream else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        dst = torch.from_numpy(decoy).cuda()
        torch.cuda.synchronize()
        _busy(torch)
        dst.copy_(src)  # queued behind the products
        r = gpu_ctx.cells(dst, k0=0)
        vol = r.vol.cpu().numpy()
        nu = torch.full((20000,), 1.0 / 20000, dtype=torch.float64, device="cuda")
        w0 = torch.zeros(20000, dtype=torch.float64, device="cuda")
        _busy(torch)
        w0.fill_(1e-9)  # queued behind the products
        w, st = gpu_ctx.ot_solve(dst, nu, w0=w0, tol=1e-3)
    assert np.array_equal(vol, ref.vol)
    assert st["converged"] == 1 and st["used_fallback_start"] == 0

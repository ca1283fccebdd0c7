"""Learned warm start on the GPU: the gfx950 MLP forward / backward / Adam
kernels and the feature kernel reproduce their CPU emulation (which
tests/test_warmstart.py checks against the FP64 oracle), and a short training
run on the GPU matches the emulated one step for step."""
import numpy as np
import pytest

import geogram_amd as ga
from geogram_amd import warmstart
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    p = warmstart.init_params(seed)
    p[omlp.OFF_B1:omlp.OFF_W2] = rng.normal(0, 0.1, 64)
    X = rng.normal(0, 1, (n, 6)).astype(np.float32)
    T = rng.normal(0, 0.3, n).astype(np.float32)
    return p, X, T


@pytest.mark.parametrize("n", [1, 37, 64, 20000])
def test_gpu_mlp_matches_emulation(gpu_ctx, emu_ctx, n):
    import torch
    p, X, T = _problem(n, n)
    y_e = emu_ctx.mlp_forward(p, X)
    g_e, l_e = emu_ctx.mlp_grad(p, X, T)
    # host pointers through the staging path
    y_g = gpu_ctx.mlp_forward(p, X)
    g_g, l_g = gpu_ctx.mlp_grad(p, X, T)
    assert np.array_equal(y_g, y_e)
    assert np.array_equal(g_g, g_e)
    assert l_g == l_e
    # device pointers
    pd, Xd, Td = (torch.from_numpy(a).cuda() for a in (p, X, T))
    assert np.array_equal(gpu_ctx.mlp_forward(pd, Xd).cpu().numpy(), y_e)
    gd, ld = gpu_ctx.mlp_grad(pd, Xd, Td)
    assert np.array_equal(gd.cpu().numpy(), g_e) and ld == l_e


def test_gpu_adam_matches_emulation(gpu_ctx, emu_ctx):
    import torch
    p, X, T = _problem(5000, 3)
    g, _ = emu_ctx.mlp_grad(p, X, T)
    pe, me, ve = p.copy(), np.zeros_like(p), np.zeros_like(p)
    pd = torch.from_numpy(p.copy()).cuda()
    md, vd = torch.zeros_like(pd), torch.zeros_like(pd)
    gd = torch.from_numpy(g).cuda()
    for step in range(1, 4):
        emu_ctx.mlp_adam(pe, g, me, ve, step, 1e-3)
        gpu_ctx.mlp_adam(pd, gd, md, vd, step, 1e-3)
    # sqrt and divide: correctly rounded on both sides, but allow one ulp
    assert np.allclose(pd.cpu().numpy(), pe, rtol=2e-7, atol=0)
    assert np.allclose(md.cpu().numpy(), me, rtol=0, atol=0)


def test_gpu_features_match_emulation(gpu_ctx, emu_ctx):
    x, nu = warmstart.make_instance(5, 30000)
    fe = emu_ctx.ot_features(x, nu)
    fg = gpu_ctx.ot_features(x, nu)
    # log / sqrt of the device math library against glibc: a few ulp of FP64,
    # then rounded to FP32
    assert np.max(np.abs(fg - fe)) <= 1e-6


def test_gpu_training_matches_emulation(gpu_ctx, emu_ctx):
    import torch
    rng = np.random.default_rng(9)
    X = rng.normal(0, 1, (3000, 6)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    pe, me, ve = warmstart.init_params(1), np.zeros(4673, np.float32), np.zeros(4673, np.float32)
    pd = torch.from_numpy(pe.copy()).cuda()
    md, vd = torch.zeros_like(pd), torch.zeros_like(pd)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    losses = []
    for step in range(1, 41):
        ge, le = emu_ctx.mlp_grad(pe, X, T)
        emu_ctx.mlp_adam(pe, ge, me, ve, step, 3e-3)
        gd, ld = gpu_ctx.mlp_grad(pd, Xd, Td)
        gpu_ctx.mlp_adam(pd, gd, md, vd, step, 3e-3)
        losses.append((le, ld))
    le = np.array([a for a, _ in losses])
    ld = np.array([b for _, b in losses])
    assert np.allclose(ld, le, rtol=1e-4)
    assert le[-1] < 0.5 * le[0]
    assert np.allclose(pd.cpu().numpy(), pe, rtol=1e-4, atol=1e-6)

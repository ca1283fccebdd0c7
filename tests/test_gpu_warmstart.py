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


def test_learned_start_against_plain(gpu_ctx):
    """What the trained network gains over the plain start on a held-out
    200k-seed 16-Gaussian instance (log-normal masses) solved to 1e-4.  The
    judged start may fall back to w = 0; it must then cost at most the one
    extra cell pass of the judgement, and never more Newton steps."""
    import time
    import torch

    x, nu = warmstart.make_instance(4242, 200_000)
    xd, nud = torch.from_numpy(x).cuda(), torch.from_numpy(nu).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, sp = gpu_ctx.ot_solve(xd, nud, tol=1e-4)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    params = torch.from_numpy(warmstart.load_params()).cuda()
    t0 = time.perf_counter()
    w0 = warmstart.predict_weights(gpu_ctx, xd, nud, params=params)
    _, sl = gpu_ctx.ot_solve(xd, nud, w0=w0, tol=1e-4, judge_start=True)
    torch.cuda.synchronize()
    tl = time.perf_counter() - t0
    print(f"plain: {sp['newton_iters']} Newton, {tp:.2f} s; learned: {sl['newton_iters']} Newton, "
          f"{tl:.2f} s, fell back to w = 0: {bool(sl['used_fallback_start'])}")
    assert sp["converged"] == 1 and sl["converged"] == 1
    assert sl["newton_iters"] <= sp["newton_iters"]
    if sl["used_fallback_start"]:
        assert sl["newton_iters"] == sp["newton_iters"]
        assert sl["cell_passes"] <= sp["cell_passes"] + 2


def test_gpu_convex_start_matches_emulation(gpu_ctx, emu_ctx):
    """The convex-start kernel on the GPU against its CPU emulation (the
    emulation is checked against the definition in test_warmstart.py); exp
    and log come from different math libraries, hence a tolerance."""
    import torch

    rng = np.random.default_rng(5)
    x = rng.random((50_000, 3))
    m = 3000
    pc = np.column_stack([rng.normal(size=m), rng.normal(size=(m, 3)), rng.random(m) * 2,
                          rng.random((m, 3)), rng.random(m) * 0.3])
    for s in (1e-3, 0.05):
        we = emu_ctx.ot_convex_start(x, pc, s)
        wd = gpu_ctx.ot_convex_start(torch.from_numpy(x).cuda(), torch.from_numpy(pc).cuda(), s)
        wh = gpu_ctx.ot_convex_start(x, pc, s)
        assert np.array_equal(wd.cpu().numpy(), wh)
        assert np.abs(wh - we).max() <= 1e-13 * max(1.0, np.abs(we).max())


def test_coarse_start_against_plain(gpu_ctx):
    """The coarse-plan start on a held-out 200k-seed 16-Gaussian instance
    (log-normal masses) solved to 1e-4: no empty cell, kept by the judge, at
    least 30 % fewer Newton steps than the zero start and less time in all
    (the coarse solve and the start included)."""
    import time
    import torch

    x, nu = warmstart.make_instance(4242, 200_000)
    xd, nud = torch.from_numpy(x).cuda(), torch.from_numpy(nu).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, sp = gpu_ctx.ot_solve(xd, nud, tol=1e-4, inexact=True)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    t0 = time.perf_counter()
    w0, info = warmstart.coarse_start(gpu_ctx, xd, nud, ratio=128)
    vol = gpu_ctx.cells(xd, weights=w0, neighbors=False).vol
    assert bool((vol > 0).all())
    w, sc = gpu_ctx.ot_solve(xd, nud, w0=w0, tol=1e-4, inexact=True, judge_start=True)
    torch.cuda.synchronize()
    tc = time.perf_counter() - t0
    info = {k: v for k, v in info.items() if k != "mu_fine"}
    print(f"plain: {sp['newton_iters']} Newton, {tp:.2f} s; coarse start ({info}): "
          f"{sc['newton_iters']} Newton, {tc:.2f} s")
    assert sp["converged"] == 1 and sc["converged"] == 1
    assert sc["used_fallback_start"] == 0
    assert sc["newton_iters"] <= 0.7 * sp["newton_iters"]
    assert tc < tp
    m = gpu_ctx.cells(xd, weights=w, neighbors=False).vol
    assert float(((m - nud).abs() / nud).max()) <= 1e-4


def test_coarse_start_batch_against_plain(gpu_ctx):
    """Eight 16-Gaussian instances with log-normal masses (20k to 100k seeds)
    solved in one batch to 1e-4: with the coarse-plan starts no instance
    restarts from 0, all converge, and the slowest needs fewer Newton steps
    than the slowest from w = 0."""
    import torch

    rng = np.random.default_rng(8)
    sizes = np.exp(rng.uniform(np.log(20_000), np.log(100_000), 8)).astype(np.int64)
    parts = [warmstart.make_instance(700 + k, int(n)) for k, n in enumerate(sizes)]
    xd = torch.from_numpy(np.vstack([p[0] for p in parts])).cuda()
    nud = torch.from_numpy(np.concatenate([p[1] for p in parts])).cuda()
    boxes = np.array([[0, 0, 0, 1, 1, 1]] * 8, float)
    _, s0 = gpu_ctx.ot_solve_batch(xd, nud, sizes, boxes, tol=1e-4, inexact=True)
    w0, info = warmstart.coarse_start_batch(gpu_ctx, xd, nud, sizes, boxes)
    _, s1 = gpu_ctx.ot_solve_batch(xd, nud, sizes, boxes, w0=w0, tol=1e-4, inexact=True)
    print("zero:", [s["newton_iters"] for s in s0], "coarse:", [s["newton_iters"] for s in s1])
    assert all(s["converged"] == 1 and s["used_fallback_start"] == 0 for s in s1)
    assert max(s["newton_iters"] for s in s1) < max(s["newton_iters"] for s in s0)

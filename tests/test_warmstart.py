"""Learned warm start of the OT solver (CPU): the FP64 oracle of the MLP is
pinned by finite differences and by torch autograd; the kernels (CPU
emulation, bit-identical to the gfx950 build) are checked against it: forward,
back-propagated gradient, Adam step, seed features, and the judged start."""
import numpy as np
import pytest

import geogram_amd as ga
from geogram_amd import warmstart
import oracle
from oracle import mlp as omlp


def _rand_problem(n, seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    p = warmstart.init_params(seed).astype(np.float64)
    p[omlp.OFF_B1:omlp.OFF_W2] = rng.normal(0, 0.1, 64)
    p[omlp.OFF_B2:omlp.OFF_W3] = rng.normal(0, 0.1, 64)
    p[omlp.OFF_B3] = 0.05
    X = rng.normal(0, 1, (n, 6))
    T = rng.normal(0, scale, n)
    return p, X, T


def test_oracle_grad_finite_differences():
    p, X, T = _rand_problem(40, 1)
    _, g = omlp.loss_and_grad(p, X, T)
    rng = np.random.default_rng(2)
    # every block of the layout: W1, b1, W2, b2, W3, b3
    probe = np.r_[rng.choice(omlp.OFF_B1, 8, replace=False),
                  omlp.OFF_B1 + rng.choice(64, 4, replace=False),
                  omlp.OFF_W2 + rng.choice(4096, 8, replace=False),
                  omlp.OFF_B2 + rng.choice(64, 4, replace=False),
                  omlp.OFF_W3 + rng.choice(64, 4, replace=False), [omlp.OFF_B3]]
    h = 1e-6
    for i in probe:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (omlp.loss_and_grad(pp, X, T)[0] - omlp.loss_and_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_oracle_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    p, X, T = _rand_problem(300, 3)
    loss, g = omlp.loss_and_grad(p, X, T)
    W1, b1, W2, b2, W3, b3 = [torch.tensor(np.array(a), dtype=torch.float64, requires_grad=True)
                              for a in omlp.unpack(p)]
    Xt = torch.tensor(X)
    lk = torch.nn.functional.leaky_relu
    y = lk(lk(Xt @ W1 + b1, 0.01) @ W2 + b2, 0.01) @ W3 + b3
    lt = ((y - torch.tensor(T)) ** 2).mean()
    lt.backward()
    gt = np.concatenate([a.grad.numpy().ravel() for a in (W1, b1, W2, b2, W3)] + [[b3.grad.item()]])
    assert abs(lt.item() - loss) <= 1e-14 * loss
    assert np.allclose(g, gt, rtol=1e-12, atol=1e-15)
    # Adam against torch.optim.Adam, three steps
    q = torch.tensor(p, requires_grad=True)
    opt = torch.optim.Adam([q], lr=1e-2)
    pr, m, v = p.copy(), np.zeros_like(p), np.zeros_like(p)
    for step in range(1, 4):
        gs = g * step
        q.grad = torch.tensor(gs)
        opt.step()
        pr, m, v = omlp.adam(pr, gs, m, v, step, 1e-2)
    assert np.allclose(pr, q.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_kernels_match_oracle(emu_ctx):
    p, X, T = _rand_problem(5000, 4)
    p32 = p.astype(np.float32)
    X32 = X.astype(np.float32)
    T32 = T.astype(np.float32)
    # oracle evaluated on the FP32-rounded inputs: the remaining gap is FP32
    # arithmetic inside the kernel
    pr, Xr, Tr = p32.astype(np.float64), X32.astype(np.float64), T32.astype(np.float64)
    y = emu_ctx.mlp_forward(p32, X32)
    yr = omlp.forward(pr, Xr)
    assert np.max(np.abs(y - yr)) <= 2e-6 * (1 + np.abs(yr).max())
    g, loss = emu_ctx.mlp_grad(p32, X32, T32)
    lr_, gr = omlp.loss_and_grad(pr, Xr, Tr)
    assert abs(loss - lr_) <= 1e-5 * lr_
    assert np.max(np.abs(g - gr)) <= 2e-5 * np.abs(gr).max()
    # per block: the relative error of each parameter block stays small
    for lo, hi in [(0, omlp.OFF_B1), (omlp.OFF_W2, omlp.OFF_B2), (omlp.OFF_W3, omlp.OFF_B3)]:
        assert np.linalg.norm(g[lo:hi] - gr[lo:hi]) <= 1e-5 * np.linalg.norm(gr[lo:hi])
    # Adam
    m = np.zeros(omlp.N_PARAMS, np.float32)
    v = np.zeros(omlp.N_PARAMS, np.float32)
    q = p32.copy()
    emu_ctx.mlp_adam(q, g, m, v, 1, 1e-3)
    qr, _, _ = omlp.adam(pr, g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    assert np.max(np.abs(q - qr)) <= 1e-6 * np.abs(qr).max() + 1e-9


def test_kernel_grad_rows_not_multiple_of_tile(emu_ctx):
    # 37 rows: partial tile of the backward pass and of the forward pass
    p, X, T = _rand_problem(37, 5)
    g, loss = emu_ctx.mlp_grad(p.astype(np.float32), X.astype(np.float32), T.astype(np.float32))
    lr_, gr = omlp.loss_and_grad(p.astype(np.float32).astype(np.float64),
                                 X.astype(np.float32).astype(np.float64),
                                 T.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(g - gr)) <= 2e-5 * np.abs(gr).max()


def test_features_match_oracle(emu_ctx):
    x, nu = warmstart.make_instance(11, 4000)
    f = emu_ctx.ot_features(x, nu)
    _, d2 = oracle.knn(x, 16)
    fr = omlp.features(x, nu, (0, 0, 0, 1, 1, 1), d2)
    assert np.max(np.abs(f - fr)) <= 1e-5
    # a box other than the unit cube
    box = (-1.0, 0.0, 2.0, 1.0, 3.0, 2.5)
    xb = np.asarray(box[:3]) + x * (np.asarray(box[3:]) - np.asarray(box[:3]))
    vol = 2.0 * 3.0 * 0.5
    f = emu_ctx.ot_features(xb, nu * vol, box)
    _, d2 = oracle.knn(xb, 16)
    fr = omlp.features(xb, nu * vol, box, d2)
    assert np.max(np.abs(f - fr)) <= 1e-5


def test_fold_normalisation():
    p, X, _ = _rand_problem(100, 6)
    mu, sd = X.mean(0) + 0.3, X.std(0) * 1.7
    raw = warmstart.fold_normalisation(p.astype(np.float32), mu, sd, 0.25, 3.0)
    y_norm = omlp.forward(p.astype(np.float32).astype(np.float64), (X - mu) / sd)
    y_raw = omlp.forward(raw.astype(np.float64), X)
    assert np.allclose(y_raw, 3.0 * y_norm + 0.25, rtol=1e-5, atol=1e-5)


def test_training_reduces_loss(emu_ctx):
    # a smooth synthetic target of the six features: Adam on the emulated
    # kernels must fit it well below its variance
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (4000, 6)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 5]).astype(np.float32)
    params, rep = warmstart.train(emu_ctx, [X], [T], steps=300, batch=1024, lr=3e-3,
                                  log=lambda *_: None)
    y = omlp.forward(params.astype(np.float64), X.astype(np.float64))
    assert np.mean((y - T) ** 2) < 0.05 * np.var(T)


def test_judged_start_falls_back_to_zero(emu_ctx):
    x, nu = warmstart.make_instance(21, 1500)
    # a start with a huge residual (no empty cell): the judge keeps w = 0
    rng = np.random.default_rng(0)
    w_bad = rng.normal(0, 1e-3, x.shape[0])
    w1, s1 = emu_ctx.ot_solve(x, nu, w0=w_bad, tol=1e-5, judge_start=True)
    w0, s0 = emu_ctx.ot_solve(x, nu, tol=1e-5)
    assert s1["used_fallback_start"] == 1
    assert s1["newton_iters"] == s0["newton_iters"]
    assert np.array_equal(w1, w0)
    # the converged weights as start: kept, and nothing left to do
    w2, s2 = emu_ctx.ot_solve(x, nu, w0=w0, tol=1e-5, judge_start=True)
    assert s2["used_fallback_start"] == 0 and s2["newton_iters"] == 0


# ------------------------------------------------ convex start (coarse plan)
def _lse_reference(x, pc, s):
    """w = |p|^2 - s log sum_j exp(piece_j(p) / s), straight from the
    definition (all terms, FP64)."""
    d = x[:, None, :] - pc[None, :, 5:8]
    r2 = (d ** 2).sum(2)
    R = pc[None, :, 8]
    h = np.where(r2 <= R * R, r2, R * (2.0 * np.sqrt(r2) - R))
    v = pc[None, :, 0] + x @ pc[:, 1:4].T + pc[None, :, 4] * h
    mx = v.max(1)
    return (x * x).sum(1) - (mx + s * np.log(np.exp((v - mx[:, None]) / s).sum(1)))


@pytest.mark.parametrize("s", [1e-4, 0.05, 3.0])
def test_convex_start_kernel_matches_definition(emu_ctx, s):
    rng = np.random.default_rng(3)
    x = rng.random((700, 3))
    m = 150
    pc = np.column_stack([rng.normal(size=m), rng.normal(size=(m, 3)), rng.random(m) * 2,
                          rng.random((m, 3)), rng.random(m) * 0.3])
    w = emu_ctx.ot_convex_start(x, pc, s)
    ref = _lse_reference(x, pc, s)
    assert np.abs(w - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    # one piece: the soft maximum is that piece
    w1 = emu_ctx.ot_convex_start(x, pc[:1], s)
    assert np.allclose(w1, _lse_reference(x, pc[:1], s), rtol=0, atol=1e-14)
    with pytest.raises(ga.VgcError):
        emu_ctx.ot_convex_start(x, pc, 0.0)


def test_coarse_start_no_empty_cell_and_fewer_newton_steps(emu_ctx):
    """On a 16-Gaussian instance with log-normal masses (the family the zero
    start handles worst) the coarse-plan start leaves no cell empty, the
    judge keeps it, and Newton needs at most 70 % of the zero start's steps."""
    x, nu = warmstart.make_instance(4, 6000)
    _, sp = emu_ctx.ot_solve(x, nu, tol=1e-4, inexact=True)
    w0, info = warmstart.coarse_start(emu_ctx, x, nu, ratio=16)
    assert info["coarse_seeds"] == 375
    vol = np.asarray(emu_ctx.cells(x, weights=w0, neighbors=False).vol)
    assert (vol > 0).all()
    w, sl = emu_ctx.ot_solve(x, nu, w0=w0, tol=1e-4, inexact=True, judge_start=True)
    assert sp["converged"] == 1 and sl["converged"] == 1
    assert sl["used_fallback_start"] == 0
    assert sl["newton_iters"] <= 0.7 * sp["newton_iters"]
    m = np.asarray(emu_ctx.cells(x, weights=w, neighbors=False).vol)
    assert np.max(np.abs(m - nu) / nu) <= 1e-4


def test_coarse_start_batch(emu_ctx):
    """Batched instances (16-Gaussian, log-normal masses): the coarse plans
    are solved together, every instance keeps its start (no restart from 0),
    converges, and needs fewer Newton steps than from w = 0."""
    sizes = [1500, 3000, 2200]
    parts = [warmstart.make_instance(50 + k, n) for k, n in enumerate(sizes)]
    x = np.vstack([p[0] for p in parts])
    nu = np.concatenate([p[1] for p in parts])
    boxes = np.array([[0, 0, 0, 1, 1, 1]] * 3, float)
    w0, info = warmstart.coarse_start_batch(emu_ctx, x, nu, sizes, boxes, ratio=16)
    assert info["coarse_converged"] == 3
    _, s0 = emu_ctx.ot_solve_batch(x, nu, sizes, boxes, tol=1e-4, inexact=True)
    _, s1 = emu_ctx.ot_solve_batch(x, nu, sizes, boxes, w0=w0, tol=1e-4, inexact=True)
    assert all(s["converged"] == 1 and s["used_fallback_start"] == 0 for s in s1)
    assert all(b["newton_iters"] < a["newton_iters"] for a, b in zip(s0, s1))

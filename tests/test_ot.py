"""Semi-discrete optimal transport: the Laguerre Hessian against finite
differences of the oracle's cell masses, the solver against a numpy/scipy
restatement of the same damped Newton method (direct sparse solves), and the
solution's masses re-measured by the oracle.  The Exploragram solver that
compute_OTM calls is not in the geogram tree, so its iteration trajectory is
not pinned; the method (Kitagawa-Merigot-Thibert damped Newton) is."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import geogram_amd as ga
import oracle


def hessian_from_cells(x, r):
    """dm/dw from a CellResult: -A_ij / (2 |p_i - p_j|) off the diagonal."""
    n = x.shape[0]
    ptr = r.nbr_ptr.astype(np.int64)
    rows = np.repeat(np.arange(n), np.diff(ptr))
    cols = r.nbr.astype(np.int64)
    d = np.linalg.norm(x[cols] - x[rows], axis=1)
    v = -r.farea / (2 * d)
    H = sp.csr_matrix((v, (rows, cols)), shape=(n, n))
    H = H - sp.diags(np.asarray(H.sum(axis=1)).ravel())
    return H


def test_hessian_matches_finite_differences(emu_ctx):
    rng = np.random.default_rng(0)
    n = 400
    x = rng.random((n, 3))
    w = rng.random(n) * 1e-3
    r = emu_ctx.cells(x, weights=w)
    H = hessian_from_cells(x, r).toarray()
    h = 1e-7
    for j in rng.choice(n, 12, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        mp = oracle.cells(x, w=wp, centred=True)["vol"]
        mm = oracle.cells(x, w=wm, centred=True)["vol"]
        fd = (mp - mm) / (2 * h)
        assert np.allclose(H[:, j], fd, rtol=2e-5, atol=1e-9 * np.abs(H[:, j]).max())
    # Laplacian structure: symmetric, rows sum to zero
    assert np.allclose(H, H.T, rtol=1e-9, atol=1e-12 * np.abs(H).max())
    assert np.abs(H.sum(axis=1)).max() < 1e-10 * np.abs(H).max()


def kmt_newton_reference(x, nu, tol, w0=None, max_newton=100):
    """The damped Newton method with exact (sparse direct) Newton steps and the
    oracle's Laguerre cells."""
    n = x.shape[0]
    w = np.zeros(n) if w0 is None else w0.copy()

    def cells(w):
        o = oracle.cells(x, w=w, centred=True, stride=256)
        class R:  # CellResult-like view
            pass
        R.vol = o["vol"]
        cnt = o["cnt"]
        R.nbr_ptr = np.r_[0, np.cumsum(cnt)]
        R.nbr = np.concatenate([o["nbr"][i, :cnt[i]] for i in range(n)])
        R.farea = np.concatenate([o["farea"][i, :cnt[i]] for i in range(n)])
        return R

    r = cells(w)
    m = r.vol
    eps0 = 0.5 * min(nu.min(), m.min())
    err = np.max(np.abs(m - nu) / nu)
    g2 = np.sum((m - nu) ** 2)
    it = 0
    while err > tol and it < max_newton:
        H = hessian_from_cells(x, r).tocsc()
        # fix the additive constant: ground the last weight
        d = np.zeros(n)
        d[:-1] = spla.spsolve(H[:-1, :-1], (nu - m)[:-1])
        d -= d.mean()
        t = 1.0
        while True:
            r2 = cells(w + t * d)
            m2 = r2.vol
            if m2.min() >= eps0 and np.sqrt(np.sum((m2 - nu) ** 2)) <= (1 - t / 2) * np.sqrt(g2):
                break
            t /= 2
        w = w + t * d
        r, m = r2, m2
        err = np.max(np.abs(m - nu) / nu)
        g2 = np.sum((m - nu) ** 2)
        it += 1
    return w, it, err


@pytest.mark.parametrize("kind", [0, 2])
def test_ot_solver_vs_reference_newton(emu_ctx, kind):
    n = 1500
    x = ga.sample_points(kind, 30 + kind, n)
    rng = np.random.default_rng(kind)
    nu = np.exp(rng.normal(0.0, 0.5, n)) if kind == 2 else np.ones(n)
    nu /= nu.sum()
    w, st = emu_ctx.ot_solve(x, nu, tol=1e-6)
    assert st["converged"] == 1 and st["final_err"] <= 1e-6
    wr, itr, errr = kmt_newton_reference(x, nu, 1e-6)
    assert st["newton_iters"] == itr
    # weights are defined up to a constant
    dw = (w - w.mean()) - (wr - wr.mean())
    assert np.abs(dw).max() <= 1e-6 * np.abs(wr - wr.mean()).max()
    # masses of the solution, re-measured by the oracle
    o = oracle.cells(x, w=w, centred=True, stride=256)
    assert np.max(np.abs(o["vol"] - nu) / nu) <= 1.01e-6


def test_ot_fallback_start_and_argument_checks(emu_ctx):
    n = 800
    x = ga.sample_points(0, 5, n)
    nu = np.full(n, 1.0 / n)
    w0 = np.zeros(n)
    w0[:5] = 1.0  # five seeds swallow everything: empty cells at the start
    w, st = emu_ctx.ot_solve(x, nu, w0=w0, tol=1e-5)
    assert st["used_fallback_start"] == 1 and st["converged"] == 1
    with pytest.raises(ga.VgcError):
        emu_ctx.ot_solve(x, nu * 2, tol=1e-5)  # masses do not sum to the volume


def test_ot_batch_equals_single_solves(emu_ctx):
    from geogram_amd import warmstart
    sizes = [1500, 800, 2500]
    xs, nus = [], []
    for i, s in enumerate(sizes):
        x, nu = warmstart.make_instance(50 + i, s)
        xs.append(x)
        nus.append(nu)
    boxes = np.array([[0, 0, 0, 1, 1, 1]] * 3, dtype=float)
    # the second instance in another box (masses scaled to its volume)
    boxes[1] = [1.0, -1.0, 0.0, 3.0, 0.0, 0.5]
    xs[1] = boxes[1, :3] + xs[1] * (boxes[1, 3:] - boxes[1, :3])
    nus[1] = nus[1] * 1.0
    x, nu = np.vstack(xs), np.concatenate(nus)
    w, sts = emu_ctx.ot_solve_batch(x, nu, sizes, boxes, tol=1e-5)
    o = 0
    for k, s in enumerate(sizes):
        wk, sk = emu_ctx.ot_solve(x[o:o + s], nu[o:o + s], box=boxes[k], tol=1e-5)
        assert sts[k]["converged"] == 1 and sts[k]["newton_iters"] == sk["newton_iters"]
        a, b = w[o:o + s] - w[o:o + s].mean(), wk - wk.mean()
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
        o += s
    with pytest.raises(ga.VgcError):
        emu_ctx.ot_solve_batch(x, nu * 2, sizes, boxes, tol=1e-5)


def test_power_nearest_start_weights(emu_ctx):
    """The start of a finer level's new seeds: the weight of the seed of the
    previous level nearest in power distance (brute force check)."""
    import ctypes
    rng = np.random.default_rng(1)
    m, nq = 3000, 2000
    x = rng.random((m, 3))
    w = rng.random(m) * 0.01
    q = rng.random((nq, 3))
    out = np.zeros(nq)
    box = np.array([0, 0, 0, 1, 1, 1.0])
    vp = ctypes.c_void_p
    rc = emu_ctx.lib.cdll.vgc_emu_power_nearest(
        emu_ctx.ctx, x.ctypes.data_as(vp), w.ctypes.data_as(vp), m, box.ctypes.data_as(vp),
        q.ctypes.data_as(vp), nq, out.ctypes.data_as(vp))
    assert rc == 0
    d = ((q[:, None, :] - x[None]) ** 2).sum(-1) - w[None]
    assert np.array_equal(out, w[d.argmin(1)])


def test_ot_solve_levels(emu_ctx):
    """optimize_levels: each level converges; the masses of the final weights,
    re-measured by the oracle, meet the tolerance; bad levels are refused."""
    n = 4000
    x, nu = ga.sample_points(0, 41, n), np.full(n, 1.0 / n)
    perm, lv = ga.brio_levels(n, min_level=400, seed=3)
    x, nu = x[perm], nu[perm]
    assert lv.tolist() == [0, 500, 4000]
    w, sts = emu_ctx.ot_solve_levels(x, nu, lv, tol=1e-5)
    assert len(sts) == 2 and all(s["converged"] == 1 for s in sts)
    o = oracle.cells(x, w=w, centred=True, stride=256)
    assert np.max(np.abs(o["vol"] - nu) / nu) <= 1.01e-5
    for bad in ([0, 500], [0, 500, 500, 4000], [1, 4000]):
        with pytest.raises(ga.VgcError):
            emu_ctx.ot_solve_levels(x, nu, np.array(bad), tol=1e-5)


def test_ot_solve_levels_convex_start(emu_ctx):
    """optimize_levels with VGC_OT_CONVEX_LEVELS on a 16-Gaussian instance
    with log-normal masses, where copied weights leave cells empty: the
    finer level starts from the convex extension of the coarser plan, keeps
    it (no fallback), needs at most 70 % of a zero-start solve's Newton steps,
    and its masses, re-measured by the oracle, meet the tolerance."""
    from geogram_amd import warmstart

    n = 6000
    x, nu = warmstart.make_instance(4, n)
    perm, lv = ga.brio_levels(n, min_level=700, seed=3)
    x, nu = np.ascontiguousarray(x[perm]), np.ascontiguousarray(nu[perm])
    assert lv.tolist() == [0, 750, 6000]
    _, sp = emu_ctx.ot_solve(x, nu, tol=1e-4, inexact=True)
    w, sts = emu_ctx.ot_solve_levels(x, nu, lv, tol=1e-4, inexact=True, convex=True)
    assert all(st["converged"] == 1 for st in sts)
    assert sts[1]["used_fallback_start"] == 0
    assert sts[1]["newton_iters"] <= 0.7 * sp["newton_iters"]
    o = oracle.cells(x, w=w, centred=True, stride=256)
    assert np.max(np.abs(o["vol"] - nu) / nu) <= 1.01e-4


def test_repaired_start_keeps_a_start_with_a_few_empty_cells(emu_ctx):
    # a converged solution with one weight lowered until its cell empties:
    # the repair raises that weight back (one round), the solve keeps the
    # start and needs fewer Newton steps than from w = 0
    x = ga.sample_points(0, 31, 3000)
    nu = np.full(3000, 1.0 / 3000)
    ws, s0 = emu_ctx.ot_solve(x, nu, tol=1e-6)
    w = ws.copy()
    w[[10, 20]] -= 0.05
    r = emu_ctx.cells(x, w)
    assert (r.vol[[10, 20]] == 0).all()
    _, plain = emu_ctx.ot_solve(x, nu, w0=w, tol=1e-6)
    assert plain["used_fallback_start"] == 1
    w2, st = emu_ctx.ot_solve(x, nu, w0=w, tol=1e-6, repair=True)
    assert st["used_fallback_start"] == 0 and st["repair_rounds"] >= 1 and st["converged"]
    assert st["newton_iters"] < plain["newton_iters"]
    o = oracle.cells(x, w=w2, centred=True)
    assert np.abs(o["vol"] - nu).max() <= 1.01e-6 * nu[0]

"""Quasi-Newton CVT (vgc_cvt_newton, CentroidalVoronoiTesselation::
Newton_iterations, voronoi/CVT.cpp:272-302).

The oracle optimiser (oracle/lbfgs.py, HLBFGS + MINPACK line search restated)
is pinned by mathematics: the minimiser of a strictly convex quadratic, the
minimum of Rosenbrock's function (also found by scipy's L-BFGS-B), and the
strong Wolfe conditions at every accepted step.  The library's emulated
driver is then held to the oracle's iterates on the CVT energy, with the
energy and gradient of the oracle's cells."""
import numpy as np
import pytest

import geogram_amd as ga
import oracle
from oracle import lbfgs as olb


def _wolfe_checked(fg):
    """Wrap fg to record every evaluation."""
    log = []

    def f(x):
        v, g = fg(x)
        log.append((x.copy(), v, g.copy()))
        return v, g
    return f, log


def test_oracle_lbfgs_quadratic_minimiser():
    rng = np.random.default_rng(0)
    n = 30
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = Q @ np.diag(np.logspace(0, 2, n)) @ Q.T
    b = rng.normal(size=n)
    xs = np.linalg.solve(A, b)
    fg = lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b)
    x, f, g, hist, nev, info = olb.lbfgs(fg, np.zeros(n), 200)
    # it runs until the line search meets rounding (info 3 or 6)
    assert np.linalg.norm(g) <= 1e-6 * np.linalg.norm(b)
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)
    assert all(b_ <= a_ for a_, b_ in zip(hist, hist[1:]))


def test_oracle_lbfgs_rosenbrock_and_wolfe():
    from scipy.optimize import minimize, rosen, rosen_der
    fg = lambda x: (rosen(x), rosen_der(x))
    start = np.array([-1.2, 1.0, -0.5, 0.8])
    x, f, g, hist, nev, info = olb.lbfgs(fg, start, 200)
    ref = minimize(rosen, start, jac=rosen_der, method="L-BFGS-B",
                   options={"gtol": 1e-12, "ftol": 1e-15, "maxiter": 1000})
    assert np.abs(x - 1.0).max() < 1e-6 and np.abs(ref.x - 1.0).max() < 1e-5
    assert hist[-1] < 1e-12
    # a line search that reports convergence meets the strong Wolfe
    # conditions (ftol 1e-4, gtol 0.9); one that is started uphill refuses
    rng = np.random.default_rng(1)
    n_conv = 0
    for _ in range(200):
        x0 = rng.normal(size=4)
        f0, g0 = fg(x0)
        d = -g0 * rng.uniform(0.01, 3.0) + 0.1 * rng.normal(size=4) * np.linalg.norm(g0)
        if g0 @ d >= 0:
            assert olb.line_search(fg, x0, f0, g0, d, 1.0)[4] == -1
            continue
        x1, f1, g1, stp, info, nfev = olb.line_search(fg, x0, f0, g0, d, 1.0)
        assert nfev <= 20
        if info == 1:
            n_conv += 1
            assert f1 <= f0 + 1e-4 * stp * (g0 @ d)
            assert abs(g1 @ d) <= 0.9 * abs(g0 @ d)
            assert np.allclose(x1, x0 + stp * d)
    assert n_conv > 150


def _oracle_fg(box=oracle.UNIT_BOX):
    def fg(xf):
        x = xf.reshape(-1, 3)
        o = oracle.cells(x, box=box, k0=20, centred=True, e2=True)
        g = 2.0 * (o["vol"][:, None] * x - o["mg"])
        return float(o["e2"].sum()), g.reshape(-1)
    return fg


@pytest.mark.parametrize("kind", [0, 1])
def test_emulated_cvt_newton_vs_oracle(emu_ctx, kind):
    n, iters = 1500, 6
    x0 = ga.sample_points(kind, 30 + kind, n)
    xg, st = emu_ctx.cvt_newton(x0, iters)
    xo, fo, go, hist, nev, info = olb.lbfgs(_oracle_fg(), x0.reshape(-1), iters)
    assert st["iterations"] == len(hist) and st["evaluations"] == nev
    assert st["line_search_info"] == info
    assert abs(st["f"] - fo) <= 1e-11 * fo
    assert np.abs(xg.reshape(-1) - xo).max() <= 1e-9
    assert st["f"] < st["f0"] and st["gnorm"] < st["gnorm0"]


def test_cvt_newton_arguments(emu_ctx):
    x = ga.sample_points(0, 1, 100)
    with pytest.raises(ga.VgcError):
        emu_ctx.cvt_newton(x, 3, m=0)

"""ORACLE — test infrastructure only.

A numpy restatement of the optimiser behind
CentroidalVoronoiTesselation::Newton_iterations (voronoi/CVT.cpp:272-302):
HLBFGS (third_party/HLBFGS/HLBFGS.cpp) in its standard L-BFGS mode, with
the MINPACK line search of Moré and Thuente (third_party/HLBFGS/LineSearch.cpp,
MCSRCH / MCSTEP).  Parameters as Newton_iterations sets them through
HLBFGSOptimizer::optimize (numerics/lbfgs_optimizers.cpp:159-196):
INIT_HLBFGS defaults (HLBFGS.cpp:30-54; ftol 1e-4, xtol 1e-16, gtol 0.9,
stpmin 1e-20, stpmax 1e20, 20 evaluations per line search) and
||g|| tolerances switched off.
"""
from __future__ import annotations

import math

import numpy as np


def _mcstep(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    """LineSearch.cpp MCSTEP: safeguarded cubic / quadratic step.  Returns
    (stx, fx, dx, sty, fy, dy, stp, brackt, info)."""
    if (brackt and (stp <= min(stx, sty) or stp >= max(stx, sty))) or dx * (stp - stx) >= 0.0 \
            or stpmax < stpmin:
        return stx, fx, dx, sty, fy, dy, stp, brackt, 0
    sgnd = dp * (dx / abs(dx))
    if fp > fx:
        info, bound = 1, True
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        stpc = stx + (p / q) * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0.0:
        info, bound = 2, False
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        stpc = stp + (p / q) * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(dp) < abs(dx):
        info, bound = 3, True
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt(max(0.0, (theta / s) ** 2 - (dx / s) * (dp / s)))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        if r < 0.0 and gamma != 0.0:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stp - stpc) < abs(stp - stpq) else stpq
        else:
            stpf = stpc if abs(stp - stpc) > abs(stp - stpq) else stpq
    else:
        info, bound = 4, False
        if brackt:
            theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp
            s = max(abs(theta), abs(dy), abs(dp))
            gamma = s * math.sqrt((theta / s) ** 2 - (dy / s) * (dp / s))
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            stpf = stp + (p / q) * (sty - stp)
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0.0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    stp = max(stpmin, min(stpmax, stpf))
    if brackt and bound:
        m = stx + 0.66 * (sty - stx)
        stp = min(m, stp) if sty > stx else max(m, stp)
    return stx, fx, dx, sty, fy, dy, stp, brackt, info


def line_search(fg, x0, f0, g0, d, stp, ftol=1e-4, gtol=0.9, xtol=1e-16, stpmin=1e-20,
                stpmax=1e20, maxfev=20):
    """LineSearch.cpp MCSRCH along d from x0.  Returns (x, f, g, stp, info, nfev)."""
    dginit = float(g0 @ d)
    if not (stp > 0.0) or not (dginit < 0.0):
        return x0, f0, g0, stp, -1, 0
    brackt, stage1, infoc, nfev = False, True, 1, 0
    finit, dgtest = f0, ftol * dginit
    width = stpmax - stpmin
    width1 = width / 0.5
    stx, fx, dgx = 0.0, finit, dginit
    sty, fy, dgy = 0.0, finit, dginit
    x, f, g = x0, f0, g0
    while True:
        if brackt:
            stmin, stmax = min(stx, sty), max(stx, sty)
        else:
            stmin, stmax = stx, stp + 4.0 * (stp - stx)
        stp = min(max(stp, stpmin), stpmax)
        if (brackt and (stp <= stmin or stp >= stmax)) or nfev >= maxfev - 1 or infoc == 0 \
                or (brackt and stmax - stmin <= xtol * stmax):
            stp = stx
        x = x0 + stp * d
        f, g = fg(x)
        nfev += 1
        dg = float(g @ d)
        ftest1 = finit + stp * dgtest
        info = 0
        if (brackt and (stp <= stmin or stp >= stmax)) or infoc == 0:
            info = 6
        if stp == stpmax and f <= ftest1 and dg <= dgtest:
            info = 5
        if stp == stpmin and (f > ftest1 or dg >= dgtest):
            info = 4
        if nfev >= maxfev:
            info = 3
        if brackt and stmax - stmin <= xtol * stmax:
            info = 2
        if f <= ftest1 and abs(dg) <= gtol * (-dginit):
            info = 1
        if info != 0:
            return x, f, g, stp, info, nfev
        if stage1 and f <= ftest1 and dg >= min(ftol, gtol) * dginit:
            stage1 = False
        if stage1 and f <= fx and f > ftest1:
            fm, fxm, fym = f - stp * dgtest, fx - stx * dgtest, fy - sty * dgtest
            dgm, dgxm, dgym = dg - dgtest, dgx - dgtest, dgy - dgtest
            stx, fxm, dgxm, sty, fym, dgym, stp, brackt, infoc = _mcstep(
                stx, fxm, dgxm, sty, fym, dgym, stp, fm, dgm, brackt, stmin, stmax)
            fx, fy = fxm + stx * dgtest, fym + sty * dgtest
            dgx, dgy = dgxm + dgtest, dgym + dgtest
        else:
            stx, fx, dgx, sty, fy, dgy, stp, brackt, infoc = _mcstep(
                stx, fx, dgx, sty, fy, dgy, stp, f, dg, brackt, stmin, stmax)
        if brackt:
            if abs(sty - stx) >= 0.66 * width1:
                stp = stx + 0.5 * (sty - stx)
            width1 = width
            width = abs(sty - stx)


def lbfgs(fg, x, max_iter, m=7):
    """HLBFGS standard mode (HLBFGS.cpp:352-585): two-loop recursion with the
    initial matrix (s.y / y.y) I of the newest pair (HLBFGS.cpp:90-118,
    157-196), first step 1/|g|, then 1; stops when a line search fails, the
    gradient vanishes, or after max_iter + 1 iterations (HLBFGS.cpp:578).
    fg(x) -> (f, g).  Returns (x, f, g, history of f after each iteration,
    evaluations, last line-search info)."""
    x = np.array(x, dtype=np.float64, copy=True)
    f, g = fg(x)
    nev = 1
    S, Y, rho = [], [], []
    hist = []
    info = 0
    it = 0
    while True:
        q = -g
        alpha = []
        for s, y, r in zip(reversed(S), reversed(Y), reversed(rho)):
            a = r * float(s @ q)
            q = q - a * y
            alpha.append(a)
        if S:
            q = (float(Y[-1] @ S[-1]) / float(Y[-1] @ Y[-1])) * q
        for (s, y, r), a in zip(zip(S, Y, rho), reversed(alpha)):
            b = r * float(y @ q)
            q = q + (a - b) * s
        stp = 1.0 / math.sqrt(float(g @ g)) if it == 0 else 1.0
        x0, g0 = x, g
        x, f, g, stp, info, nfev = line_search(fg, x0, f, g0, q, stp)
        nev += nfev
        it += 1
        hist.append(f)
        if info != 1 or not (float(g @ g) > 0.0) or it > max_iter:
            break
        S.append(x - x0)
        Y.append(g - g0)
        rho.append(1.0 / float(Y[-1] @ S[-1]))
        if len(S) > m:
            S.pop(0)
            Y.pop(0)
            rho.pop(0)
    return x, f, g, hist, nev, info

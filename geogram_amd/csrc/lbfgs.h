// Limited-memory BFGS on device vectors with the More-Thuente line search:
// the quasi-Newton CVT of CentroidalVoronoiTesselation::Newton_iterations
// (voronoi/CVT.cpp:272-302), which runs HLBFGS with m = 7 pairs, no
// gradient/step tolerances (epsg = epsf = epsx = 0) and the energy and
// gradient of RestrictedVoronoiDiagram::compute_CVT_func_grad.
//
// The iteration follows HLBFGS's standard update
// (third_party/HLBFGS/HLBFGS.cpp:352-585): two-loop recursion over the last
// m pairs (s, y), the initial matrix scaled by s.y / y.y of the newest pair
// (HLBFGS.cpp:90-118), a first trial step 1/|g| then 1, a line search of at
// most 20 evaluations with ftol 1e-4, gtol 0.9, xtol 1e-16, step bounds
// [1e-20, 1e20] (HLBFGS.cpp:30-54), and it stops when the line search fails,
// the gradient vanishes, or the iteration count exceeds the limit
// (HLBFGS.cpp:560-583).  The line search is the algorithm of More and
// Thuente (ACM TOMS 20(3), 1994; MINPACK's mcsrch / mcstep), written here
// as a host-side state machine over scalars; vectors stay on the device.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace vgc {

// One-dimensional safeguarded step of More-Thuente: updates the interval of
// uncertainty [stx, sty] (function values f, derivatives d) with the trial
// (stp, fp, dp) and returns the next trial in stp.
struct MtStep {
    static void step(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy,
                     double& stp, double fp, double dp, bool& brackt, double stpmin,
                     double stpmax, int& info) {
        info = 0;
        if ((brackt && (stp <= std::min(stx, sty) || stp >= std::max(stx, sty))) ||
            dx * (stp - stx) >= 0.0 || stpmax < stpmin)
            return;
        const double sgnd = dp * (dx / std::fabs(dx));
        double stpf, stpc, stpq;
        bool bound;
        if (fp > fx) {
            // higher function value: the minimum is bracketed; cubic step
            // closer to stx, or halfway to the quadratic one
            info = 1;
            bound = true;
            const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
            const double s = std::max(std::fabs(theta), std::max(std::fabs(dx), std::fabs(dp)));
            double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
            if (stp < stx) gamma = -gamma;
            const double p = (gamma - dx) + theta, q = ((gamma - dx) + gamma) + dp;
            stpc = stx + (p / q) * (stp - stx);
            stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
            stpf = std::fabs(stpc - stx) < std::fabs(stpq - stx) ? stpc
                                                                  : stpc + (stpq - stpc) / 2.0;
            brackt = true;
        } else if (sgnd < 0.0) {
            // derivatives of opposite signs: bracketed; the cubic or the
            // secant step, the farther from stp
            info = 2;
            bound = false;
            const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
            const double s = std::max(std::fabs(theta), std::max(std::fabs(dx), std::fabs(dp)));
            double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
            if (stp > stx) gamma = -gamma;
            const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dx;
            stpc = stp + (p / q) * (stx - stp);
            stpq = stp + (dp / (dp - dx)) * (stx - stp);
            stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
            brackt = true;
        } else if (std::fabs(dp) < std::fabs(dx)) {
            // same signs, the derivative shrinks: cubic step if it tends to
            // a minimum beyond stp, else towards the bound
            info = 3;
            bound = true;
            const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
            const double s = std::max(std::fabs(theta), std::max(std::fabs(dx), std::fabs(dp)));
            double gamma =
                s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
            if (stp > stx) gamma = -gamma;
            const double p = (gamma - dp) + theta, q = (gamma + (dx - dp)) + gamma;
            const double r = p / q;
            if (r < 0.0 && gamma != 0.0)
                stpc = stp + r * (stx - stp);
            else
                stpc = stp > stx ? stpmax : stpmin;
            stpq = stp + (dp / (dp - dx)) * (stx - stp);
            if (brackt)
                stpf = std::fabs(stp - stpc) < std::fabs(stp - stpq) ? stpc : stpq;
            else
                stpf = std::fabs(stp - stpc) > std::fabs(stp - stpq) ? stpc : stpq;
        } else {
            // same signs, the derivative does not shrink: cubic step on the
            // bracket, or the bound
            info = 4;
            bound = false;
            if (brackt) {
                const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
                const double s =
                    std::max(std::fabs(theta), std::max(std::fabs(dy), std::fabs(dp)));
                double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
                if (stp > sty) gamma = -gamma;
                const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dy;
                stpf = stp + (p / q) * (sty - stp);
            } else {
                stpf = stp > stx ? stpmax : stpmin;
            }
        }
        if (fp > fx) {
            sty = stp;
            fy = fp;
            dy = dp;
        } else {
            if (sgnd < 0.0) {
                sty = stx;
                fy = fx;
                dy = dx;
            }
            stx = stp;
            fx = fp;
            dx = dp;
        }
        stpf = std::max(stpmin, std::min(stpmax, stpf));
        stp = stpf;
        if (brackt && bound) {
            const double m = stx + 0.66 * (sty - stx);
            stp = sty > stx ? std::min(m, stp) : std::max(m, stp);
        }
    }
};

// Line search along direction d from x0 (function value f0, slope dg0 < 0):
// start() gives the first trial step; after each evaluation at x0 + stp d,
// next(f, dg) returns 0 to evaluate again at the new stp, or the outcome:
// 1 sufficient decrease and curvature conditions hold; 2 relative width of
// the interval below xtol; 3 maxfev evaluations; 4 step at stpmin; 5 step at
// stpmax; 6 rounding errors prevent progress; -1 bad input / not a descent
// direction.
struct MoreThuente {
    double ftol = 1e-4, gtol = 0.9, xtol = 1e-16, stpmin = 1e-20, stpmax = 1e20;
    int maxfev = 20;
    double stp = 0.0;
    int nfev = 0;

    int start(double f0, double dg0, double stp0) {
        if (!(stp0 > 0.0) || !(dg0 < 0.0)) return -1;
        stp = stp0;
        infoc = 1;
        brackt = false;
        stage1 = true;
        nfev = 0;
        finit = f0;
        dginit = dg0;
        dgtest = ftol * dginit;
        width = stpmax - stpmin;
        width1 = width / 0.5;
        stx = 0.0;
        fx = finit;
        dgx = dginit;
        sty = 0.0;
        fy = finit;
        dgy = dginit;
        set_interval();
        return 0;
    }

    int next(double f, double dg) {
        ++nfev;
        const double ftest1 = finit + stp * dgtest;
        int info = 0;
        if ((brackt && (stp <= stmin || stp >= stmax)) || infoc == 0) info = 6;
        if (stp == stpmax && f <= ftest1 && dg <= dgtest) info = 5;
        if (stp == stpmin && (f > ftest1 || dg >= dgtest)) info = 4;
        if (nfev >= maxfev) info = 3;
        if (brackt && stmax - stmin <= xtol * stmax) info = 2;
        if (f <= ftest1 && std::fabs(dg) <= gtol * (-dginit)) info = 1;
        if (info != 0) return info;
        if (stage1 && f <= ftest1 && dg >= std::min(ftol, gtol) * dginit) stage1 = false;
        if (stage1 && f <= fx && f > ftest1) {
            // modified function psi(stp) = f - stp * dgtest while no step
            // has sufficient decrease yet
            double fm = f - stp * dgtest, fxm = fx - stx * dgtest, fym = fy - sty * dgtest;
            double dgm = dg - dgtest, dgxm = dgx - dgtest, dgym = dgy - dgtest;
            MtStep::step(stx, fxm, dgxm, sty, fym, dgym, stp, fm, dgm, brackt, stmin, stmax,
                         infoc);
            fx = fxm + stx * dgtest;
            fy = fym + sty * dgtest;
            dgx = dgxm + dgtest;
            dgy = dgym + dgtest;
        } else {
            MtStep::step(stx, fx, dgx, sty, fy, dgy, stp, f, dg, brackt, stmin, stmax, infoc);
        }
        if (brackt) {
            if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
            width1 = width;
            width = std::fabs(sty - stx);
        }
        set_interval();
        return 0;
    }

private:
    void set_interval() {
        if (brackt) {
            stmin = std::min(stx, sty);
            stmax = std::max(stx, sty);
        } else {
            stmin = stx;
            stmax = stp + 4.0 * (stp - stx);
        }
        stp = std::max(stp, stpmin);
        stp = std::min(stp, stpmax);
        if ((brackt && (stp <= stmin || stp >= stmax)) || nfev >= maxfev - 1 || infoc == 0 ||
            (brackt && stmax - stmin <= xtol * stmax))
            stp = stx;
    }

    int infoc = 1;
    bool brackt = false, stage1 = true;
    double finit = 0, dginit = 0, dgtest = 0, width = 0, width1 = 0;
    double stx = 0, fx = 0, dgx = 0, sty = 0, fy = 0, dgy = 0, stmin = 0, stmax = 0;
};

struct LbfgsStats {
    int iterations = 0;      // accepted line searches
    int evaluations = 0;     // energy / gradient evaluations
    int line_search_info = 0;  // outcome of the last line search
    double f0 = 0.0, f = 0.0, gnorm0 = 0.0, gnorm = 0.0;
};

// Minimise f over x (nx device doubles) with m correction pairs for at most
// max_iter iterations.  FG(x, g) evaluates f at device x, writes the gradient
// to device g and returns f.  V is the backend: dot(n, a, b) -> double,
// axpy(n, a, x, y): y += a x, scal(n, a, x): x *= a, copy(n, src, dst),
// neg(n, src, dst): dst = -src, and dalloc / dfree.
template <class B, class FG>
int lbfgs_minimize(B& be, double* x, long long nx, int m, int max_iter, FG&& fg,
                   LbfgsStats& st) {
    const int n = (int)nx;
    std::vector<double*> bufs;
    auto alloc = [&](void) {
        double* p = (double*)be.dalloc(sizeof(double) * (size_t)n);
        bufs.push_back(p);
        return p;
    };
    struct Free {
        B& be;
        std::vector<double*>& b;
        ~Free() {
            for (double* p : b) be.dfree(p);
        }
    } guard{be, bufs};
    double* g = alloc();
    double* q = alloc();
    double* x0 = alloc();
    double* g0 = alloc();
    std::vector<double*> S(m), Y(m);
    for (int k = 0; k < m; ++k) {
        S[k] = alloc();
        Y[k] = alloc();
    }
    std::vector<double> rho(m), alpha(m);
    double f = fg(x, g);
    ++st.evaluations;
    st.f0 = f;
    st.gnorm0 = std::sqrt(be.vdot(n, g, g));
    int stored = 0, newest = -1;
    for (int iter = 0;; ++iter) {
        // search direction q = -H g (two-loop recursion, newest pair first)
        be.vscale(n, -1.0, g, q);
        for (int j = 0; j < stored; ++j) {
            const int k = (newest - j + m) % m;
            alpha[k] = rho[k] * be.vdot(n, S[k], q);
            be.vaxpy(n, -alpha[k], Y[k], q);
        }
        if (stored > 0) {
            const double ys = be.vdot(n, Y[newest], S[newest]);
            const double yy = be.vdot(n, Y[newest], Y[newest]);
            be.vscale(n, ys / yy, q, q);
        }
        for (int j = stored - 1; j >= 0; --j) {
            const int k = (newest - j + m) % m;
            const double beta = rho[k] * be.vdot(n, Y[k], q);
            be.vaxpy(n, alpha[k] - beta, S[k], q);
        }
        be.vscale(n, 1.0, x, x0);
        be.vscale(n, 1.0, g, g0);
        const double f0 = f;
        const double dg0 = be.vdot(n, g, q);
        MoreThuente ls;
        const double stp0 = iter == 0 ? 1.0 / std::sqrt(be.vdot(n, g, g)) : 1.0;
        int info = ls.start(f0, dg0, stp0);
        while (info == 0) {
            // x = x0 + stp q
            be.vscale(n, 1.0, x0, x);
            be.vaxpy(n, ls.stp, q, x);
            f = fg(x, g);
            ++st.evaluations;
            info = ls.next(f, be.vdot(n, g, q));
        }
        st.line_search_info = info;
        st.iterations = iter + 1;
        st.f = f;
        const double gnorm = std::sqrt(be.vdot(n, g, g));
        st.gnorm = gnorm;
        if (info != 1) break;  // line search failed: x holds its last trial
        if (!(gnorm > 0.0)) break;
        if (iter + 1 > max_iter) break;
        // new pair s = x - x0, y = g - g0
        newest = (newest + 1) % m;
        be.vscale(n, 1.0, x, S[newest]);
        be.vaxpy(n, -1.0, x0, S[newest]);
        be.vscale(n, 1.0, g, Y[newest]);
        be.vaxpy(n, -1.0, g0, Y[newest]);
        rho[newest] = 1.0 / be.vdot(n, Y[newest], S[newest]);
        stored = std::min(stored + 1, m);
    }
    return 0;
}

}  // namespace vgc

// Semi-discrete optimal transport solver (host driver, written once against a
// backend): uniform density on a box to Dirac masses nu_i at the seeds.
//
// Damped Newton with the step rule of Kitagawa, Merigot and Thibert
// ("Convergence of a Newton algorithm for semi-discrete optimal transport",
// JEMS 2019): from weights w, solve H d = nu - m(w) with the Laguerre
// Hessian H (ot_kernels.h) by Jacobi-preconditioned conjugate gradients in
// FP64; accept w + t d for the largest t in {1, 1/2, 1/4, ...} such that
// every cell keeps mass >= eps0 = min(min nu, min m(w0)) / 2 and
// |m - nu|_2 decreases by the factor (1 - t/2).  Stops when
// max_i |m_i - nu_i| / nu_i <= tol.  The Exploragram solver behind
// OptimalTransportMap3d::optimize (called at
// examples/exploragram/compute_OTM/main.cpp:283-297) is not part of the
// geogram tree; this driver follows the published method.
#pragma once

#include <cmath>
#include <cstdio>
#include <vector>

#include "pipeline.h"

namespace vgc {

struct OtStats {
    int newton_iters = 0;
    int cg_iters = 0;
    int cell_passes = 0;
    int halvings = 0;
    int used_fallback_start = 0;
    double err0 = 0, final_err = 0;
    double t_total_ms = 0, t_cells_ms = 0, t_cg_ms = 0;
    int converged = 0;
};

template <class B>
class OtSolver {
public:
    OtSolver(B& be, Engine<B>& E) : be_(be), E_(E) {}
    ~OtSolver() {
        for (auto& s : slots_)
            if (s.p) be_.dfree(s.p);
    }

    // d_xyz (3n), d_nu (n), d_w (n, in: start, out: solution): device memory,
    // caller order.  cg_tol: relative residual of the Newton systems.  judge:
    // also measure the zero start and keep whichever start has the smaller
    // |m - nu|_2 (one more cell pass).
    int solve(Problem pb, const double* d_nu, double* d_w, double tol, int max_newton,
              double cg_tol, int cg_max, OtStats& st, bool judge = false, bool verbose = false) {
        auto t0 = std::chrono::steady_clock::now();
        const int n = pb.n;
        pb.want_neighbors = true;
        pb.d_w = d_w;
        double* d_b = (double*)slot(0, sizeof(double) * n);
        double* d_x = (double*)slot(1, sizeof(double) * n);
        double* d_r = (double*)slot(2, sizeof(double) * n);
        double* d_z = (double*)slot(3, sizeof(double) * n);
        double* d_p = (double*)slot(4, sizeof(double) * n);
        double* d_q = (double*)slot(5, sizeof(double) * n);
        double* d_diag = (double*)slot(6, sizeof(double) * n);
        double* d_wt = (double*)slot(7, sizeof(double) * n);
        double* d_scal = (double*)slot(8, sizeof(double) * 8);
        double* d_cg = (double*)slot(10, sizeof(double) * 8);
        EngineStats es;

        // 0: cells of w computed; 1: some cell empty or failed (the masses
        // are not all valid); -1: error
        long long rej_empty = 0, rej_failed = 0, rej_test = 0;
        auto cells = [&](const double* w, bool trial) -> int {
            auto tc = std::chrono::steady_clock::now();
            pb.d_w = w;
            pb.abort_on_empty = trial;
            const long long failed0 = es.nb_failed, empty0 = es.nb_empty;
            int rc = E_.cells(pb, es);
            st.cell_passes++;
            st.t_cells_ms += ms_since(tc);
            if (rc < 0) return -1;
            if (es.nb_failed > failed0) {
                rej_failed++;
                return 1;
            }
            if (rc > 0 || es.nb_empty > empty0) {
                rej_empty++;
                return 1;
            }
            return 0;
        };
        // mass statistics: [0] max |m-nu|/nu, [1] |m-nu|_2^2, [2] min m
        auto masses = [&](double out[3]) {
            be_.ot_mass_stats(n, E_.d_vol, d_nu, d_scal);
            be_.d2h(out, d_scal, sizeof(double) * 3);
        };

        // zero start first when judging: the usual winner, the given start,
        // is then the last one measured and its CSR feeds the first step
        double m0[3] = {0, 0, 0};
        if (judge) {
            be_.dzero(d_wt, sizeof(double) * n);
            if (cells(d_wt, false)) return fail("the cells of the zero start failed");
            masses(m0);
        }
        int rc0 = cells(d_w, true);
        if (rc0 < 0) return -1;
        double ms[3] = {0, 0, 0};
        if (rc0 == 0) masses(ms);
        if (rc0 > 0 || !(ms[2] > 0.0) || (judge && m0[1] < ms[1])) {
            // the start leaves a cell empty, or is worse than w = 0: use w = 0
            be_.dzero(d_w, sizeof(double) * n);
            st.used_fallback_start = 1;
            if (cells(d_w, false)) return fail("the cells of the zero start failed");
            masses(ms);
        }
        double numin = 0;
        be_.ot_min(n, d_nu, d_scal + 3);
        be_.d2h(&numin, d_scal + 3, sizeof(double));
        const double eps0 = 0.5 * std::min(numin, ms[2]);
        st.err0 = ms[0];
        double err = ms[0], g2 = ms[1];
        while (err > tol && st.newton_iters < max_newton) {
            // Newton system H d = nu - m
            auto tg = std::chrono::steady_clock::now();
            const long long nnz = E_.total_neighbors;
            double* d_val = (double*)slot(9, sizeof(double) * std::max<long long>(nnz, 1));
            be_.ot_hessian(n, pb.d_xyz, E_.d_ptr, E_.d_nbr, E_.d_area, d_val, d_diag);
            be_.ot_rhs(n, d_nu, E_.d_vol, d_b);
            int it = pcg(n, E_.d_ptr, E_.d_nbr, d_val, d_diag, d_b, d_x, d_r, d_z, d_p, d_q, d_cg,
                         cg_tol, cg_max);
            st.cg_iters += it;
            st.t_cg_ms += ms_since(tg);
            // damped step
            double t = 1.0;
            bool ok = false;
            for (int h = 0; h < 40; ++h) {
                be_.ot_axpy(n, d_w, t, d_x, d_wt);
                // a trial with an empty cell fails the mass test whatever the
                // other cells: its evaluation stops after the first pass
                const int rc = cells(d_wt, true);
                if (rc < 0) return -1;
                double m2[3] = {0, 0, 0};
                if (rc == 0) masses(m2);
                if (rc == 0 && !(m2[2] >= eps0 &&
                                 std::sqrt(m2[1]) <= (1.0 - 0.5 * t) * std::sqrt(g2)))
                    rej_test++;
                if (rc == 0 && m2[2] >= eps0 &&
                    std::sqrt(m2[1]) <= (1.0 - 0.5 * t) * std::sqrt(g2)) {
                    err = m2[0];
                    g2 = m2[1];
                    ok = true;
                    break;
                }
                t *= 0.5;
                st.halvings++;
            }
            if (verbose) {
                std::fprintf(stderr,
                             "[vgc ot] newton %d: err %.3e step %.4g cg %d cells %.0f ms cg %.0f ms"
                             " rejected trials: %lld failed cells (%llu over capacity), %lld empty,"
                             " %lld mass test\n",
                             st.newton_iters + 1, ok ? err : -1.0, ok ? t : 0.0, it, st.t_cells_ms,
                             st.t_cg_ms, rej_failed, es.nb_fail_capacity, rej_empty, rej_test);
                std::fflush(stderr);
            }
            if (!ok) break;
            be_.d2d(d_w, d_wt, sizeof(double) * n);
            st.newton_iters++;
        }
        be_.sync();
        st.final_err = err;
        st.converged = err <= tol;
        st.t_total_ms = ms_since(t0);
        return 0;
    }

private:
    // Jacobi-preconditioned CG from x = 0; returns the iterations done.  The
    // iterations run on the device (scalars in device memory), kCgCheck at a
    // time as one graph launch; the host reads the state between them.
    static constexpr int kCgCheck = 16;
    int pcg(int n, const long long* ptr, const int* col, const double* val, const double* diag,
            const double* b, double* x, double* r, double* z, double* p, double* q, double* scal,
            double tol, int maxit) {
        be_.cg_start(n, b, diag, x, r, z, p, scal, tol * tol);
        int issued = 0;
        double h[8];
        be_.d2h(h, scal, sizeof(double) * 8);
        while (h[6] == 0.0 && issued < maxit) {
            const int k = std::min(kCgCheck, maxit - issued);
            be_.cg_iters(k, n, ptr, col, val, diag, x, r, z, p, q, scal);
            issued += k;
            be_.d2h(h, scal, sizeof(double) * 8);
        }
        return (int)h[7];
    }

    int fail(const char* msg) {
        E_.error = msg;
        return -1;
    }
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    void* slot(int id, size_t bytes) {
        if ((int)slots_.size() <= id) slots_.resize(id + 1);
        Buf& b = slots_[id];
        if (b.bytes < bytes) {
            if (b.p) be_.dfree(b.p);
            b.bytes = std::max(bytes, b.bytes + b.bytes / 4);
            b.p = be_.dalloc(b.bytes);
        }
        return b.p;
    }
    static double ms_since(std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    }

    B& be_;
    Engine<B>& E_;
    std::vector<Buf> slots_;
};

}  // namespace vgc

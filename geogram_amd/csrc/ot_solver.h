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

#include "ot_kernels.h"
#include "pipeline.h"

namespace vgc {

struct OtStats {
    int newton_iters = 0;
    int cg_iters = 0;
    int cell_passes = 0;
    int halvings = 0;
    int used_fallback_start = 0;
    int repair_rounds = 0;
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
    // forcing term of the inexact Newton variant: CG relative residual for a
    // step from max relative mass error err, once the previous step was a
    // full one (the damped phase needs accurate directions: measured, loose
    // directions there cost more step halvings than they save)
    static double forcing(double err, double cg_tol, bool full_step) {
        return full_step ? std::max(cg_tol, std::min(0.01, 0.01 * err)) : cg_tol;
    }

    int solve(Problem pb, const double* d_nu, double* d_w, double tol, int max_newton,
              double cg_tol, int cg_max, OtStats& st, bool judge = false, bool verbose = false,
              bool inexact = false, bool repair = false) {
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
        double* d_cg = (double*)slot(10, sizeof(double) * 16);
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
        // with repair, full passes (every empty cell found): the weights of
        // empty cells are raised until each seed lies in its cell
        int rc0 = cells(d_w, !repair);
        if (rc0 < 0) return -1;
        if (repair) {
            const double* b = pb.box.data();
            const double h2 = std::pow((b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]) / n, 2.0 / 3.0);
            for (int round = 0; round < kRepairRounds && rc0 > 0 && rej_failed == 0; ++round) {
                E_.repair_empty(pb, kRepairMargin * h2, d_wt);
                be_.d2d(d_w, d_wt, sizeof(double) * n);
                st.repair_rounds++;
                const long long e0 = es.nb_empty;
                rc0 = cells(d_w, false);
                if (rc0 < 0) return -1;
                if (verbose)
                    std::fprintf(stderr, "[vgc ot] start repair %d: %lld empty cells\n", round + 1,
                                 es.nb_empty - e0);
            }
        }
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
        double last_t = 0.0;  // step of the previous Newton iteration
        while (err > tol && st.newton_iters < max_newton) {
            // Newton system H d = nu - m
            auto tg = std::chrono::steady_clock::now();
            const long long nnz = E_.total_neighbors;
            double* d_val = (double*)slot(9, sizeof(double) * std::max<long long>(nnz, 1));
            double* d_per = nullptr;
            if (pb.periodic) {
                const double per[3] = {pb.box[3] - pb.box[0], pb.box[4] - pb.box[1],
                                       pb.box[5] - pb.box[2]};
                d_per = (double*)slot(29, sizeof(double) * 3);
                be_.h2d(d_per, per, sizeof(double) * 3);
            }
            be_.ot_hessian(n, pb.d_xyz, E_.d_ptr, E_.d_nbr, E_.d_area, d_val, d_diag, d_per);
            // Newton system H d = nu - m, projected on the range of H (the
            // masses and nu can differ in sum by rounding; the constants are
            // H's kernel, and CG on an inconsistent system stalls)
            be_.ot_rhs(n, d_nu, E_.d_vol, d_b, d_scal + 4);
            const double ctol = inexact ? forcing(err, cg_tol, last_t == 1.0) : cg_tol;
            int it;
            if (!pb.periodic) {
                // the system in the cells' Morton order: neighbouring rows,
                // neighbouring columns, so the product's gathers stay in cache
                const MortonSystem M = to_morton(n, nnz, d_val, d_diag, d_b);
                it = pcg(n, M.ptr, M.col, M.val, M.diag, M.b, M.x, d_r, d_z, d_p, d_q, d_cg, ctol,
                         cg_max);
                be_.scatter_f64(n, E_.d_orig, M.x, d_x);
            } else {
                it = pcg(n, E_.d_ptr, E_.d_nbr, d_val, d_diag, d_b, d_x, d_r, d_z, d_p, d_q, d_cg,
                         ctol, cg_max);
            }
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
            last_t = t;
            be_.d2d(d_w, d_wt, sizeof(double) * n);
            st.newton_iters++;
        }
        be_.sync();
        st.final_err = err;
        st.converged = err <= tol;
        st.t_total_ms = ms_since(t0);
        return 0;
    }

    // n_inst independent problems (pb.inst_n, pb.box) solved together: one
    // cell pass covers every instance, the Laguerre system is block diagonal
    // and its PCG keeps per-instance scalars (segmented reductions), each
    // instance has its own damped-Newton state and step.  st[k] per instance.
    int solve_batch(Problem pb, const double* d_nu, double* d_w, double tol, int max_newton,
                    double cg_tol, int cg_max, std::vector<OtStats>& st, bool verbose = false,
                    bool inexact = false) {
        auto t0 = std::chrono::steady_clock::now();
        const int n = pb.n, K = pb.n_inst;
        st.assign(K, OtStats());
        pb.want_neighbors = true;
        pb.abort_on_empty = false;
        // instance ranges (caller order) and the reduction block table
        std::vector<long long> off(K + 1, 0);
        for (int k = 0; k < K; ++k) off[k + 1] = off[k] + pb.inst_n[k];
        std::vector<int> blk_inst, blk_idx, cnt(K), first(K), p0(K);
        for (int k = 0; k < K; ++k) {
            p0[k] = (int)off[k];
            cnt[k] = (int)std::min<long long>(kSegMaxBlocks, std::max<long long>(1, (pb.inst_n[k] + 2047) / 2048));
            first[k] = (int)blk_inst.size();
            for (int b = 0; b < cnt[k]; ++b) {
                blk_inst.push_back(k);
                blk_idx.push_back(b);
            }
        }
        const int nblk = (int)blk_inst.size();
        long long* d_off = (long long*)slot(20, sizeof(long long) * (K + 1));
        int* d_tab = (int*)slot(21, sizeof(int) * (2 * (size_t)nblk + 3 * (size_t)K));
        int* d_inst = (int*)slot(22, sizeof(int) * (size_t)n);
        be_.h2d(d_off, off.data(), sizeof(long long) * (K + 1));
        std::vector<int> tab;
        tab.insert(tab.end(), blk_inst.begin(), blk_inst.end());
        tab.insert(tab.end(), blk_idx.begin(), blk_idx.end());
        tab.insert(tab.end(), cnt.begin(), cnt.end());
        tab.insert(tab.end(), first.begin(), first.end());
        tab.insert(tab.end(), p0.begin(), p0.end());
        be_.h2d(d_tab, tab.data(), sizeof(int) * tab.size());
        be_.fill_inst(d_inst, n, d_tab + 2 * nblk + 2 * K, K);
        SegTable T{K, nblk, d_off, d_tab, d_tab + nblk, d_tab + 2 * nblk, d_tab + 2 * nblk + K, d_inst};

        double* d_b = (double*)slot(0, sizeof(double) * n);
        double* d_x = (double*)slot(1, sizeof(double) * n);
        double* d_r = (double*)slot(2, sizeof(double) * n);
        double* d_z = (double*)slot(3, sizeof(double) * n);
        double* d_p = (double*)slot(4, sizeof(double) * n);
        double* d_q = (double*)slot(5, sizeof(double) * n);
        double* d_diag = (double*)slot(6, sizeof(double) * n);
        double* d_wt = (double*)slot(7, sizeof(double) * n);
        double* d_ms = (double*)slot(23, sizeof(double) * 3 * (size_t)K);
        double* d_cg = (double*)slot(24, sizeof(double) * kCgStride * (size_t)K);
        double* d_t = (double*)slot(25, sizeof(double) * (size_t)K);
        int* d_sel = (int*)slot(26, sizeof(int) * (size_t)K);
        EngineStats es;
        std::vector<double> ms(3 * (size_t)K), tk(K), err(K), g2(K), eps0(K);
        std::vector<double> last_t(K, 0.0);  // each instance's previous step
        std::vector<int> sel(K), bad;
        int passes = 0;
        std::vector<int> compute;  // instances whose cells a pass computes
        // trial passes: an instance with an empty cell fails its mass test
        // whatever its other cells, so its evaluation stops there
        auto cells = [&](const double* w, bool trial = false) -> int {
            auto tc = std::chrono::steady_clock::now();
            pb.d_w = w;
            pb.abort_on_empty = trial;
            pb.active = compute.empty() ? nullptr : &compute;
            int rc = E_.cells(pb, es);
            ++passes;
            bad = E_.inst_bad;
            be_.seg_mass_stats(T, E_.d_vol, d_nu, d_ms);
            be_.d2h(ms.data(), d_ms, sizeof(double) * 3 * (size_t)K);
            const double dt = ms_since(tc);
            for (auto& s : st) s.t_cells_ms += dt;
            return rc < 0 ? -1 : 0;
        };
        // start: instances whose start leaves a cell empty or failed restart from 0
        if (cells(d_w)) return -1;
        bool any0 = false;
        for (int k = 0; k < K; ++k) {
            sel[k] = bad[k] > 0 || !(ms[3 * k + 2] > 0.0);
            if (sel[k]) {
                st[k].used_fallback_start = 1;
                any0 = true;
            }
        }
        if (any0) {
            be_.dzero(d_wt, sizeof(double) * n);
            be_.h2d(d_sel, sel.data(), sizeof(int) * K);
            be_.seg_copy(n, d_inst, d_sel, d_wt, d_w);
            if (cells(d_w)) return -1;
            for (int k = 0; k < K; ++k)
                if (bad[k] > 0) return fail("the cells of the zero start failed");
        }
        std::vector<double> numin(3 * (size_t)K);
        be_.seg_mass_stats(T, d_nu, d_nu, d_ms);  // [2] = min nu
        be_.d2h(numin.data(), d_ms, sizeof(double) * 3 * (size_t)K);
        be_.seg_mass_stats(T, E_.d_vol, d_nu, d_ms);
        std::vector<int> active(K);
        for (int k = 0; k < K; ++k) {
            eps0[k] = 0.5 * std::min(numin[3 * k + 2], ms[3 * k + 2]);
            err[k] = ms[3 * k + 0];
            g2[k] = ms[3 * k + 1];
            st[k].err0 = err[k];
            active[k] = err[k] > tol;
        }
        int* d_act = (int*)slot(27, sizeof(int) * (size_t)K);
        for (int it = 0; it < max_newton; ++it) {
            bool any = false;
            for (int k = 0; k < K; ++k) any = any || active[k];
            if (!any) break;
            auto tg = std::chrono::steady_clock::now();
            const long long nnz = E_.total_neighbors;
            double* d_val = (double*)slot(9, sizeof(double) * std::max<long long>(nnz, 1));
            be_.ot_hessian(n, pb.d_xyz, E_.d_ptr, E_.d_nbr, E_.d_area, d_val, d_diag);
            be_.h2d(d_act, active.data(), sizeof(int) * K);
            // per instance projected on the range of its Laplacian (as solve)
            be_.seg_rhs(T, n, d_inst, d_act, d_nu, E_.d_vol, d_b);
            std::vector<double> tol2(K);
            for (int k = 0; k < K; ++k) {
                const double tk2 = inexact ? forcing(err[k], cg_tol, last_t[k] == 1.0) : cg_tol;
                tol2[k] = tk2 * tk2;
            }
            double* d_tol2 = (double*)slot(28, sizeof(double) * (size_t)K);
            be_.h2d(d_tol2, tol2.data(), sizeof(double) * (size_t)K);
            // Morton order (instances keep their ranges: the order leads
            // with the instance), restricted to the rows of the instances
            // still active: the others' CG would only skip their rows
            int n_sub = n;
            const int* omap = E_.d_orig;
            SegTable Ts = T;
            int n_act = 0;
            for (int k = 0; k < K; ++k) n_act += active[k] ? 1 : 0;
            if (n_act < K) {
                std::vector<long long> soff(K + 1, 0);
                std::vector<int> sbi, sbx, scnt(K, 0), sfirst(K, 0), sp0(K);
                for (int k = 0; k < K; ++k) {
                    sp0[k] = (int)soff[k];
                    soff[k + 1] = soff[k] + (active[k] ? pb.inst_n[k] : 0);
                    scnt[k] = active[k] ? cnt[k] : 0;
                    sfirst[k] = (int)sbi.size();
                    for (int b = 0; b < scnt[k]; ++b) {
                        sbi.push_back(k);
                        sbx.push_back(b);
                    }
                }
                n_sub = (int)soff[K];
                const int nbs = (int)sbi.size();
                long long* d_soff = (long long*)slot(41, sizeof(long long) * (K + 1));
                int* d_stab = (int*)slot(42, sizeof(int) * (2 * (size_t)nbs + 3 * (size_t)K));
                int* d_sinst = (int*)slot(43, sizeof(int) * std::max(n_sub, 1));
                int* d_omap = (int*)slot(44, sizeof(int) * std::max(n_sub, 1));
                be_.h2d(d_soff, soff.data(), sizeof(long long) * (K + 1));
                std::vector<int> stab;
                stab.insert(stab.end(), sbi.begin(), sbi.end());
                stab.insert(stab.end(), sbx.begin(), sbx.end());
                stab.insert(stab.end(), scnt.begin(), scnt.end());
                stab.insert(stab.end(), sfirst.begin(), sfirst.end());
                stab.insert(stab.end(), sp0.begin(), sp0.end());
                be_.h2d(d_stab, stab.data(), sizeof(int) * stab.size());
                be_.fill_inst(d_sinst, n_sub, d_stab + 2 * nbs + 2 * K, K);
                be_.range_map(n_sub, d_sinst, d_soff, d_off, E_.d_orig, d_omap);
                omap = d_omap;
                Ts = SegTable{K, nbs, d_soff, d_stab, d_stab + nbs, d_stab + 2 * nbs,
                              d_stab + 2 * nbs + K, d_sinst};
            }
            const MortonSystem M = to_morton(n_sub, nnz, d_val, d_diag, d_b, omap);
            be_.cgb_start(Ts, M.b, M.diag, M.x, d_r, d_z, d_p, d_cg, d_tol2);
            std::vector<double> h(kCgStride * (size_t)K);
            be_.d2h(h.data(), d_cg, sizeof(double) * kCgStride * K);
            int issued = 0;
            auto all_done = [&]() {
                for (int k = 0; k < K; ++k)
                    if (h[kCgStride * k + 6] == 0.0) return false;
                return true;
            };
            while (!all_done() && issued < cg_max) {
                const int kk = std::min(kCgCheck, cg_max - issued);
                be_.cgb_iters(kk, Ts, n_sub, M.ptr, M.col, M.val, M.diag, M.x, d_r, d_z, d_p, d_q,
                              d_cg);
                issued += kk;
                be_.d2h(h.data(), d_cg, sizeof(double) * kCgStride * K);
            }
            if (n_sub < n) be_.dzero(d_x, sizeof(double) * n);
            be_.scatter_f64(n_sub, omap, M.x, d_x);
            const double tcg = ms_since(tg);
            for (int k = 0; k < K; ++k) {
                st[k].t_cg_ms += tcg;
                if (active[k]) st[k].cg_iters += (int)h[kCgStride * k + 7];
            }
            // per-instance damped steps
            std::vector<int> pending(active), accepted(K, 0);
            for (int k = 0; k < K; ++k) tk[k] = active[k] ? 1.0 : 0.0;
            be_.d2d(d_wt, d_w, sizeof(double) * n);
            for (int hv = 0; hv < 40; ++hv) {
                bool anyp = false;
                for (int k = 0; k < K; ++k) anyp = anyp || pending[k];
                if (!anyp) break;
                compute = pending;  // the other instances keep their cells
                be_.h2d(d_t, tk.data(), sizeof(double) * K);
                be_.seg_axpy(n, d_inst, d_w, d_t, d_x, d_wt);
                if (cells(d_wt, true)) return -1;
                for (int k = 0; k < K; ++k) {
                    if (!pending[k]) continue;
                    const bool okk = bad[k] == 0 && ms[3 * k + 2] >= eps0[k] &&
                                     std::sqrt(ms[3 * k + 1]) <= (1.0 - 0.5 * tk[k]) * std::sqrt(g2[k]);
                    if (okk) {
                        pending[k] = 0;
                        accepted[k] = 1;
                        last_t[k] = tk[k];
                        err[k] = ms[3 * k + 0];
                        g2[k] = ms[3 * k + 1];
                    } else {
                        tk[k] *= 0.5;
                        st[k].halvings++;
                    }
                }
                // accepted instances keep their new weights from here on
                for (int k = 0; k < K; ++k) sel[k] = accepted[k] && tk[k] != 0.0;
                be_.h2d(d_sel, sel.data(), sizeof(int) * K);
                be_.seg_copy(n, d_inst, d_sel, d_wt, d_w);
                for (int k = 0; k < K; ++k)
                    if (sel[k]) tk[k] = 0.0;
            }
            for (int k = 0; k < K; ++k) {
                if (!active[k]) continue;
                if (accepted[k]) {
                    st[k].newton_iters++;
                    active[k] = err[k] > tol;
                } else {
                    active[k] = 0;  // line search exhausted
                }
            }
            if (verbose) {
                int na = 0;
                double emax = 0.0;
                for (int k = 0; k < K; ++k) {
                    na += active[k];
                    emax = std::max(emax, err[k]);
                }
                std::fprintf(stderr, "[vgc ot batch] newton %d: %d of %d instances active, max err %.3e\n",
                             it + 1, na, K, emax);
                std::fflush(stderr);
            }
        }
        be_.sync();
        const double tt = ms_since(t0);
        for (int k = 0; k < K; ++k) {
            st[k].final_err = err[k];
            st[k].converged = err[k] <= tol;
            st[k].cell_passes = passes;
            st[k].t_total_ms = tt;
        }
        return 0;
    }

private:
    // Jacobi-preconditioned CG from x = 0; returns the iterations done.  The
    // iterations run on the device (scalars in device memory), kCgCheck at a
    // time as one graph launch; the host reads the state between them.
    static constexpr int kCgCheck = 16;
    static constexpr int kRepairRounds = 16;
    // margin of the repair, in units of h^2 (h = (box volume / n)^(1/3))
    static constexpr double kRepairMargin = 0.01;
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

    // The Newton system H d = b of the last cell pass (caller-order CSR of
    // the engine, values val, diagonal diag) permuted to the pass's Morton
    // order; x is its solution buffer, in that order.
    struct MortonSystem {
        const long long* ptr;
        const int* col;
        const double* val;
        const double* diag;
        const double* b;
        double* x;
    };
    MortonSystem to_morton(int n, long long nnz, const double* val, const double* diag,
                           const double* b, const int* orig = nullptr) {
        if (!orig) orig = E_.d_orig;  // rows: orig[m] is the caller's row of row m
        int* inv = (int*)slot(33, sizeof(int) * (size_t)E_.total_points());
        int* len = (int*)slot(34, sizeof(int) * (size_t)n);
        long long* ptr = (long long*)slot(35, sizeof(long long) * ((size_t)n + 1));
        int* col = (int*)slot(36, sizeof(int) * (size_t)std::max<long long>(nnz, 1));
        double* v = (double*)slot(37, sizeof(double) * (size_t)std::max<long long>(nnz, 1));
        double* dg = (double*)slot(38, sizeof(double) * (size_t)n);
        double* bb = (double*)slot(39, sizeof(double) * (size_t)n);
        double* x = (double*)slot(40, sizeof(double) * (size_t)n);
        be_.inv_perm(n, orig, inv);
        be_.perm_row_len(n, orig, E_.d_ptr, len);
        be_.scan_i32(len, ptr, n);
        be_.perm_rows(n, orig, inv, E_.d_ptr, E_.d_nbr, val, ptr, col, v);
        be_.gather_f64(n, orig, diag, dg);
        be_.gather_f64(n, orig, b, bb);
        return MortonSystem{ptr, col, v, dg, bb, x};
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

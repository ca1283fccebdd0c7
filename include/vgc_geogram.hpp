// C++ mirror of geogram's entry points for the meshless cell path, header-only
// over the C ABI of vgc.h (link with libvgc.so).  Class and method names follow
// geogram so a caller switches by changing the type:
//   NearestNeighborSearch          points/nn_search.h:69-200
//   RestrictedVoronoiDiagram       voronoi/RVD.h:267-290 (volumetric, box domain)
//   CentroidalVoronoiTesselation   voronoi/CVT.h:139-164
//   OptimalTransportMap3d          examples/exploragram/compute_OTM/main.cpp:283-297
// Errors are reported as std::runtime_error (geogram's shims would map them to
// geo_assert or its own exceptions).
#ifndef VGC_GEOGRAM_HPP
#define VGC_GEOGRAM_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "vgc.h"

namespace vgc {
namespace geogram {

typedef unsigned int index_t;

class Context {
public:
    explicit Context(int device = 0) {
        int rc = vgc_create(&ctx_, device);
        if (rc != VGC_OK) throw std::runtime_error("vgc: no usable GPU (vgc_create failed)");
    }
    ~Context() { vgc_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    vgc_ctx* get() const { return ctx_; }
    void check(int rc) const {
        if (rc < 0) throw std::runtime_error(std::string("vgc: ") + vgc_last_error(ctx_));
    }

private:
    vgc_ctx* ctx_ = nullptr;
};

// NearestNeighborSearch (points/nn_search.h:69-200): exact FP64 squared
// distances, ascending; equal distances by ascending index.
class NearestNeighborSearch {
public:
    explicit NearestNeighborSearch(Context& ctx) : ctx_(ctx) {}
    // nn_search.h:96
    void set_points(index_t nb_points, const double* points) {
        n_ = nb_points;
        pts_ = points;
    }
    index_t nb_points() const { return n_; }
    // nn_search.h:129-134: nb_neighbors nearest points of query_point
    void get_nearest_neighbors(index_t nb_neighbors, const double* query_point,
                               index_t* neighbors, double* neighbors_sq_dist) const {
        get_nearest_neighbors(nb_neighbors, 1, query_point, neighbors, neighbors_sq_dist);
    }
    // nn_search.h:180-185: by the index of a stored point (itself included,
    // at distance 0)
    void get_nearest_neighbors(index_t nb_neighbors, index_t query_point, index_t* neighbors,
                               double* neighbors_sq_dist) const {
        get_nearest_neighbors(nb_neighbors, 1, pts_ + 3 * (size_t)query_point, neighbors,
                              neighbors_sq_dist);
    }
    // nn_search.h:192-198
    index_t get_nearest_neighbor(const double* query_point) const {
        index_t result;
        double sq_dist;
        get_nearest_neighbors(1, query_point, &result, &sq_dist);
        return result;
    }
    // the same for nq query points in one call: neighbors / sq_dist are
    // nq * nb_neighbors (the form a GPU caller should use)
    void get_nearest_neighbors(index_t nb_neighbors, index_t nq, const double* query_points,
                               index_t* neighbors, double* sq_dist) const {
        ctx_.check(vgc_knn_search(ctx_.get(), pts_, n_, query_points, nq, nb_neighbors, 0,
                                  reinterpret_cast<int32_t*>(neighbors), sq_dist));
    }
    // every stored point's nb_neighbors nearest others, with the coincident-
    // point rule of Delaunay_NearestNeighbors::get_neighbors_internal
    // (delaunay/delaunay_nn.cpp:105-143); -1 pads rows that end early
    void get_neighbors_of_all(index_t nb_neighbors, int32_t* neighbors, double* sq_dist) const {
        ctx_.check(vgc_knn(ctx_.get(), pts_, n_, nb_neighbors, 0, neighbors, sq_dist));
    }

private:
    Context& ctx_;
    index_t n_ = 0;
    const double* pts_ = nullptr;
};

// Volumetric restricted Voronoi diagram of seeds in a box.
class RestrictedVoronoiDiagram {
public:
    RestrictedVoronoiDiagram(Context& ctx, const double box[6]) : ctx_(ctx) {
        for (int i = 0; i < 6; ++i) box_[i] = box[i];
    }
    void set_points(index_t nb_points, const double* points) {
        n_ = nb_points;
        pts_ = points;
    }
    // mg[3n], m[n]: caller-allocated and zeroed; the cell integrals are added
    // (RVD_Nd_Impl::compute_centroids_in_volume accumulates the same way)
    void compute_centroids(double* mg, double* m) {
        std::vector<double> vol(n_), g(3 * (size_t)n_);
        vgc_stats st;
        ctx_.check(vgc_cells(ctx_.get(), pts_, nullptr, n_, box_, 0, VGC_NO_NEIGHBORS, vol.data(),
                             g.data(), nullptr, nullptr, nullptr, 0, &st));
        for (index_t i = 0; i < n_; ++i) {
            m[i] += vol[i];
            for (int c = 0; c < 3; ++c) mg[3 * (size_t)i + c] += g[3 * (size_t)i + c];
        }
    }
    // compute_CVT_func_grad_in_volume (voronoi/RVD.h:318): f and g[3n]
    // caller-zeroed; the energy and its gradient are added
    void compute_CVT_func_grad(double& f, double* g) {
        vgc_stats st;
        ctx_.check(vgc_cvt_func_grad(ctx_.get(), pts_, n_, box_, 0, &f, g, &st));
    }
    // Laguerre variant (power distance |x - p|^2 - w)
    void compute_centroids(const double* weights, double* mg, double* m) {
        std::vector<double> vol(n_), g(3 * (size_t)n_);
        vgc_stats st;
        ctx_.check(vgc_cells(ctx_.get(), pts_, weights, n_, box_, 0, VGC_NO_NEIGHBORS, vol.data(),
                             g.data(), nullptr, nullptr, nullptr, 0, &st));
        for (index_t i = 0; i < n_; ++i) {
            m[i] += vol[i];
            for (int c = 0; c < 3; ++c) mg[3 * (size_t)i + c] += g[3 * (size_t)i + c];
        }
    }
    // per-cell neighbour CSR and facet areas (VBW::ConvexCell outputs)
    void compute_cells(std::vector<double>& vol, std::vector<double>& mg,
                       std::vector<uint64_t>& ptr, std::vector<int32_t>& nbr,
                       std::vector<double>& farea, const double* weights = nullptr) {
        vol.resize(n_);
        mg.resize(3 * (size_t)n_);
        ptr.resize((size_t)n_ + 1);
        vgc_stats st;
        uint64_t cap = 24ull * n_ + 64;
        for (int attempt = 0; attempt < 2; ++attempt) {
            nbr.resize(cap);
            farea.resize(cap);
            int rc = vgc_cells(ctx_.get(), pts_, weights, n_, box_, 0, 0, vol.data(), mg.data(),
                               ptr.data(), nbr.data(), farea.data(), cap, &st);
            if (rc == VGC_ERR_CAPACITY) {
                cap = st.nb_neighbors;
                continue;
            }
            ctx_.check(rc);
            nbr.resize(st.nb_neighbors);
            farea.resize(st.nb_neighbors);
            return;
        }
        throw std::runtime_error("vgc: neighbour capacity");
    }

private:
    Context& ctx_;
    double box_[6];
    index_t n_ = 0;
    const double* pts_ = nullptr;
};

// Lloyd relaxation of seeds in a box (security radius on).
class CentroidalVoronoiTesselation {
public:
    CentroidalVoronoiTesselation(Context& ctx, const double box[6]) : ctx_(ctx) {
        for (int i = 0; i < 6; ++i) box_[i] = box[i];
    }
    void set_points(index_t nb_points, const double* points) {
        points_.assign(points, points + 3 * (size_t)nb_points);
    }
    index_t nb_points() const { return index_t(points_.size() / 3); }
    const double* embedding(index_t i) const { return points_.data() + 3 * (size_t)i; }
    virtual void Lloyd_iterations(index_t nb_iter) {
        vgc_stats st;
        ctx_.check(vgc_lloyd(ctx_.get(), points_.data(), nb_points(), box_, nb_iter, 0, 0, &st));
    }
    // voronoi/CVT.h:164 (HLBFGS with m pairs in geogram)
    virtual void Newton_iterations(index_t nb_iter, index_t m = 7) {
        ctx_.check(vgc_cvt_newton(ctx_.get(), points_.data(), nb_points(), box_, nb_iter, m, 0,
                                  &opt_stats_));
    }
    const vgc_opt_stats& optimizer_stats() const { return opt_stats_; }
    virtual ~CentroidalVoronoiTesselation() {}

private:
    Context& ctx_;
    double box_[6];
    std::vector<double> points_;
    vgc_opt_stats opt_stats_{};
};

// Semi-discrete optimal transport, uniform density on a box -> seeds.
class OptimalTransportMap3d {
public:
    OptimalTransportMap3d(Context& ctx, const double box[6]) : ctx_(ctx) {
        for (int i = 0; i < 6; ++i) box_[i] = box[i];
    }
    void set_points(index_t nb_points, const double* points, const double* masses = nullptr) {
        points_.assign(points, points + 3 * (size_t)nb_points);
        const double vol = (box_[3] - box_[0]) * (box_[4] - box_[1]) * (box_[5] - box_[2]);
        if (masses)
            nu_.assign(masses, masses + nb_points);
        else
            nu_.assign(nb_points, vol / nb_points);
        w_.assign(nb_points, 0.0);
    }
    void set_epsilon(double eps) { eps_ = eps; }
    void set_initial_weights(const double* w) {
        w_.assign(w, w + w_.size());
        judge_ = false;
    }
    // Start Newton from the weights an MLP (VGC_MLP_PARAMS floats, e.g. the
    // trained geogram_amd/warmstart_params.npy payload) predicts from the
    // seeds' features; the zero start remains the fallback (judged by the
    // mass residual).  No counterpart in geogram.
    void set_learned_initial_weights(const float* params) {
        const index_t n = index_t(nu_.size());
        std::vector<float> feats((size_t)n * VGC_MLP_FEATURES), y(n);
        ctx_.check(vgc_ot_features(ctx_.get(), points_.data(), nu_.data(), n, box_, 0,
                                   feats.data()));
        ctx_.check(vgc_mlp_forward(ctx_.get(), params, feats.data(), n, 0, y.data()));
        for (index_t i = 0; i < n; ++i) w_[i] = y[i];
        judge_ = true;
    }
    // Start Newton from a coarse transport plan (no counterpart in geogram;
    // the start of optimize_levels' finer levels is its analogue): every
    // ratio-th seed of a fixed random subset carries the masses of the seeds
    // nearest to it and is solved to 1e-3; vgc_ot_convex_start extends the
    // plan to all seeds as a convex soft maximum (no empty cell).  The zero
    // start stays the fallback (judged by the mass residual).
    void set_coarse_initial_weights(index_t ratio = 256, uint64_t seed = 12345) {
        const index_t n = index_t(nu_.size());
        const index_t m = std::max<index_t>(64, n / ratio);
        if (m >= n) return;  // too few seeds: the zero start
        std::vector<index_t> idx(n);
        for (index_t i = 0; i < n; ++i) idx[i] = i;
        std::mt19937_64 rng(seed);
        for (index_t i = 0; i < m; ++i) {  // partial Fisher-Yates
            std::uniform_int_distribution<index_t> pick(i, n - 1);
            std::swap(idx[i], idx[pick(rng)]);
        }
        std::sort(idx.begin(), idx.begin() + m);
        std::vector<double> q(3 * (size_t)m), nuc(m, 0.0), wc(m, 0.0);
        for (index_t j = 0; j < m; ++j)
            for (int a = 0; a < 3; ++a) q[3 * (size_t)j + a] = points_[3 * (size_t)idx[j] + a];
        std::vector<int32_t> nn(n);
        std::vector<double> d2(n);
        ctx_.check(vgc_knn_search(ctx_.get(), q.data(), m, points_.data(), n, 1, 0, nn.data(),
                                  d2.data()));
        for (index_t i = 0; i < n; ++i) nuc[nn[i]] += nu_[i];
        vgc_ot_stats cst{};
        ctx_.check(vgc_ot_solve(ctx_.get(), q.data(), nuc.data(), wc.data(), m, box_, 1e-3, 100,
                                VGC_OT_INEXACT, &cst));
        std::vector<double> vl(m), ml(3 * (size_t)m), vv(m), mv(3 * (size_t)m);
        ctx_.check(vgc_cells(ctx_.get(), q.data(), wc.data(), m, box_, 0, VGC_NO_NEIGHBORS,
                             vl.data(), ml.data(), nullptr, nullptr, nullptr, 0, nullptr));
        ctx_.check(vgc_cells(ctx_.get(), q.data(), nullptr, m, box_, 0, VGC_NO_NEIGHBORS,
                             vv.data(), mv.data(), nullptr, nullptr, nullptr, 0, nullptr));
        const double kPi = 3.14159265358979323846;
        std::vector<double> prod(m), pieces((size_t)VGC_START_PIECE * m);
        for (index_t j = 0; j < m; ++j) {
            if (!(vl[j] > 0.0)) return;  // an empty coarse cell: the zero start
            prod[j] = 4.0 * std::cbrt(3.0 * vv[j] / (4.0 * kPi)) *
                      std::cbrt(3.0 * vl[j] / (4.0 * kPi));
        }
        std::vector<double> srt(prod);
        std::sort(srt.begin(), srt.end());
        const double s = (m % 2) ? srt[m / 2] : 0.5 * (srt[m / 2 - 1] + srt[m / 2]);
        for (index_t j = 0; j < m; ++j) {
            double c[3], b[3], dwall = 1e300, r2 = 0.0, gq = 0.0, qq = 0.0;
            for (int a = 0; a < 3; ++a) {
                c[a] = ml[3 * (size_t)j + a] / vl[j];
                b[a] = mv[3 * (size_t)j + a] / vv[j];
                dwall = std::min(dwall, std::min(c[a] - box_[a], box_[3 + a] - c[a]));
                const double qa = q[3 * (size_t)j + a];
                r2 += (qa - b[a]) * (qa - b[a]);
                gq += 2.0 * c[a] * qa;
                qq += qa * qa;
            }
            const double mu = std::cbrt(vl[j] / vv[j]);
            const double R = std::min(std::cbrt(3.0 * vv[j] / (4.0 * kPi)), 0.98 * dwall / mu);
            const double h = r2 <= R * R ? r2 : R * (2.0 * std::sqrt(r2) - R);
            double* pj = pieces.data() + (size_t)VGC_START_PIECE * j;
            pj[0] = qq - wc[j] - gq - mu * h;
            for (int a = 0; a < 3; ++a) {
                pj[1 + a] = 2.0 * c[a];
                pj[5 + a] = b[a];
            }
            pj[4] = mu;
            pj[8] = R;
        }
        ctx_.check(vgc_ot_convex_start(ctx_.get(), points_.data(), n, pieces.data(), m, s, 0,
                                       w_.data()));
        judge_ = true;
    }
    void optimize(index_t max_iter) {
        ctx_.check(vgc_ot_solve(ctx_.get(), points_.data(), nu_.data(), w_.data(),
                                index_t(nu_.size()), box_, eps_, max_iter,
                                judge_ ? VGC_OT_JUDGE_START : 0u, &stats_));
    }
    // optimize_levels (compute_OTM/main.cpp:294): the seeds set with
    // set_points are in level order, levels = {0, ..., n}; convex: coarse
    // levels carry their nearest seeds' masses and each level starts from the
    // convex extension of the previous one (VGC_OT_CONVEX_LEVELS)
    void optimize_levels(const std::vector<uint32_t>& levels, index_t max_iter,
                         bool convex = false) {
        std::vector<vgc_ot_stats> st(levels.size() > 1 ? levels.size() - 1 : 1);
        ctx_.check(vgc_ot_solve_levels(ctx_.get(), points_.data(), nu_.data(), w_.data(),
                                       index_t(nu_.size()), box_, levels.data(),
                                       index_t(levels.size() - 1), eps_, max_iter,
                                       convex ? VGC_OT_CONVEX_LEVELS : 0u, st.data()));
        stats_ = st.back();
    }
    const std::vector<double>& weights() const { return w_; }
    const vgc_ot_stats& stats() const { return stats_; }

private:
    Context& ctx_;
    double box_[6];
    double eps_ = 1e-4;
    bool judge_ = false;
    std::vector<double> points_, nu_, w_;
    vgc_ot_stats stats_{};
};

}  // namespace geogram
}  // namespace vgc

#endif

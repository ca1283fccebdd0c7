// Host driver of the cell pipeline, written once against a backend (the HIP
// backend on gfx950, or the CPU emulation used by the tests).
//
//   1. Morton keys (instance id above a 48-bit Morton code) and a radix sort;
//      seeds are gathered into Morton order (FP64 x,y,z,w and an FP32 copy).
//   2. An implicit bounding-volume tree per instance over buckets of 16
//      consecutive seeds.
//   3. Pass A: every cell, one wavefront each, 64 corner slots.  The wave
//      finds its neighbours itself (breadth-first walk of the tree into LDS,
//      one ring of distance at a time, sorted, nearest first) until the
//      security radius is reached; this replaces the kd-tree neighbourhoods
//      of Delaunay_NearestNeighbors (delaunay/delaunay_nn.cpp) and their
//      k += k/8 enlargements (voronoi/generic_RVD.h:2335-2345).  With k0 > 0
//      the first pass reads kNN lists of k0 instead (FP32 ranking).
//   4. Pass B: cells whose list ran out before the security radius (list
//      mode), or whose ring did not fit the wave's buffers.  Their
//      partial cell bounds the final one, so every seed within
//      sqrt(4.1) R_partial is gathered (radius query), sorted by distance and
//      the cell is clipped again: the list is complete by construction.
//   5. Pass C: cells that overflowed 64 corners, with 128 corner slots, then
//      256 for the few that overflow again (huge Laguerre cells).
//   6. Pass E: cells with a sign inside its error bound, again with exact
//      signs (ties resolved as geogram's exact predicates do).
//   7. Neighbour lists and facet areas compacted into CSR (caller's order).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace vgc {

struct EngineStats {
    unsigned long long nb_cells = 0, nb_neighbors = 0, nb_pass_b = 0, nb_pass_c = 0, nb_pass_e = 0;
    unsigned long long nb_uncertain = 0, nb_failed = 0, nb_empty = 0, nb_disconnected = 0;
    unsigned long long nb_clips = 0, nb_tests = 0, radius_entries = 0;
    double t_build_ms = 0, t_knn_ms = 0, t_clip_ms = 0, t_pass_bc_ms = 0, t_compact_ms = 0,
           t_total_ms = 0;
    // pass A clip launch: GPU time (events on the launch stream) and the
    // algorithmic work it did (live corners per test, corners created,
    // integration tetrahedra, facets)
    double t_clip_a_gpu_ms = 0, t_knn_gpu_ms = 0;
    unsigned long long a_live = 0, a_new = 0, a_tets = 0, a_facets = 0, a_tests = 0, a_cells = 0;
    unsigned long long nb_fail_capacity = 0;  // failed cells over the corner/plane/facet limits
    unsigned long long nb_pieces = 0;  // cells over the plane limit computed in octant pieces
    unsigned long long why_rings = 0, why_walk = 0, why_narrow = 0;  // pass A -> lists
    unsigned long long nb_regather = 0;  // weighted cells gathered again without the ring limit
    // cells of the FP32 first pass it could not certify, redone by the FP64
    // first pass, and that launch's GPU time
    unsigned long long nb_fp64_redo = 0;
    double t_redo_gpu_ms = 0;
    unsigned long long nb_periodic_retries = 0;  // periodic passes redone with a wider margin
    int tuned_waves = 0;                 // autotuner's choice (0: not tuned by this call)
};

struct Problem {
    int n = 0;                      // total seeds (all instances)
    int n_inst = 1;
    std::vector<int> inst_n;        // seeds per instance (caller order = instance order)
    std::vector<double> box;        // 6 per instance
    const double* d_xyz = nullptr;  // device (backend) memory, 3n, caller order
    const double* d_w = nullptr;    // device, n, or nullptr (Voronoi)
    int k0 = 0;                     // 0: neighbours gathered inside the clipping kernel;
                                    // 16..64: kNN lists of k0 for the first pass
    bool want_neighbors = true;
    bool want_e2 = false;           // integral of |x - p|^2 per cell (CVT energy)
    bool count_work = false;        // algorithmic work counters (global atomics)
    bool abort_on_empty = false;    // an instance with an empty cell stops (and is bad)
                                    // (cells() then returns 1; outputs unset)
    // per instance, nonzero = compute its cells; the others keep the outputs
    // (volumes, centroids, neighbours) of the last pass that computed them
    const std::vector<int>* active = nullptr;
    // periodic domain (one instance): the box is the period; cells are
    // computed with images of the seeds near the faces as extra candidates
    bool periodic = false;
    // internal: the first n_real seeds are the period's, src maps every
    // seed (images included) to its seed of the period
    int n_real = 0;
    const int* src = nullptr;
    // cell meshes / primal tetrahedra (ClipParams::mesh_*), one instance
    int mesh_mode = 0;
    double mesh_shrink = 0.0;
    int* mesh_cnt = nullptr;
    const long long* mesh_off = nullptr;
    double* mesh_v = nullptr;
    long long* mesh_fptr = nullptr;
    int* mesh_fv = nullptr;
    int* mesh_fnbr = nullptr;
    int* mesh_tet = nullptr;
};

// Outputs of Engine::mesh (device memory; null pointers are not written).
struct MeshOut {
    double* v = nullptr;        // 3 per vertex
    long long* fptr = nullptr;  // facets + 1 (the last entry = facet vertex entries)
    int* fv = nullptr;
    int* fnbr = nullptr;
    int* tet = nullptr;         // 4 per tetrahedron
    long long* cell_off = nullptr;  // 4 (n + 1): per cell first vertex, facet,
                                    // facet entry, tetrahedron (and the totals)
};

template <class B>
class Engine {
public:
    explicit Engine(B& be) : be_(be) {}
    ~Engine() {
        for (auto& s : slots_)
            if (s.p) be_.dfree(s.p);
    }

    std::string error;

    // outputs of the last cells() call (backend memory)
    double* d_vol = nullptr;
    double* d_mg = nullptr;
    double* d_e2 = nullptr;  // if want_e2
    long long* d_ptr = nullptr;  // n + 1
    int* d_nbr = nullptr;
    double* d_area = nullptr;
    long long total_neighbors = 0;
    long long list_chunk = 1ll << 27;  // entries per chunk of the list tiers
    bool autotune = false;             // pick the first pass's work-group size by timing
    bool waves4 = false;               // first pass with 4-wavefront work-groups (no tuner)
    bool fp64_first = false;           // first pass with FP64 corner state (default FP32)
    // start of a library call: the FP32 / FP64 choice of the first pass
    // follows the passes of this call only
    void begin_call() { fp32_skip_ = 0; }
    std::vector<int> inst_bad;         // per instance: failed or empty cells of the last pass
    std::vector<uint32_t> h_flags;  // by Morton position, current for census_ cells
    std::vector<int> census_;       // cells of the last pass that are not plainly done

    // layout of the last build
    InstLayout IL{};
    pt4* d_pts = nullptr;
    pf4* d_pf = nullptr;
    int* d_orig = nullptr;
    int* d_inst_of = nullptr;
    bnode* d_nodes = nullptr;

    static constexpr int kFaceStride = 256;  // faces of a cell <= 256 planes

    // ------------------------------------------------------------ build
    int build(const Problem& pb, EngineStats& st) {
        auto t0 = now();
        n_ = pb.n;
        ni_ = pb.n_inst;
        if ((int)pb.inst_n.size() != ni_ || (int)pb.box.size() != 6 * ni_) {
            error = "instance description does not match n_inst";
            return -1;
        }
        h_p0_.assign(ni_, 0);
        h_n_ = pb.inst_n;
        h_b0_.assign(ni_, 0);
        h_nb_.assign(ni_, 0);
        h_L_.assign(ni_, 1);
        h_t0_.assign(ni_, 0);
        h_leaf_off_.assign(ni_ + 1, 0);
        long long pp = 0, bb = 0, tt = 0;
        max_h_ = 0;
        for (int i = 0; i < ni_; ++i) {
            h_p0_[i] = (int)pp;
            pp += h_n_[i];
            int nb = (h_n_[i] + kBucket - 1) / kBucket;
            h_nb_[i] = nb;
            h_b0_[i] = (int)bb;
            bb += nb;
            int L = 1, h = 0;
            while (L < nb) {
                L *= 2;
                ++h;
            }
            h_L_[i] = L;
            max_h_ = std::max(max_h_, h);
            h_t0_[i] = tt;
            tt += 2LL * L;
            h_leaf_off_[i + 1] = h_leaf_off_[i] + L;
        }
        if (pp != n_) {
            error = "instance sizes do not add up to n";
            return -1;
        }
        n_nodes_ = tt;
        int* d_p0 = (int*)slot(S_P0, sizeof(int) * ni_);
        int* d_n = (int*)slot(S_N, sizeof(int) * ni_);
        int* d_b0 = (int*)slot(S_B0, sizeof(int) * ni_);
        int* d_nb = (int*)slot(S_NB, sizeof(int) * ni_);
        int* d_L = (int*)slot(S_L, sizeof(int) * ni_);
        long long* d_t0 = (long long*)slot(S_T0, sizeof(long long) * ni_);
        long long* d_leaf_off = (long long*)slot(S_LEAFOFF, sizeof(long long) * (ni_ + 1));
        double* d_box = (double*)slot(S_BOX, sizeof(double) * 6 * ni_);
        be_.h2d(d_p0, h_p0_.data(), sizeof(int) * ni_);
        be_.h2d(d_n, h_n_.data(), sizeof(int) * ni_);
        be_.h2d(d_b0, h_b0_.data(), sizeof(int) * ni_);
        be_.h2d(d_nb, h_nb_.data(), sizeof(int) * ni_);
        be_.h2d(d_L, h_L_.data(), sizeof(int) * ni_);
        be_.set_tree_leaves(*std::max_element(h_L_.begin(), h_L_.begin() + ni_));
        be_.h2d(d_t0, h_t0_.data(), sizeof(long long) * ni_);
        be_.h2d(d_leaf_off, h_leaf_off_.data(), sizeof(long long) * (ni_ + 1));
        be_.h2d(d_box, pb.box.data(), sizeof(double) * 6 * ni_);
        d_box_ = d_box;

        // Morton sort
        int* d_inst_orig = nullptr;
        if (ni_ > 1) {
            d_inst_orig = (int*)slot(S_INSTORIG, sizeof(int) * (size_t)n_);
            be_.fill_inst(d_inst_orig, n_, d_p0, ni_);
        }
        uint64_t* keys = (uint64_t*)slot(S_KEYS, sizeof(uint64_t) * (size_t)n_);
        uint64_t* keys2 = (uint64_t*)slot(S_KEYS2, sizeof(uint64_t) * (size_t)n_);
        int* vals = (int*)slot(S_VALS, sizeof(int) * (size_t)n_);
        int* vals2 = (int*)slot(S_VALS2, sizeof(int) * (size_t)n_);
        be_.morton(pb.d_xyz, d_inst_orig, d_box, n_, keys, vals);
        int end_bit = 48;
        while ((1LL << (end_bit - 48)) < ni_) ++end_bit;
        be_.sort_pairs(keys, vals, keys2, vals2, n_, end_bit);  // result in keys2 / vals2
        d_pts = (pt4*)slot(S_PTS, sizeof(pt4) * (size_t)n_);
        d_pf = (pf4*)slot(S_PF, sizeof(pf4) * (size_t)n_);
        d_orig = (int*)slot(S_ORIG, sizeof(int) * (size_t)n_);
        d_inst_of = ni_ > 1 ? (int*)slot(S_INSTOF, sizeof(int) * (size_t)n_) : nullptr;
        be_.gather(pb.d_xyz, pb.d_w, d_inst_orig, vals2, n_, d_pts, d_pf, d_orig, d_inst_of);

        IL.inst_of = d_inst_of;
        IL.p0 = d_p0;
        IL.n = d_n;
        IL.b0 = d_b0;
        IL.nb = d_nb;
        IL.t0 = d_t0;
        IL.L = d_L;
        d_nodes = (bnode*)slot(S_NODES, sizeof(bnode) * (size_t)n_nodes_);
        be_.bvh_leaves(d_pts, IL, d_nodes, d_leaf_off, ni_, h_leaf_off_[ni_]);
        std::vector<long long> off(ni_ + 1);
        long long* d_lvl = (long long*)slot(S_LVLOFF, sizeof(long long) * (ni_ + 1));
        for (int h = 1; h <= max_h_; ++h) {
            off[0] = 0;
            for (int i = 0; i < ni_; ++i) off[i + 1] = off[i] + (h_L_[i] >> h);
            if (off[ni_] == 0) break;
            be_.h2d(d_lvl, off.data(), sizeof(long long) * (ni_ + 1));
            be_.bvh_level(IL, d_nodes, d_lvl, ni_, off[ni_], h);
        }
        // weights: max per instance (security radius of Laguerre cells)
        weighted_ = pb.d_w != nullptr;
        d_wmax_ = (double*)slot(S_WMAX, sizeof(double) * ni_);
        if (weighted_) be_.wmax(d_pts, IL, ni_, d_wmax_);
        be_.sync();
        st.t_build_ms += ms_since(t0);
        return 0;
    }

    // Weights of the power-nearest seeds of pb (one instance, weighted) for
    // the nq points d_q (FP64 xyz): the start of the new seeds of a finer
    // level in the multilevel transport solver.
    int power_nearest_weights(const Problem& pb, const double* d_q, int nq, double* d_wout,
                              EngineStats& st) {
        if (build(pb, st)) return -1;
        double cs = 0.0;
        for (int k = 0; k < 6; ++k) cs = std::max(cs, std::fabs(pb.box[k]));
        be_.power_nearest(d_pts, d_nodes, IL, d_q, nq, float(cs * cs), d_wout);
        return 0;
    }

    // seeds of the last build
    int total_points() const { return n_; }

    // Weights of the empty cells of the last cell pass (built for pb.d_w)
    // raised until each seed lies in its cell (ot_repair_row), into d_wout.
    void repair_empty(const Problem& pb, double delta, double* d_wout) {
        double cs = 0.0;
        for (int k = 0; k < 6; ++k) cs = std::max(cs, std::fabs(pb.box[k]));
        be_.ot_repair(pb.n, pb.d_xyz, pb.d_w, d_vol, d_pts, d_nodes, IL, float(cs * cs), delta,
                      d_wout);
    }

    // ------------------------------------------------------------ kNN
    // K nearest of every seed (Morton positions), n x K.
    int knn_lists(int K, int*& d_list, float* d_d2 = nullptr) {
        d_list = (int*)slot(S_LIST, sizeof(int) * (size_t)n_ * K);
        be_.knn(K, d_pf, d_nodes, IL, nullptr, n_, d_list, d_d2);
        return 0;
    }

    // ------------------------------------------------------- exact kNN
    // Exact nearest neighbours among the points of the last build (one
    // instance) of nq query points: per query, the kk points of smallest
    // (FP64 squared distance, caller's index) — FP32 radius lists (see
    // common.h), grown until the kk-th entry is certified.  d_q: FP64 query
    // points in the caller's order (device); nullptr: the built points
    // themselves.  S: largest coordinate magnitude of points and queries.
    // Output rows (query order, kout = kk, or kk - 1 with self_rule) of caller
    // indices / FP64 squared distances, -1 / -1.0 where there is none.
    // self_rule: each point's own entry is dropped and coincident points
    // follow Delaunay_NearestNeighbors::get_neighbors_internal
    // (delaunay/delaunay_nn.cpp:105-143): a coincident point of smaller index
    // leaves the row empty, one of larger index is skipped.
    int knn_exact(const double* d_q, int nq, int kk, bool self_rule, double S, int* d_idx,
                  double* d_d2, EngineStats& st) {
        auto t0 = now();
        if (ni_ != 1 || kk < 1 || kk > n_) {
            error = "knn: one instance and 1 <= k <= number of points";
            return -1;
        }
        const pt4* qp = d_pts;
        const pf4* qf = d_pf;
        const int* qrow = d_orig;
        const int* qpos = nullptr;  // Morton place of each query (built points: itself)
        if (d_q) {
            uint64_t* k1 = (uint64_t*)slot(S_QKEYS, sizeof(uint64_t) * (size_t)nq);
            uint64_t* k2 = (uint64_t*)slot(S_QKEYS2, sizeof(uint64_t) * (size_t)nq);
            int* v1 = (int*)slot(S_QVALS, sizeof(int) * (size_t)nq);
            int* v2 = (int*)slot(S_QVALS2, sizeof(int) * (size_t)nq);
            be_.morton(d_q, nullptr, d_box_, nq, k1, v1);
            be_.sort_pairs(k1, v1, k2, v2, nq, 48);
            pt4* qp_ = (pt4*)slot(S_QPTS, sizeof(pt4) * (size_t)nq);
            pf4* qf_ = (pf4*)slot(S_QPF, sizeof(pf4) * (size_t)nq);
            int* qrow_ = (int*)slot(S_QROW, sizeof(int) * (size_t)nq);
            int* qpos_ = (int*)slot(S_QPOS, sizeof(int) * (size_t)nq);
            be_.gather(d_q, nullptr, nullptr, v2, nq, qp_, qf_, qrow_, nullptr);
            be_.lower_bound_u64((const uint64_t*)slot(S_KEYS2, 0), n_, k2, nq, qpos_);
            qp = qp_;
            qf = qf_;
            qrow = qrow_;
            qpos = qpos_;
        }
        const double e = knn_vec_err(S);
        float* rr = (float*)slot(S_QR2, sizeof(float) * (size_t)nq);
        be_.knn_seed_radius(d_pf, d_nodes, IL, qf, qpos, nq, kk, e, rr);
        std::vector<int> items;
        bool all = true;
        int* d_redo = (int*)slot(S_QREDO, sizeof(int) * (size_t)nq);
        unsigned* d_nredo = (unsigned*)slot(S_CENS_N, sizeof(unsigned));
        for (int round = 0;; ++round) {
            const int ni = all ? nq : (int)items.size();
            if (ni == 0) break;
            if (round >= 96) {
                error = "knn: radius did not converge";
                return -1;
            }
            int* d_items = (int*)slot(S_ITEMS, sizeof(int) * (size_t)ni);
            if (all)
                be_.iota(d_items, ni);
            else
                be_.h2d(d_items, items.data(), sizeof(int) * (size_t)ni);
            int* d_cnt = (int*)slot(S_RCNT, sizeof(int) * (size_t)ni);
            be_.knn_count(d_pf, d_nodes, IL, qf, d_items, ni, rr, d_cnt);
            long long* d_off = (long long*)slot(S_ROFF, sizeof(long long) * ((size_t)ni + 1));
            be_.scan_i32(d_cnt, d_off, ni);
            std::vector<long long> off(ni + 1);
            be_.d2h(off.data(), d_off, sizeof(long long) * ((size_t)ni + 1));
            st.radius_entries += off[ni];
            be_.dzero(d_nredo, sizeof(unsigned));
            for (int a = 0; a < ni;) {
                int b = a + 1;
                while (b < ni && off[b + 1] - off[a] <= list_chunk) ++b;
                const long long base = off[a], total = off[b] - base;
                const int nc = b - a;
                const size_t cap = (size_t)std::max<long long>(total, 1);
                int* l_m = (int*)slot(S_RLIST, sizeof(int) * cap);
                float* l_f = (float*)slot(S_RD2, sizeof(float) * cap);
                int* l_i = (int*)slot(S_RLIST2, sizeof(int) * cap);
                int* l_i2 = (int*)slot(S_QLI2, sizeof(int) * cap);
                double* l_d = (double*)slot(S_QLD, sizeof(double) * cap);
                double* l_d2 = (double*)slot(S_QLD2, sizeof(double) * cap);
                long long* d_coff = (long long*)slot(S_COFF, sizeof(long long) * ((size_t)nc + 1));
                be_.rebase(d_off + a, nc + 1, base, d_coff);
                be_.knn_fill(d_pf, d_nodes, IL, qf, d_items + a, nc, rr, d_coff, l_m, l_f);
                be_.knn_d64(qp, d_pts, d_orig, d_items + a, nc, d_coff, l_m, l_i, l_d);
                // (FP64 squared distance, caller's index): index first, then a
                // stable sort by distance
                be_.seg_sort_kv(l_i, l_d, l_i2, l_d2, total, nc, d_coff);
                be_.seg_sort_kv(l_d2, l_i2, l_d, l_i, total, nc, d_coff);
                be_.knn_select(d_items + a, nc, d_coff, d_cnt + a, l_d, l_i, rr, e, kk, n_,
                               self_rule ? 1 : 0, qrow, d_idx, d_d2, d_redo, d_nredo);
                a = b;
            }
            unsigned nr = 0;
            be_.d2h(&nr, d_nredo, sizeof(unsigned));
            items.resize(nr);
            if (nr) be_.d2h(items.data(), d_redo, sizeof(int) * nr);
            std::sort(items.begin(), items.end());
            all = false;
        }
        be_.sync();
        st.t_knn_ms += ms_since(t0);
        return 0;
    }

    // ------------------------------------------------------------ cells
    int cells(const Problem& pb, EngineStats& st) {
        if (pb.periodic) return cells_periodic(pb, st);
        return cells_core(pb, st);
    }

    // Periodic cells (the torus of the box; PeriodicDelaunay3d with
    // periodic = true, delaunay/periodic_delaunay_3d.cpp:4017-4320, uses 27
    // translated copies): images of the seeds within a margin delta of the
    // faces join as candidates, the cells are clipped in the box grown by
    // delta, and a cell counts only if everything that could still cut it
    // (the radius of its security test) lies within the margin; otherwise
    // the pass is redone with 1.1 times the largest such radius.  The first
    // margin, 8 mean spacings, covers the security radius of Voronoi cells
    // of spread-out seeds (about 2 cell radii), so one pass is the common
    // case.  Outputs as cells(): the first n entries; neighbour
    // ids are seeds of the period (an image's seed).
    int cells_periodic(const Problem& pb, EngineStats& st) {
        if (pb.n_inst != 1 || pb.mesh_mode) {
            error = "periodic: one instance, no mesh export";
            return -1;
        }
        const long long n = pb.n;
        const double* b = pb.box.data();
        const double Lx = b[3] - b[0], Ly = b[4] - b[1], Lz = b[5] - b[2];
        const double Lmax = std::max(Lx, std::max(Ly, Lz));
        double delta = std::min(Lmax, 8.0 * std::cbrt(Lx * Ly * Lz / double(n)));
        for (int round = 0;; ++round) {
            // delta = Lmax keeps all 26 images of every seed.  That set is
            // enough for any cell: a cell lies within half a period of its
            // seed on each axis (its own images bound it), and for a point
            // v there the nearest image of any seed lies within half a
            // period of v on each axis, i.e. among the images one period
            // away; farther images cut nothing the nearest one does not.
            const bool full = delta >= Lmax;
            int* gcnt = (int*)slot(S_GCNT, sizeof(int) * (size_t)n);
            long long* goff = (long long*)slot(S_GOFF, sizeof(long long) * ((size_t)n + 1));
            double* d_b = (double*)slot(S_GBOX, sizeof(double) * 6);
            be_.h2d(d_b, b, sizeof(double) * 6);
            be_.ghost_count(pb.d_xyz, (int)n, d_b, delta, gcnt);
            be_.scan_i32(gcnt, goff, (int)n);
            long long ng = 0;
            be_.d2h(&ng, goff + n, sizeof(long long));
            const long long ne = n + ng;
            if (ne > (1ll << 30)) {
                error = "periodic: too many images";
                return -1;
            }
            double* xe = (double*)slot(S_GXYZ, sizeof(double) * 3 * (size_t)ne);
            double* we = pb.d_w ? (double*)slot(S_GW, sizeof(double) * (size_t)ne) : nullptr;
            int* src = (int*)slot(S_GSRC, sizeof(int) * (size_t)ne);
            be_.ghost_fill(pb.d_xyz, pb.d_w, (int)n, d_b, delta, goff, xe, we, src);
            Problem q = pb;
            q.periodic = false;
            q.n = (int)ne;
            q.inst_n = {(int)ne};
            q.box = {b[0] - delta, b[1] - delta, b[2] - delta, b[3] + delta, b[4] + delta,
                     b[5] + delta};
            q.d_xyz = xe;
            q.d_w = we;
            q.n_real = (int)n;
            q.src = src;
            const int rc = cells_core(q, st);
            if (rc) return rc;  // error, or a transport trial stopped by an empty cell
            // the reach of the cells: the radius within which a seed could
            // still cut one (its security radius)
            float* d_reach = (float*)slot(S_GREACH, sizeof(float));
            be_.periodic_reach(d_pts, d_orig, (int)ne, (int)n, d_r2_, weighted_ ? d_wmax_ : nullptr,
                               d_reach);
            float reach2 = 0.0f;
            be_.d2h(&reach2, d_reach, sizeof(float));
            if (full || double(reach2) < delta * delta) break;
            // reach2 >= delta^2 here, so the margin grows
            delta = std::min(Lmax, 1.1 * std::sqrt(double(reach2)));
            st.nb_periodic_retries++;
        }
        return 0;
    }

    int cells_core(const Problem& pb, EngineStats& st) {
        auto tstart = now();
        if (build(pb, st)) return -1;
        piece_cells_.clear();
        x_total_ = 0;
        const int K0 = pb.k0;
        const bool gather = K0 == 0;
        auto t1 = now();
        int* d_list = nullptr;
        if (!gather) {
            be_.event_begin(0);
            knn_lists(K0, d_list);
            be_.event_end(0);
            be_.sync();
            st.t_knn_ms += ms_since(t1);
            st.t_knn_gpu_ms += be_.event_ms(0);
        }

        d_vol = (double*)slot(S_VOL, sizeof(double) * (size_t)n_);
        d_mg = (double*)slot(S_MG, sizeof(double) * 3 * (size_t)n_);
        int* d_cnt = (int*)slot(S_CNT, sizeof(int) * (size_t)n_);
        int* d_nbr_tmp = (int*)slot(S_NBRTMP, sizeof(int) * (size_t)n_ * kFaceStride);
        double* d_area_tmp = (double*)slot(S_AREATMP, sizeof(double) * (size_t)n_ * kFaceStride);
        uint32_t* d_flags = (uint32_t*)slot(S_FLAGS, sizeof(uint32_t) * (size_t)n_);
        double* d_r2 = (double*)slot(S_R2, sizeof(double) * (size_t)n_);
        unsigned long long* d_ctr = (unsigned long long*)slot(S_CTR, sizeof(unsigned long long) * 8);
        be_.dzero(d_ctr, sizeof(unsigned long long) * 8);
        be_.dzero(d_flags, sizeof(uint32_t) * (size_t)n_);
        // a subset of instances: the others keep their neighbour counts
        std::vector<int> aitems;
        if (pb.active) {
            for (int k = 0; k < ni_; ++k)
                if ((*pb.active)[k])
                    for (int c = h_p0_[k]; c < h_p0_[k] + h_n_[k]; ++c) aitems.push_back(c);
        } else {
            be_.dzero(d_cnt, sizeof(int) * (size_t)n_);
        }

        ClipParams P{};
        P.pts = d_pts;
        P.orig = d_orig;
        P.inst_of = d_inst_of;
        P.box = d_box_;
        P.wmax = d_wmax_;
        P.inst_n = IL.n;
        P.weighted = weighted_ ? 1 : 0;
        P.vol = d_vol;
        P.mg = d_mg;
        d_e2 = pb.want_e2 ? (double*)slot(S_E2, sizeof(double) * (size_t)n_) : nullptr;
        P.e2 = d_e2;
        P.nbr_cnt = d_cnt;
        P.nbr_tmp = d_nbr_tmp;
        P.area_tmp = d_area_tmp;
        P.face_stride = kFaceStride;
        P.flags = d_flags;
        P.r2 = d_r2;
        P.ball = (fball*)slot(S_BALL, sizeof(fball) * (size_t)n_);
        P.counters = pb.count_work ? d_ctr : nullptr;
        P.pf = d_pf;
        P.nodes = d_nodes;
        P.IL = IL;
        P.mesh_mode = pb.mesh_mode;
        P.mesh_n = n_;
        P.mesh_cnt = pb.mesh_cnt;
        P.mesh_off = pb.mesh_off;
        P.mesh_shrink = pb.mesh_shrink;
        P.mesh_v = pb.mesh_v;
        P.mesh_fptr = pb.mesh_fptr;
        P.mesh_fv = pb.mesh_fv;
        P.mesh_fnbr = pb.mesh_fnbr;
        P.mesh_tet = pb.mesh_tet;
        P.n_real = pb.n_real;
        P.src = pb.src;
        d_r2_ = d_r2;

        // pass A over every cell, then the tiers for the flagged ones
        auto t2 = now();
        P.items = nullptr;
        P.n_items = n_;
        if (pb.active) {
            int* d_ai = (int*)slot(S_AITEMS, sizeof(int) * std::max<size_t>(1, aitems.size()));
            if (!aitems.empty()) be_.h2d(d_ai, aitems.data(), sizeof(int) * aitems.size());
            P.items = d_ai;
            P.n_items = (int)aitems.size();
        }
        P.list = d_list;
        P.list_ptr = nullptr;
        P.list_len = nullptr;
        P.list_stride = K0;
        P.list_complete = 0;
        P.ring_limit = weighted_ ? kMaxFilteredRings : kMaxVoronoiRings;
        P.corner = 0;
        // FP32 corner state for the first pass, except for cell meshes (those
        // are emitted by the FP64 pass) and for a few passes after one that
        // had to redo more than 5 % of its cells in FP64 (large Laguerre
        // weights: there the FP64 pass alone is faster; the outputs are the
        // same either way)
        P.fp32 = (gather && !fp64_first && !pb.mesh_mode && fp32_skip_ == 0) ? 1 : 0;
        if (fp32_skip_ > 0) --fp32_skip_;
        if (pb.abort_on_empty) {
            // the caller only needs to know whether an instance has an empty
            // cell: its first empty cell stops it
            unsigned* d_ab = (unsigned*)slot(S_ABORT, sizeof(unsigned) * (size_t)ni_);
            be_.dzero(d_ab, sizeof(unsigned) * (size_t)ni_);
            P.abort_flag = d_ab;
        }
        if (!autotune) {
            // without the tuner every call runs the default work-group size
            be_.set_clip_waves(waves4 ? 4 : kDefaultClipWaves);
            tuned_n_ = -1;
        }
        if (autotune && P.n_items > 0 && tuned_n_ != n_) {
            // autotuner: time the first pass with 1 and 4 wavefronts per
            // work-group (same results) and keep the faster for this size
            // (interleaved, twice each, minimum kept: the first launch of a
            // problem runs on cold caches).  Large problems are timed on a
            // sample of 128 evenly spaced runs of 2048 consecutive cells, so
            // that tuning costs about half of one full pass.
            ClipParams T = P;
            T.counters = nullptr;
            T.abort_flag = nullptr;
            const int kRun = 2048, kRuns = 128;
            if (P.n_items > 2 * kRun * kRuns) {
                std::vector<int> titems;
                titems.reserve((size_t)kRun * kRuns);
                for (int j = 0; j < kRuns; ++j) {
                    const long long b = (long long)P.n_items * j / kRuns;
                    for (int i = 0; i < kRun; ++i)
                        titems.push_back(pb.active ? aitems[b + i] : (int)(b + i));
                }
                int* d_t = (int*)slot(S_TITEMS, sizeof(int) * titems.size());
                be_.h2d(d_t, titems.data(), sizeof(int) * titems.size());
                T.items = d_t;
                T.n_items = (int)titems.size();
            }
            double t1w = 1e300, t4w = 1e300;
            for (int wv : {4, 1, 4, 1}) {
                be_.set_clip_waves(wv);
                be_.event_begin(1);
                be_.clip(1, false, gather, T);
                be_.event_end(1);
                be_.sync();
                const double ms = be_.event_ms(1);
                double& t = wv == 1 ? t1w : t4w;
                t = ms < t ? ms : t;
            }
            // the default unless the other is clearly faster
            const int best_w = kDefaultClipWaves == 1 ? (t4w < 0.97 * t1w ? 4 : 1)
                                                      : (t1w < 0.97 * t4w ? 1 : 4);
            be_.set_clip_waves(best_w);
            tuned_n_ = n_;
            st.tuned_waves = best_w;
        }
        be_.event_begin(1);
        if (P.n_items > 0) be_.clip(1, false, gather, P);
        be_.event_end(1);
        be_.sync();
        st.t_clip_ms += ms_since(t2);
        st.t_clip_a_gpu_ms += be_.event_ms(1);
        {
            unsigned long long c[8];
            be_.d2h(c, d_ctr, sizeof(c));
            st.a_tests += c[1];
            st.a_live += c[2];
            st.a_new += c[3];
            st.a_tets += c[4];
            st.a_facets += c[5];
            st.a_cells += P.n_items;
        }
        if (P.fp32) {
            // cells with a sign or corner the FP32 corners could not certify
            // (flagged uncertain and left without outputs): the FP64 first
            // pass, which writes every output of them
            P.fp32 = 0;
            const size_t cap = std::max(1, P.n_items);
            int* d_cc = (int*)slot(S_CENS_C, sizeof(int) * cap);
            uint32_t* d_cf = (uint32_t*)slot(S_CENS_F, sizeof(uint32_t) * cap);
            unsigned* d_cn = (unsigned*)slot(S_CENS_N, sizeof(unsigned));
            const unsigned m = be_.census(P.items, P.n_items, d_flags, CELL_UNCERTAIN, P.abort_flag,
                                          P.inst_of, d_cc, d_cf, d_cn);
            std::vector<int> cc(m), redo;
            std::vector<uint32_t> cf(m);
            if (m) {
                be_.d2h(cc.data(), d_cc, sizeof(int) * m);
                be_.d2h(cf.data(), d_cf, sizeof(uint32_t) * m);
            }
            for (unsigned i = 0; i < m; ++i)
                if (cf[i] & CELL_UNCERTAIN) redo.push_back(cc[i]);
            std::sort(redo.begin(), redo.end());
            st.nb_fp64_redo += redo.size();
            if (redo.size() * 20 > (size_t)P.n_items) fp32_skip_ = kFp32Skip;
            if (!redo.empty()) {
                int* d_items = (int*)slot(S_ITEMS, sizeof(int) * redo.size());
                be_.h2d(d_items, redo.data(), sizeof(int) * redo.size());
                ClipParams Q = P;
                Q.items = d_items;
                Q.n_items = (int)redo.size();
                be_.event_begin(1);
                be_.clip(1, false, gather, Q);
                be_.event_end(1);
                be_.sync();
                st.t_redo_gpu_ms += be_.event_ms(1);
            }
        }
        auto t3 = now();
        // instances stopped by an empty cell: bad; no further pass for them
        std::vector<unsigned> stopped;
        if (P.abort_flag) {
            stopped.resize(ni_);
            be_.d2h(stopped.data(), P.abort_flag, sizeof(unsigned) * (size_t)ni_);
            int nstop = 0, nlive = 0;
            for (int k = 0; k < ni_; ++k) {
                nstop += stopped[k] != 0;
                nlive += !pb.active || (*pb.active)[k];
            }
            st.nb_empty += nstop;
            if (nstop == nlive) {
                inst_bad.assign(ni_, 0);
                for (int k = 0; k < ni_; ++k) inst_bad[k] = stopped[k] ? 1 : 0;
                st.t_total_ms += ms_since(tstart);
                return 1;
            }
        }
        // the cells that are not plainly done, with their flags, sorted by
        // Morton position: the later passes only touch these, and h_flags is
        // kept current for them alone
        h_flags.resize(n_);
        census_.clear();
        {
            const size_t cap = std::max(1, P.n_items);
            int* d_cc = (int*)slot(S_CENS_C, sizeof(int) * cap);
            uint32_t* d_cf = (uint32_t*)slot(S_CENS_F, sizeof(uint32_t) * cap);
            unsigned* d_cn = (unsigned*)slot(S_CENS_N, sizeof(unsigned));
            const unsigned m = be_.census(P.items, P.n_items, d_flags, kCensusMask, P.abort_flag,
                                          P.inst_of, d_cc, d_cf, d_cn);
            std::vector<int> cc(m);
            std::vector<uint32_t> cf(m);
            if (m) {
                be_.d2h(cc.data(), d_cc, sizeof(int) * m);
                be_.d2h(cf.data(), d_cf, sizeof(uint32_t) * m);
            }
            std::vector<unsigned> ord(m);
            for (unsigned i = 0; i < m; ++i) ord[i] = i;
            std::sort(ord.begin(), ord.end(), [&](unsigned a, unsigned b) { return cc[a] < cc[b]; });
            census_.resize(m);
            for (unsigned i = 0; i < m; ++i) {
                census_[i] = cc[ord[i]];
                h_flags[cc[ord[i]]] = cf[ord[i]];
            }
        }
        std::vector<int> need, over;
        long long nempty = 0;
        for (int c : census_) {
            uint32_t f = h_flags[c];
            if (f & CELL_OVERFLOW)
                over.push_back(c);
            else if (f & CELL_NEED_K)
                need.push_back(c);
            if (f & CELL_EMPTY) ++nempty;
            if (f & CELL_NEED_K) {
                st.why_rings += (f & CELL_WHY_RINGS) != 0;
                st.why_walk += (f & CELL_WHY_WALK) != 0;
                st.why_narrow += (f & CELL_WHY_NARROW) != 0;
            }
        }
        if (pb.abort_on_empty && ni_ == 1 && nempty > 0) {
            inst_bad.assign(1, 1);
            st.nb_empty += nempty;
            st.t_total_ms += ms_since(tstart);
            return 1;
        }
        P.abort_flag = nullptr;
        if (gather) {
            // cells that used up the first pass's ring limit: gather again
            // without the limit (a wave per cell) before any list tier
            std::vector<int> again, rest;
            for (int c : need) (h_flags[c] & CELL_WHY_RINGS ? again : rest).push_back(c);
            if (!again.empty()) {
                st.nb_regather += again.size();
                int* d_items = (int*)slot(S_ITEMS, sizeof(int) * again.size());
                be_.h2d(d_items, again.data(), sizeof(int) * again.size());
                ClipParams Q = P;
                Q.items = d_items;
                Q.n_items = (int)again.size();
                Q.ring_limit = 1 << 30;
                Q.corner = 1;
                be_.clip(1, false, true, Q);
                refresh_flags(again, d_flags);
                for (int c : again) {
                    if (h_flags[c] & CELL_OVERFLOW)
                        over.push_back(c);
                    else if (h_flags[c] & CELL_NEED_K)
                        rest.push_back(c);
                }
                need.swap(rest);
            }
        }
        if (tiers(need, over, false, P, st)) return -1;
        // cells with signs inside their error bound: exact re-clip
        std::vector<int> unsure;
        for (int c : census_)
            if (h_flags[c] & (CELL_UNCERTAIN | CELL_BAD_TOPO)) unsure.push_back(c);
        st.nb_pass_e += unsure.size();
        if (!unsure.empty()) {
            int ni = (int)unsure.size();
            int* d_items = (int*)slot(S_ITEMS, sizeof(int) * ni);
            be_.h2d(d_items, unsure.data(), sizeof(int) * ni);
            ClipParams Q = P;
            Q.items = d_items;
            Q.n_items = ni;
            be_.clip(1, true, gather, Q);
            refresh_flags(unsure, d_flags);
            std::vector<int> need2, over2;
            for (int c : unsure) {
                uint32_t f = h_flags[c];
                if (f & CELL_OVERFLOW)
                    over2.push_back(c);
                else if (f & CELL_NEED_K)
                    need2.push_back(c);
            }
            if (tiers(need2, over2, true, P, st)) return -1;
        }
        st.t_pass_bc_ms += ms_since(t3);

        // statistics
        unsigned long long ctr[8];
        be_.d2h(ctr, d_ctr, sizeof(ctr));
        st.nb_clips += ctr[0];
        st.nb_tests += ctr[1];
        st.nb_cells += P.n_items;
        // the other computed cells are done with no flag
        inst_bad.assign(ni_, 0);
        for (size_t k = 0; k < stopped.size(); ++k)
            if (stopped[k]) inst_bad[k]++;
        for (int c : census_) {
            const uint32_t f = h_flags[c];
            if (!(f & CELL_DONE)) st.nb_failed++;
            if (!(f & CELL_DONE) && (f & CELL_OVERFLOW)) st.nb_fail_capacity++;
            if (f & CELL_UNCERTAIN) st.nb_uncertain++;
            if (f & CELL_EMPTY) st.nb_empty++;
            if (f & CELL_DISCONNECTED) st.nb_disconnected++;
            if (!(f & CELL_DONE) || (f & CELL_EMPTY)) {
                const int k = int(std::upper_bound(h_p0_.begin(), h_p0_.begin() + ni_, c) -
                                  h_p0_.begin()) - 1;
                inst_bad[k]++;
            }
        }

        // CSR
        auto t4 = now();
        d_ptr = (long long*)slot(S_PTR, sizeof(long long) * ((size_t)n_ + 1));
        be_.scan_i32(d_cnt, d_ptr, n_);
        be_.d2h(&total_neighbors, d_ptr + n_, sizeof(long long));
        st.nb_neighbors += total_neighbors;
        if (pb.want_neighbors) {
            d_nbr = (int*)slot(S_NBR, sizeof(int) * std::max<long long>(1, total_neighbors));
            d_area = (double*)slot(S_AREA, sizeof(double) * std::max<long long>(1, total_neighbors));
            be_.compact(n_, d_cnt, d_ptr, d_nbr_tmp, d_area_tmp, kFaceStride, d_nbr, d_area);
            if (!piece_cells_.empty())
                be_.piece_copy((int)piece_cells_.size(), (const int*)slots_[S_XCELLS].p, d_orig,
                               (const long long*)slots_[S_XOFF].p, (const int*)slots_[S_XNBR].p,
                               (const double*)slots_[S_XAREA].p, d_cnt, d_ptr, d_nbr, d_area);
        }
        be_.sync();
        st.t_compact_ms += ms_since(t4);
        st.t_total_ms += ms_since(tstart);
        return 0;
    }

    // Cell meshes and primal tetrahedra of one instance: a counting pass of
    // the cells (every tier), exclusive offsets, then, if `cap` holds the
    // totals (vertices, facets, facet entries, tetrahedra), a writing pass
    // into `out`.  tot receives the totals.  Returns 0, or 1 when the
    // capacity is too small (nothing written but the totals), -1 on error.
    int mesh(Problem pb, double shrink, const long long cap[4], MeshOut& out, long long tot[4],
             EngineStats& st) {
        if (pb.n_inst != 1) {
            error = "mesh: one instance";
            return -1;
        }
        const long long n = pb.n;
        int* cnt = (int*)slot(S_MCNT, sizeof(int) * 4 * (size_t)n);
        long long* off = (long long*)slot(S_MOFF, sizeof(long long) * 4 * (size_t)(n + 1));
        pb.want_neighbors = false;
        pb.mesh_mode = 1;
        pb.mesh_cnt = cnt;
        if (cells(pb, st)) return -1;
        for (int k = 0; k < 4; ++k) be_.scan_i32(cnt + k * n, off + k * (n + 1), (int)n);
        for (int k = 0; k < 4; ++k) be_.d2h(&tot[k], off + k * (n + 1) + n, sizeof(long long));
        if (tot[0] > 0x7FFFFFFFll || tot[2] > 0x7FFFFFFFll) {
            error = "mesh: more than 2^31 - 1 vertices or facet entries";
            return -1;
        }
        if (out.cell_off) be_.d2d(out.cell_off, off, sizeof(long long) * 4 * (size_t)(n + 1));
        for (int k = 0; k < 4; ++k)
            if (tot[k] > cap[k]) return 1;
        pb.mesh_mode = 2;
        pb.mesh_off = off;
        pb.mesh_shrink = shrink;
        pb.mesh_v = out.v;
        pb.mesh_fptr = out.fptr;
        pb.mesh_fv = out.fv;
        pb.mesh_fnbr = out.fnbr;
        pb.mesh_tet = out.tet;
        if (cells(pb, st)) return -1;
        if (out.fptr) be_.d2d(out.fptr + tot[1], off + 2 * (n + 1) + n, sizeof(long long));
        be_.sync();
        return 0;
    }

    // Lloyd relaxation in place on d_xyz (caller order), security radius on.
    int lloyd(Problem pb, double* d_xyz, int iters, EngineStats& st) {
        pb.d_xyz = d_xyz;
        pb.want_neighbors = false;
        double* wrap = nullptr;
        if (pb.periodic) {
            wrap = (double*)slot(S_GWRAP, sizeof(double) * 6);
            be_.h2d(wrap, pb.box.data(), sizeof(double) * 6);
        }
        for (int it = 0; it < iters; ++it) {
            if (cells(pb, st)) return -1;
            be_.lloyd_update(d_xyz, d_vol, d_mg, pb.n, wrap);
        }
        be_.sync();
        return 0;
    }

    int n() const { return n_; }

private:
    enum Slot {
        S_P0, S_N, S_B0, S_NB, S_L, S_T0, S_LEAFOFF, S_BOX, S_INSTORIG, S_KEYS, S_KEYS2,
        S_VALS, S_VALS2, S_PTS, S_PF, S_ORIG, S_INSTOF, S_NODES, S_LVLOFF, S_WMAX, S_LIST,
        S_VOL, S_MG, S_CNT, S_NBRTMP, S_AREATMP, S_FLAGS, S_R2, S_CTR, S_PTR, S_NBR, S_AREA,
        S_ITEMS, S_RHO2, S_RCNT, S_ROFF, S_RLIST, S_RD2, S_RLIST2, S_RD22, S_RLEN, S_BALL,
        S_FILT, S_COFF, S_AITEMS, S_TITEMS, S_CENS_C, S_CENS_F, S_CENS_N, S_ABORT, S_E2,
        S_QKEYS, S_QKEYS2, S_QVALS, S_QVALS2, S_QPTS, S_QPF, S_QROW, S_QPOS, S_QR2, S_QREDO,
        S_QLI2, S_QLD, S_QLD2, S_MCNT, S_MOFF, S_GCNT, S_GOFF, S_GBOX, S_GXYZ, S_GW, S_GSRC,
        S_GREACH, S_GWRAP, S_PCELLS, S_PITEMS, S_PLPTR, S_PLLEN, S_PBOX, S_PVOL, S_PMG, S_PE2,
        S_PCNT, S_PNBR, S_PAREA, S_PFLAGS, S_PR2, S_PPTR, S_PIDX, S_XCELLS, S_XOFF, S_XNBR,
        S_XAREA, S_COUNT
    };
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };

    // h_flags[c] = the device flag of each cell c of `cells`
    void refresh_flags(const std::vector<int>& cells, const uint32_t* d_flags) {
        if (cells.empty()) return;
        const size_t m = cells.size();
        int* d_c = (int*)slot(S_CENS_C, sizeof(int) * m);
        uint32_t* d_f = (uint32_t*)slot(S_CENS_F, sizeof(uint32_t) * m);
        be_.h2d(d_c, cells.data(), sizeof(int) * m);
        be_.gather_u32(d_c, (int)m, d_flags, d_f);
        std::vector<uint32_t> f(m);
        be_.d2h(f.data(), d_f, sizeof(uint32_t) * m);
        for (size_t i = 0; i < m; ++i) h_flags[cells[i]] = f[i];
    }

    void* slot(int id, size_t bytes) {
        if (slots_.size() < S_COUNT) slots_.resize(S_COUNT);
        Buf& b = slots_[id];
        if (b.bytes < bytes) {
            if (b.p) be_.dfree(b.p);
            size_t nb = std::max(bytes, b.bytes + b.bytes / 4);
            b.p = be_.dalloc(nb);
            b.bytes = nb;
        }
        return b.p;
    }

    // Pass B for `need` (complete radius lists, 64 corner slots), then pass C
    // with 128 and 256 corner slots for the cells that overflow.  h_flags is
    // refreshed.
    int tiers(const std::vector<int>& need, std::vector<int> over, bool exact, const ClipParams& P,
              EngineStats& st) {
        st.nb_pass_b += need.size();
        if (!need.empty()) {
            if (radius_pass(1, exact, need, P, st)) return -1;
            refresh_flags(need, P.flags);
            for (int c : need)
                if (h_flags[c] & CELL_OVERFLOW) over.push_back(c);
        }
        st.nb_pass_c += over.size();
        if (!over.empty()) {
            if (radius_pass(2, exact, over, P, st)) return -1;
            refresh_flags(over, P.flags);
            std::vector<int> over4;
            for (int c : over)
                if (h_flags[c] & CELL_OVERFLOW) over4.push_back(c);
            if (!over4.empty()) {
                if (radius_pass(4, exact, over4, P, st)) return -1;
                refresh_flags(over4, P.flags);
                std::vector<int> over8;
                for (int c : over4)
                    if (h_flags[c] & CELL_OVERFLOW) over8.push_back(c);
                if (!over8.empty()) {
                    if (radius_pass(8, exact, over8, P, st)) return -1;
                    refresh_flags(over8, P.flags);
                    std::vector<int> over16;
                    for (int c : over8)
                        if (h_flags[c] & CELL_OVERFLOW) over16.push_back(c);
                    if (!over16.empty() && !P.mesh_mode) {
                        if (piece_pass(exact, over16, P, st)) return -1;
                        refresh_flags(over16, P.flags);
                    }
                }
            }
        }
        return 0;
    }

    // Complete lists for `items`: every seed within the distance that can
    // still cut the partial cell of the previous pass and that passes the
    // (power) filter of its ball, in chunks of at most list_chunk
    // entries (or one item), each sorted by distance and clipped in list mode.
    int radius_pass(int S, bool exact, const std::vector<int>& items, ClipParams P,
                    EngineStats& st) {
        const int ni = (int)items.size();
        int* d_items = (int*)slot(S_ITEMS, sizeof(int) * ni);
        be_.h2d(d_items, items.data(), sizeof(int) * ni);
        float* d_rho2 = (float*)slot(S_RHO2, sizeof(float) * ni);
        PowerFilter* d_filt = P.ball ? (PowerFilter*)slot(S_FILT, sizeof(PowerFilter) * ni) : nullptr;
        be_.radius_r2(d_items, ni, d_pts, P.r2, d_inst_of, d_wmax_, weighted_ ? 1 : 0, P.ball,
                      d_box_, d_rho2, d_filt);
        int* d_rcnt = (int*)slot(S_RCNT, sizeof(int) * ni);
        be_.radius_count(d_pf, d_nodes, IL, d_items, ni, d_rho2, d_filt, d_rcnt);
        long long* d_roff = (long long*)slot(S_ROFF, sizeof(long long) * (ni + 1));
        be_.scan_i32(d_rcnt, d_roff, ni);
        std::vector<long long> off(ni + 1);
        be_.d2h(off.data(), d_roff, sizeof(long long) * (ni + 1));
        st.radius_entries += off[ni];
        for (int a = 0; a < ni;) {
            int b = a + 1;
            while (b < ni && off[b + 1] - off[a] <= list_chunk) ++b;
            const long long base = off[a], total = off[b] - base;
            const int nc = b - a;
            size_t cap = (size_t)std::max<long long>(total, 1);
            int* d_rl = (int*)slot(S_RLIST, sizeof(int) * cap);
            float* d_rd = (float*)slot(S_RD2, sizeof(float) * cap);
            int* d_rl2 = (int*)slot(S_RLIST2, sizeof(int) * cap);
            float* d_rd2 = (float*)slot(S_RD22, sizeof(float) * cap);
            long long* d_coff = (long long*)slot(S_COFF, sizeof(long long) * (nc + 1));
            be_.radius_fill(d_pf, d_nodes, IL, d_items + a, nc, d_rho2 + a,
                            d_filt ? d_filt + a : nullptr, d_roff + a, base, d_rl, d_rd);
            be_.rebase(d_roff + a, nc + 1, base, d_coff);
            be_.seg_sort(d_rd, d_rl, d_rd2, d_rl2, total, nc, d_coff);  // into rd2 / rl2
            ClipParams Q = P;
            Q.items = d_items + a;
            Q.n_items = nc;
            Q.list = d_rl2;
            Q.list_ptr = d_coff;
            Q.list_len = d_rcnt + a;
            Q.list_stride = 0;
            Q.list_complete = 1;
            be_.clip(S, exact, false, Q);
            be_.sync();
            a = b;
        }
        return 0;
    }

    // Cells over the plane limit (more than 256 facets), each computed in
    // pieces: its complete radius list (as radius_pass), one item per octant
    // of the instance's box around the seed (512 corner slots), a piece still
    // over the limit split into the eight sub-boxes of its box (at most
    // kPieceLevels levels), the finished pieces merged per cell.  The merged
    // neighbour lists are kept for the CSR.  Cells whose pieces fail stay
    // flagged.
    int piece_pass(bool exact, const std::vector<int>& cells, ClipParams P, EngineStats& st) {
        const int m = (int)cells.size();
        int* d_cells = (int*)slot(S_PCELLS, sizeof(int) * m);
        be_.h2d(d_cells, cells.data(), sizeof(int) * m);
        float* d_rho2 = (float*)slot(S_RHO2, sizeof(float) * m);
        PowerFilter* d_filt = P.ball ? (PowerFilter*)slot(S_FILT, sizeof(PowerFilter) * m) : nullptr;
        be_.radius_r2(d_cells, m, d_pts, P.r2, d_inst_of, d_wmax_, weighted_ ? 1 : 0, P.ball,
                      d_box_, d_rho2, d_filt);
        int* d_rcnt = (int*)slot(S_RCNT, sizeof(int) * m);
        be_.radius_count(d_pf, d_nodes, IL, d_cells, m, d_rho2, d_filt, d_rcnt);
        long long* d_roff = (long long*)slot(S_ROFF, sizeof(long long) * (m + 1));
        be_.scan_i32(d_rcnt, d_roff, m);
        std::vector<long long> roff(m + 1);
        std::vector<int> rcnt(m);
        be_.d2h(roff.data(), d_roff, sizeof(long long) * (m + 1));
        be_.d2h(rcnt.data(), d_rcnt, sizeof(int) * m);
        const long long total = roff[m];
        if (total > list_chunk) return 0;  // the cells stay failed
        st.radius_entries += total;
        const size_t cap = (size_t)std::max<long long>(total, 1);
        int* d_rl = (int*)slot(S_RLIST, sizeof(int) * cap);
        float* d_rd = (float*)slot(S_RD2, sizeof(float) * cap);
        int* d_rl2 = (int*)slot(S_RLIST2, sizeof(int) * cap);
        float* d_rd2 = (float*)slot(S_RD22, sizeof(float) * cap);
        be_.radius_fill(d_pf, d_nodes, IL, d_cells, m, d_rho2, d_filt, d_roff, 0, d_rl, d_rd);
        be_.seg_sort(d_rd, d_rl, d_rd2, d_rl2, total, m, d_roff);
        // the octants, then level after level
        double* d_b0 = (double*)slot(S_PBOX, sizeof(double) * 6 * kPieces * m);
        be_.piece_setup(kPieces * m, d_cells, d_pts, d_inst_of, d_box_, d_b0);
        std::vector<double> box(6 * (size_t)kPieces * m);
        be_.d2h(box.data(), d_b0, sizeof(double) * box.size());
        std::vector<int> owner(kPieces * (size_t)m);
        for (int q = 0; q < kPieces * m; ++q) owner[q] = q / kPieces;
        std::vector<std::vector<int>> fin(m);
        std::vector<char> bad(m, 0);
        long long npieces = 0;
        const uint32_t stop = CELL_NEED_K | CELL_BAD_TOPO;
        for (int level = 0; !owner.empty(); ++level) {
            const int nq = (int)owner.size();
            if (level == kPieceLevels || npieces + nq > kMaxPieces) {
                for (int t : owner) bad[t] = 1;
                break;
            }
            const long long base = npieces;
            npieces += nq;
            std::vector<int> items(nq);
            std::vector<long long> lptr(nq);
            std::vector<int> llen(nq);
            for (int q = 0; q < nq; ++q) {
                items[q] = cells[owner[q]];
                lptr[q] = roff[owner[q]];
                llen[q] = rcnt[owner[q]];
            }
            int* p_items = (int*)slot(S_PITEMS, sizeof(int) * nq);
            long long* p_lptr = (long long*)slot(S_PLPTR, sizeof(long long) * nq);
            int* p_llen = (int*)slot(S_PLLEN, sizeof(int) * nq);
            double* p_box = (double*)slot(S_PBOX, sizeof(double) * 6 * nq);
            be_.h2d(p_items, items.data(), sizeof(int) * nq);
            be_.h2d(p_lptr, lptr.data(), sizeof(long long) * nq);
            be_.h2d(p_llen, llen.data(), sizeof(int) * nq);
            be_.h2d(p_box, box.data(), sizeof(double) * 6 * nq);
            const size_t kb = sizeof(double) * base, kn = sizeof(double) * npieces;
            double* p_vol = (double*)grow_keep(S_PVOL, kb, kn);
            double* p_mg = (double*)grow_keep(S_PMG, 3 * kb, 3 * kn);
            double* p_e2 = P.e2 ? (double*)grow_keep(S_PE2, kb, kn) : nullptr;
            int* p_cnt = (int*)grow_keep(S_PCNT, sizeof(int) * base, sizeof(int) * npieces);
            int* p_nbr = (int*)grow_keep(S_PNBR, sizeof(int) * base * kFaceStride,
                                         sizeof(int) * npieces * kFaceStride);
            double* p_area = (double*)grow_keep(S_PAREA, kb * kFaceStride, kn * kFaceStride);
            uint32_t* p_flags = (uint32_t*)grow_keep(S_PFLAGS, sizeof(uint32_t) * base,
                                                     sizeof(uint32_t) * npieces);
            double* p_r2 = (double*)grow_keep(S_PR2, kb, kn);
            ClipParams Q = P;
            Q.items = p_items;
            Q.n_items = nq;
            Q.list = d_rl2;
            Q.list_ptr = p_lptr;
            Q.list_len = p_llen;
            Q.list_stride = 0;
            Q.list_complete = 1;
            Q.item_box = p_box;
            Q.piece = 1;
            Q.vol = p_vol + base;
            Q.mg = p_mg + 3 * base;
            Q.e2 = p_e2 ? p_e2 + base : nullptr;
            Q.nbr_cnt = p_cnt + base;
            Q.nbr_tmp = p_nbr + base * kFaceStride;
            Q.area_tmp = p_area + base * kFaceStride;
            Q.face_stride = kFaceStride;
            Q.flags = p_flags + base;
            Q.r2 = p_r2 + base;
            Q.ball = nullptr;
            be_.clip(8, exact, false, Q);
            std::vector<uint32_t> f(nq);
            be_.d2h(f.data(), p_flags + base, sizeof(uint32_t) * nq);
            std::vector<int> next_owner;
            std::vector<double> next_box;
            for (int q = 0; q < nq; ++q) {
                const int t = owner[q];
                if ((f[q] & CELL_DONE) && !(f[q] & (stop | CELL_OVERFLOW))) {
                    fin[t].push_back(int(base + q));
                } else if ((f[q] & CELL_OVERFLOW) && !(f[q] & stop)) {
                    const double* b = box.data() + 6 * (size_t)q;
                    for (int k = 0; k < kPieces; ++k) {
                        next_owner.push_back(t);
                        for (int a = 0; a < 3; ++a) {
                            const double mid = 0.5 * (b[a] + b[3 + a]);
                            const bool upper = (k >> a) & 1;
                            next_box.push_back(upper ? mid : b[a]);
                        }
                        for (int a = 0; a < 3; ++a) {
                            const double mid = 0.5 * (b[a] + b[3 + a]);
                            const bool upper = (k >> a) & 1;
                            next_box.push_back(upper ? b[3 + a] : mid);
                        }
                    }
                } else {
                    bad[t] = 1;
                }
            }
            owner.swap(next_owner);
            box.swap(next_box);
        }
        // merge the cells whose pieces all finished
        std::vector<int> cnt(npieces);
        be_.d2h(cnt.data(), slots_[S_PCNT].p, sizeof(int) * npieces);
        std::vector<int> mc;
        std::vector<long long> pptr(1, 0), xoff;
        std::vector<int> pidx;
        long long xt = x_total_;
        for (int t = 0; t < m; ++t) {
            if (bad[t]) continue;
            mc.push_back(cells[t]);
            long long c = 0;
            for (int q : fin[t]) {
                pidx.push_back(q);
                c += cnt[q];
            }
            pptr.push_back((long long)pidx.size());
            xoff.push_back(xt);
            xt += c;
        }
        const int mm = (int)mc.size();
        st.nb_pieces += mm;
        if (mm == 0) return 0;
        const size_t b0 = piece_cells_.size();
        piece_cells_.insert(piece_cells_.end(), mc.begin(), mc.end());
        const size_t b1 = piece_cells_.size();
        int* x_cells = (int*)grow_keep(S_XCELLS, sizeof(int) * b0, sizeof(int) * b1);
        long long* x_off = (long long*)grow_keep(S_XOFF, sizeof(long long) * b0, sizeof(long long) * b1);
        int* x_nbr = (int*)grow_keep(S_XNBR, sizeof(int) * x_total_,
                                     sizeof(int) * std::max<long long>(xt, 1));
        double* x_area = (double*)grow_keep(S_XAREA, sizeof(double) * x_total_,
                                            sizeof(double) * std::max<long long>(xt, 1));
        x_total_ = xt;
        be_.h2d(x_cells + b0, mc.data(), sizeof(int) * mm);
        be_.h2d(x_off + b0, xoff.data(), sizeof(long long) * mm);
        long long* d_pptr = (long long*)slot(S_PPTR, sizeof(long long) * (mm + 1));
        int* d_pidx = (int*)slot(S_PIDX, sizeof(int) * std::max<size_t>(pidx.size(), 1));
        be_.h2d(d_pptr, pptr.data(), sizeof(long long) * (mm + 1));
        if (!pidx.empty()) be_.h2d(d_pidx, pidx.data(), sizeof(int) * pidx.size());
        be_.piece_merge(mm, x_cells + b0, d_orig, d_pptr, d_pidx, (const double*)slots_[S_PVOL].p,
                        (const double*)slots_[S_PMG].p,
                        P.e2 ? (const double*)slots_[S_PE2].p : nullptr,
                        (const int*)slots_[S_PCNT].p, (const int*)slots_[S_PNBR].p,
                        (const double*)slots_[S_PAREA].p, (const uint32_t*)slots_[S_PFLAGS].p,
                        kFaceStride, x_off + b0, P.vol, P.mg, P.e2, P.nbr_cnt, P.flags, x_nbr,
                        x_area);
        be_.sync();
        return 0;
    }
    // slot id grown to `bytes`, its first `keep` bytes kept
    void* grow_keep(int id, size_t keep, size_t bytes) {
        if (slots_.size() < S_COUNT) slots_.resize(S_COUNT);
        Buf& b = slots_[id];
        if (b.bytes >= bytes) return b.p;
        size_t nb = std::max(bytes, b.bytes + b.bytes / 4);
        void* p = be_.dalloc(nb);
        if (keep) be_.d2d(p, b.p, keep);
        if (b.p) be_.dfree(b.p);
        b.p = p;
        b.bytes = nb;
        return p;
    }

    static std::chrono::steady_clock::time_point now() { return std::chrono::steady_clock::now(); }
    static double ms_since(std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(now() - t).count();
    }

    B& be_;
    std::vector<Buf> slots_;
    int n_ = 0, ni_ = 1, max_h_ = 0;
    long long n_nodes_ = 0;
    bool weighted_ = false;
    int tuned_n_ = -1;
    double* d_box_ = nullptr;
    double* d_wmax_ = nullptr;
    double* d_r2_ = nullptr;  // squared radius per cell of the last pass (Morton order)
    std::vector<int> h_p0_, h_n_, h_b0_, h_nb_, h_L_;
    // first passes left to run with FP64 corners (after an FP32 pass that
    // redid more than 5 % of its cells)
    static constexpr int kFp32Skip = 8;
    int fp32_skip_ = 0;
    std::vector<int> piece_cells_;  // cells of this pass computed in pieces
    long long x_total_ = 0;         // their merged neighbour entries
    static constexpr int kPieceLevels = 4;
    static constexpr long long kMaxPieces = 1 << 16;
    std::vector<long long> h_t0_, h_leaf_off_;
};

}  // namespace vgc

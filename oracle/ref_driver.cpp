// ORACLE — test infrastructure only.
//
// C entry points over geogram's own code (the Delaunay pluggable module that
// oracle/build_ref.sh assembles from the reference sources): the reference
// path itself, used to pin the oracle and as the CPU baseline.
//
//  georef_delaunay_cells: geogram's fastest CPU path for Voronoi / Laguerre
//    cells in a box — PeriodicDelaunay3d(false) with the seeds (and
//    weights), copy_Laguerre_cell_from_Delaunay, six box clips, then
//    compute_geometry / volume / compute_mg / facet_area / v_global_index of
//    the contributing planes (delaunay/periodic_delaunay_3d.cpp:3634-3733,
//    voronoi/convex_cell.cpp:1160-1307; compute_RVD's cube mode,
//    examples/geogram/compute_RVD/main.cpp:450-477), cells in parallel;
//  georef_periodic_cells: the same over the 3-torus of the box,
//    PeriodicDelaunay3d(period, true) and periodic_vertex_real
//    (delaunay/periodic_delaunay_3d.h, delaunay/periodic.h);
//  georef_knn: NearestNeighborSearch "BNN" (points/kd_tree.cpp),
//    get_nearest_neighbors of each point by index (the point included).
#include "Delaunay_psm.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

using namespace GEO;

namespace {

std::once_flag g_init;

void init(int nthreads) {
    std::call_once(g_init, [] {
        GEO::initialize(GEO::GEOGRAM_INSTALL_NONE);
        Logger::instance()->set_quiet(true);
    });
    if (nthreads > 0) Process::set_max_threads(index_t(nthreads));
}

double seconds_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

extern "C" {

// Cells of the n seeds (w: weights or NULL) clipped to box: vol[n],
// mg[3n] (integral of x), cnt[n] and nbr / farea [n * stride] (contributing
// neighbours, NO_INDEX box planes left out).  t[0]: Delaunay seconds,
// t[1]: cells seconds.  Returns 0, -1 if a cell has more than stride
// neighbours, -3 if a weighted seed has an empty cell (geogram's
// triangulation stops there).
int georef_delaunay_cells(const double* xyz, const double* w, int n, const double* box,
                          int nthreads, double* vol, double* mg, int* cnt, int* nbr,
                          double* farea, int stride, double* t) {
    init(nthreads);
    auto t0 = std::chrono::steady_clock::now();
    PeriodicDelaunay3d D(false);
    // Unweighted: the vertex at infinity bounds the cells of hull seeds, as
    // in compute_RVD's cube mode.  Weighted: update_v_to_cell with the
    // vertex at infinity indexes its table with the marker of deleted
    // tetrahedra (delaunay/periodic_delaunay_3d.cpp:3451-3458), so the
    // hull cells are bounded instead by 8 guard seeds of weight 0 at the
    // corners of a cube 10 times the box around it: their power planes stay
    // 5 box sizes away from every seed, outside the box, and their cells
    // are not reported.
    std::vector<double> gx, gw;
    index_t nv = index_t(n);
    if (w) {
        const double cx = 0.5 * (box[0] + box[3]), cy = 0.5 * (box[1] + box[4]),
                     cz = 0.5 * (box[2] + box[5]);
        const double R = 10.0 * std::max(box[3] - box[0], std::max(box[4] - box[1], box[5] - box[2]));
        gx.assign(xyz, xyz + 3 * size_t(n));
        gw.assign(w, w + n);
        for (int k = 0; k < 8; ++k) {
            gx.push_back(cx + ((k & 1) ? R : -R));
            gx.push_back(cy + ((k & 2) ? R : -R));
            gx.push_back(cz + ((k & 4) ? R : -R));
            gw.push_back(0.0);
        }
        nv += 8;
        D.set_vertices(nv, gx.data());
        D.set_weights(gw.data());
    } else {
        D.set_keeps_infinite(true);
        D.set_vertices(nv, xyz);
    }
    D.compute();
    t[0] = seconds_since(t0);
    // a weighted seed without a cell stops the triangulation
    // (periodic_delaunay_3d.cpp:3165-3167): the reference has no cells then
    if (D.has_empty_cells()) return -3;
    auto t1 = std::chrono::steady_clock::now();
    std::atomic<int> err(0);
    parallel_for_slice(0, index_t(n), [&](index_t b, index_t e) {
        VBW::ConvexCell C(VBW::WithVGlobal);
        PeriodicDelaunay3d::IncidentTetrahedra W;
        for (index_t v = b; v < e; ++v) {
            D.copy_Laguerre_cell_from_Delaunay(v, C, W);
            cnt[v] = 0;
            if (C.empty()) {
                // a hidden vertex of the regular triangulation: empty cell
                vol[v] = 0.0;
                mg[3 * v] = mg[3 * v + 1] = mg[3 * v + 2] = 0.0;
                continue;
            }
            // box planes with the "no seed" global index (clip_by_plane(P)
            // alone would leave the slot's index of an earlier cell)
            const VBW::global_index_t none = VBW::global_index_t(-1);
            C.clip_by_plane(vec4(1.0, 0.0, 0.0, -box[0]), none);
            C.clip_by_plane(vec4(-1.0, 0.0, 0.0, box[3]), none);
            C.clip_by_plane(vec4(0.0, 1.0, 0.0, -box[1]), none);
            C.clip_by_plane(vec4(0.0, -1.0, 0.0, box[4]), none);
            C.clip_by_plane(vec4(0.0, 0.0, 1.0, -box[2]), none);
            C.clip_by_plane(vec4(0.0, 0.0, -1.0, box[5]), none);
            cnt[v] = 0;
            if (C.empty()) {
                vol[v] = 0.0;
                mg[3 * v] = mg[3 * v + 1] = mg[3 * v + 2] = 0.0;
                continue;
            }
            C.compute_geometry();
            double m = 0.0;
            vec3 g;
            C.compute_mg(m, g);
            vol[v] = m;
            mg[3 * v] = g.x;
            mg[3 * v + 1] = g.y;
            mg[3 * v + 2] = g.z;
            for (index_t lv = 1; lv < C.nb_v(); ++lv) {
                if (!C.vertex_is_contributing(lv)) continue;
                const VBW::global_index_t j = C.v_global_index(lv);
                if (j == VBW::global_index_t(-1) || j >= VBW::global_index_t(n)) continue;
                if (cnt[v] >= stride) {
                    err = -1;
                    break;
                }
                nbr[size_t(v) * stride + cnt[v]] = int(j);
                farea[size_t(v) * stride + cnt[v]] = C.facet_area(lv);
                ++cnt[v];
            }
        }
    });
    t[1] = seconds_since(t1);
    return err;
}

// Periodic cells of the n seeds in the box (one period of a 3-torus):
// PeriodicDelaunay3d(period, true) over the seeds translated to the origin,
// copy_Laguerre_cell_from_Delaunay of each seed (no clip), the neighbours'
// real indices (periodic_vertex_real).  Outputs as georef_delaunay_cells
// (mg in the box's frame).
int georef_periodic_cells(const double* xyz, const double* w, int n, const double* box,
                          int nthreads, double* vol, double* mg, int* cnt, int* nbr,
                          double* farea, int stride, double* t) {
    init(nthreads);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> x0(3 * size_t(n));
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) x0[3 * size_t(i) + a] = xyz[3 * size_t(i) + a] - box[a];
    PeriodicDelaunay3d D(vec3(box[3] - box[0], box[4] - box[1], box[5] - box[2]), true);
    D.set_vertices(index_t(n), x0.data());
    if (w) D.set_weights(w);
    D.compute();
    t[0] = seconds_since(t0);
    if (D.has_empty_cells()) return -3;
    auto t1 = std::chrono::steady_clock::now();
    std::atomic<int> err(0);
    parallel_for_slice(0, index_t(n), [&](index_t b, index_t e) {
        VBW::ConvexCell C(VBW::WithVGlobal);
        PeriodicDelaunay3d::IncidentTetrahedra W;
        for (index_t v = b; v < e; ++v) {
            D.copy_Laguerre_cell_from_Delaunay(v, C, W);
            cnt[v] = 0;
            if (C.empty()) {
                vol[v] = 0.0;
                mg[3 * v] = mg[3 * v + 1] = mg[3 * v + 2] = 0.0;
                continue;
            }
            C.compute_geometry();
            double m = 0.0;
            vec3 g;
            C.compute_mg(m, g);
            vol[v] = m;
            mg[3 * v] = g.x + m * box[0];
            mg[3 * v + 1] = g.y + m * box[1];
            mg[3 * v + 2] = g.z + m * box[2];
            for (index_t lv = 1; lv < C.nb_v(); ++lv) {
                if (!C.vertex_is_contributing(lv)) continue;
                const VBW::global_index_t j = C.v_global_index(lv);
                if (j == VBW::global_index_t(-1)) continue;
                if (cnt[v] >= stride) {
                    err = -1;
                    break;
                }
                nbr[size_t(v) * stride + cnt[v]] = int(D.periodic_vertex_real(index_t(j)));
                farea[size_t(v) * stride + cnt[v]] = C.facet_area(lv);
                ++cnt[v];
            }
        }
    });
    t[1] = seconds_since(t1);
    return err;
}

// k nearest points of each point (the point itself first), by geogram's
// kd-tree: idx / d2 are n * k.
int georef_knn(const double* xyz, int n, int k, int* idx, double* d2) {
    init(0);
    NearestNeighborSearch_var nn = NearestNeighborSearch::create(3, "BNN");
    nn->set_points(index_t(n), xyz);
    parallel_for(0, index_t(n), [&](index_t i) {
        std::vector<index_t> ix(k);
        nn->get_nearest_neighbors(index_t(k), i, ix.data(), d2 + size_t(i) * k);
        for (int j = 0; j < k; ++j) idx[size_t(i) * k + j] = int(ix[j]);
    });
    return 0;
}

int georef_max_threads(void) {
    init(0);
    return int(Process::maximum_concurrent_threads());
}

}  // extern "C"

// geogram-side binding of the MI355X path: the translation unit a geogram
// maintainer adds to libgeogram (with include/vgc.h on the include path and
// libvgc.so on the link line).  It registers GPU-backed implementations
// behind geogram's own extension points; callers only change the algorithm
// name they select:
//
//   NearestNeighborSearch "VGC"      points/nn_search.h:69-200, factory at :288
//   VgcRestrictedVoronoiDiagram      voronoi/RVD.h:64-720 (volumetric, box domain)
//   VgcCentroidalVoronoiTesselation  voronoi/CVT.h:157,164 (Lloyd / Newton)
//   VgcOptimalTransportMap3d         the calls of examples/exploragram/
//                                    compute_OTM/main.cpp:283-297
//
// Domain: the bounding box of the mesh handed to the constructors (the
// meshless path clips cells to a box, as compute_RVD's cube mode does).
// Errors from the library become geo_assert failures, geogram's convention.
#include <geogram/basic/assert.h>
#include <geogram/basic/logger.h>
#include <geogram/delaunay/delaunay.h>
#include <geogram/mesh/mesh.h>
#include <geogram/mesh/mesh_geometry.h>
#include <geogram/points/nn_search.h>
#include <geogram/voronoi/CVT.h>
#include <geogram/voronoi/RVD.h>

#include <vector>

#include "vgc.h"

namespace GEO {

namespace {

// one library context per thread of the caller (a context is used by one
// host thread at a time)
vgc_ctx* vgc_context() {
    thread_local struct Holder {
        vgc_ctx* ctx = nullptr;
        Holder() { geo_assert(vgc_create(&ctx, 0) == VGC_OK); }
        ~Holder() { vgc_destroy(ctx); }
    } holder;
    return holder.ctx;
}

void vgc_check(int rc) {
    if (rc < 0) {
        Logger::err("VGC") << vgc_last_error(vgc_context()) << std::endl;
        geo_assert_not_reached;
    }
    if (rc == VGC_PARTIAL) Logger::warn("VGC") << vgc_last_error(vgc_context()) << std::endl;
}

void mesh_box(const Mesh* M, double box[6]) {
    geo_assert(M != nullptr && M->vertices.nb() > 0);
    get_bbox(*M, box, box + 3);
}

}  // namespace

// ------------------------------------------------------------------- kNN

// Exact FP64 nearest neighbours on the GPU.  Equal distances come in
// ascending index order (the kd-tree puts the later point first,
// points/kd_tree.h:176-190).
class VgcNearestNeighborSearch : public NearestNeighborSearch {
public:
    explicit VgcNearestNeighborSearch(coord_index_t dim) : NearestNeighborSearch(dim) {
        geo_assert(dim == 3);
    }

    void set_points(index_t nb_points, const double* points) override {
        NearestNeighborSearch::set_points(nb_points, points);
    }

    bool stride_supported() const override { return false; }

    void get_nearest_neighbors(index_t nb_neighbors, const double* query_point,
                               index_t* neighbors, double* neighbors_sq_dist) const override {
        geo_assert(nb_neighbors <= nb_points());
        vgc_check(vgc_knn_search(vgc_context(), points_, nb_points_, query_point, 1, nb_neighbors,
                                 0, reinterpret_cast<int32_t*>(neighbors), neighbors_sq_dist));
    }

    void get_nearest_neighbors(index_t nb_neighbors, const double* query_point,
                               index_t* neighbors, double* neighbors_sq_dist,
                               KeepInitialValues) const override {
        get_nearest_neighbors(nb_neighbors, query_point, neighbors, neighbors_sq_dist);
    }

    void get_nearest_neighbors(index_t nb_neighbors, index_t query_point, index_t* neighbors,
                               double* neighbors_sq_dist) const override {
        get_nearest_neighbors(nb_neighbors, point_ptr(query_point), neighbors, neighbors_sq_dist);
    }

    // every point's nearest others at once (the form Delaunay_NearestNeighbors
    // wants, delaunay/delaunay_nn.cpp:73-84): one GPU call instead of n
    void get_neighbors_of_all(index_t nb_neighbors, std::vector<int32_t>& neighbors,
                              std::vector<double>& sq_dist) const {
        neighbors.resize(size_t(nb_points()) * nb_neighbors);
        sq_dist.resize(neighbors.size());
        vgc_check(vgc_knn(vgc_context(), points_, nb_points_, nb_neighbors, 0, neighbors.data(),
                          sq_dist.data()));
    }
};

// -------------------------------------------------- restricted Voronoi

// Restricted Voronoi diagram of the Delaunay vertices: in the box of `mesh`
// (volumetric, the default) or on the triangles of `mesh` (set_volumetric
// (false)).  compute_centroids_in_volume, compute_CVT_func_grad_in_volume
// and compute_centroids_on_surface run on the GPU (accumulated into the
// caller's zeroed arrays, RVD.cpp:468-478, 878-899, 371-415; a vertex
// attribute "weight" is the surface density, as in
// ComputeCentroidsWeighted); the primal tetrahedra of compute_RDT too.
// Mesh-traversal callbacks are not part of the meshless path.
class VgcRestrictedVoronoiDiagram : public RestrictedVoronoiDiagram {
public:
    VgcRestrictedVoronoiDiagram(Delaunay* delaunay, Mesh* mesh)
        : RestrictedVoronoiDiagram(delaunay, mesh, nullptr, 0) {
        geo_assert(delaunay->dimension() == 3);
        mesh_box(mesh, box_);
        volumetric_ = true;
    }

    void set_volumetric(bool x) override { volumetric_ = x; }
    void set_check_SR(bool x) override { geo_assert(x); }
    void set_exact_predicates(bool) override {}
    bool exact_predicates() const override { return true; }
    void create_threads() override {}
    void delete_threads() override {}

    void compute_centroids_in_volume(double* mg, double* m) override {
        const index_t n = delaunay()->nb_vertices();
        std::vector<double> vol(n), g(3 * size_t(n));
        vgc_stats st;
        vgc_check(vgc_cells(vgc_context(), delaunay()->vertex_ptr(0), nullptr, n, box_, 0,
                            VGC_NO_NEIGHBORS, vol.data(), g.data(), nullptr, nullptr, nullptr, 0,
                            &st));
        for (index_t i = 0; i < n; ++i) {
            m[i] += vol[i];
            for (coord_index_t c = 0; c < 3; ++c) mg[3 * i + c] += g[3 * i + c];
        }
    }

    void compute_CVT_func_grad_in_volume(double& f, double* g) override {
        vgc_stats st;
        vgc_check(vgc_cvt_func_grad(vgc_context(), delaunay()->vertex_ptr(0),
                                    delaunay()->nb_vertices(), box_, 0, &f, g, &st));
    }

    void compute_RDT(vector<index_t>& simplices, vector<double>& embedding, RDTMode,
                     const vector<bool>&, MeshFacetsAABB*) override {
        const index_t n = delaunay()->nb_vertices();
        uint64_t nt = 0;
        vgc_stats st;
        int rc = vgc_rdt(vgc_context(), delaunay()->vertex_ptr(0), nullptr, n, box_, 0, nullptr,
                         0, &nt, &st);
        geo_assert(rc == VGC_OK || rc == VGC_ERR_CAPACITY);
        simplices.resize(index_t(4 * nt));
        vgc_check(vgc_rdt(vgc_context(), delaunay()->vertex_ptr(0), nullptr, n, box_, 0,
                          reinterpret_cast<int32_t*>(simplices.data()), nt, &nt, &st));
        embedding.assign(delaunay()->vertex_ptr(0), delaunay()->vertex_ptr(0) + 3 * size_t(n));
    }

    // the cells as polyhedra (compute_RVD's cube mode), one mesh facet per
    // cell facet, the neighbour's index in the facet attribute "nbr"
    void compute_RVD(Mesh& M, coord_index_t, bool, bool) override {
        const index_t n = delaunay()->nb_vertices();
        const double* x = delaunay()->vertex_ptr(0);
        vgc_mesh_sizes need{};
        vgc_stats st;
        vgc_check(vgc_cell_mesh(vgc_context(), x, nullptr, n, box_, 0.0, 0, nullptr, nullptr,
                                nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &need, &st));
        std::vector<double> V(3 * need.nb_vertices);
        std::vector<uint64_t> fptr(need.nb_facets + 1), cvp(n + 1), cfp(n + 1);
        std::vector<int32_t> fv(need.nb_facet_vertices), fnbr(need.nb_facets);
        vgc_check(vgc_cell_mesh(vgc_context(), x, nullptr, n, box_, 0.0, 0, cvp.data(), cfp.data(),
                                V.data(), fptr.data(), fv.data(), fnbr.data(), nullptr, &need,
                                &need, &st));
        M.clear();
        M.vertices.set_dimension(3);
        M.vertices.create_vertices(index_t(need.nb_vertices));
        for (index_t v = 0; v < index_t(need.nb_vertices); ++v)
            for (coord_index_t c = 0; c < 3; ++c) M.vertices.point_ptr(v)[c] = V[3 * v + c];
        Attribute<index_t> nbr(M.facets.attributes(), "nbr");
        for (index_t f = 0; f < index_t(need.nb_facets); ++f) {
            const index_t b = index_t(fptr[f]), e = index_t(fptr[f + 1]);
            const index_t F = M.facets.create_polygon(e - b);
            for (index_t k = b; k < e; ++k) M.facets.set_vertex(F, k - b, index_t(fv[k]));
            nbr[F] = index_t(fnbr[f]);
        }
    }

    void compute_centroids_on_surface(double* mg, double* m) override {
        const Mesh* S = mesh();
        const index_t n = delaunay()->nb_vertices(), nv = S->vertices.nb(), nt = S->facets.nb();
        std::vector<double> V(3 * size_t(nv)), w;
        std::vector<int32_t> F(3 * size_t(nt));
        for (index_t v = 0; v < nv; ++v)
            for (coord_index_t c = 0; c < 3; ++c) V[3 * v + c] = S->vertices.point_ptr(v)[c];
        for (index_t f = 0; f < nt; ++f) {
            geo_assert(S->facets.nb_vertices(f) == 3);  // triangulated surfaces
            for (index_t k = 0; k < 3; ++k) F[3 * f + k] = int32_t(S->facets.vertex(f, k));
        }
        Attribute<double> weight;
        weight.bind_if_is_defined(S->vertices.attributes(), "weight");
        if (weight.is_bound()) w.assign(weight.data(), weight.data() + nv);
        std::vector<double> g(3 * size_t(n)), mm(n);
        vgc_surface_stats st;
        vgc_check(vgc_surface_centroids(vgc_context(), delaunay()->vertex_ptr(0), n, V.data(), nv,
                                        F.data(), nt, w.empty() ? nullptr : w.data(), 0, g.data(),
                                        mm.data(), &st));
        for (index_t i = 0; i < n; ++i) {
            m[i] += mm[i];
            for (coord_index_t c = 0; c < 3; ++c) mg[3 * i + c] += g[3 * i + c];
        }
    }
    void compute_CVT_func_grad_on_surface(double&, double*) override { geo_assert_not_reached; }
    bool compute_initial_sampling_on_surface(double*, index_t, bool) override {
        geo_assert_not_reached;
    }
    bool compute_initial_sampling_in_volume(double*, index_t, bool) override {
        geo_assert_not_reached;
    }
    void project_points_on_surface(index_t, double*, vec3*, bool) override {
        geo_assert_not_reached;
    }
    void compute_RVC(index_t, Mesh&, Mesh&, bool) override { geo_assert_not_reached; }
    void for_each_polyhedron(RVDPolyhedronCallback&, bool, bool, bool) override {
        geo_assert_not_reached;
    }
    void for_each_polygon(RVDPolygonCallback&, bool, bool, bool) override {
        geo_assert_not_reached;
    }
    void set_facets_range(index_t, index_t) override {}
    void set_tetrahedra_range(index_t, index_t) override {}
    GEOGen::PointAllocator* point_allocator() override { return nullptr; }

    const double* box() const { return box_; }

private:
    double box_[6];
};

// ------------------------------------------------------------------- CVT

// Lloyd (security radius on: the stock loop turns it off, CVT.cpp:139, and
// its cells overlap) and quasi-Newton iterations on the GPU, in the box of
// `mesh`.
class VgcCentroidalVoronoiTesselation : public CentroidalVoronoiTesselation {
public:
    explicit VgcCentroidalVoronoiTesselation(Mesh* mesh)
        : CentroidalVoronoiTesselation(mesh, 3, "default") {
        mesh_box(mesh, box_);
    }

    void Lloyd_iterations(index_t nb_iter) override {
        vgc_stats st;
        vgc_check(vgc_lloyd(vgc_context(), points_.data(), nb_points(), box_, nb_iter, 0, 0, &st));
        delaunay_->set_vertices(nb_points(), points_.data());
        newiteration();
    }

    void Newton_iterations(index_t nb_iter, index_t m) override {
        vgc_opt_stats st;
        vgc_check(vgc_cvt_newton(vgc_context(), points_.data(), nb_points(), box_, nb_iter, m, 0,
                                 &st));
        delaunay_->set_vertices(nb_points(), points_.data());
        newiteration();
    }

private:
    double box_[6];
};

// ----------------------------------------------------- optimal transport

// The calls compute_OTM makes on Exploragram's OptimalTransportMap3d
// (examples/exploragram/compute_OTM/main.cpp:283-297), over the box of
// the source mesh with its uniform density.
class VgcOptimalTransportMap3d {
public:
    explicit VgcOptimalTransportMap3d(Mesh* mesh) { mesh_box(mesh, box_); }

    void set_points(index_t nb_points, const double* points, index_t stride = 3) {
        points_.resize(3 * size_t(nb_points));
        for (index_t i = 0; i < nb_points; ++i)
            for (coord_index_t c = 0; c < 3; ++c) points_[3 * i + c] = points[stride * i + c];
        const double V = (box_[3] - box_[0]) * (box_[4] - box_[1]) * (box_[5] - box_[2]);
        nu_.assign(nb_points, V / nb_points);
        weights_.assign(nb_points, 0.0);
    }
    void set_nu(index_t i, double nu) { nu_[i] = nu; }
    void set_epsilon(double eps) { eps_ = eps; }
    // optimize_levels: start each level from the convex extension of the
    // previous one, coarse levels carrying their nearest seeds' masses
    // (VGC_OT_CONVEX_LEVELS) instead of copied weights
    void set_convex_levels(bool x) { convex_levels_ = x; }
    void optimize(index_t max_iter) {
        vgc_check(vgc_ot_solve(vgc_context(), points_.data(), nu_.data(), weights_.data(),
                               index_t(nu_.size()), box_, eps_, max_iter, 0, &stats_));
    }
    void optimize_levels(const vector<index_t>& levels, index_t max_iter) {
        std::vector<uint32_t> lv(levels.begin(), levels.end());
        std::vector<vgc_ot_stats> st(lv.size() > 1 ? lv.size() - 1 : 1);
        vgc_check(vgc_ot_solve_levels(vgc_context(), points_.data(), nu_.data(), weights_.data(),
                                      index_t(nu_.size()), box_, lv.data(),
                                      index_t(lv.size() - 1), eps_, max_iter,
                                      convex_levels_ ? VGC_OT_CONVEX_LEVELS : 0u, st.data()));
        stats_ = st.back();
    }
    double weight(index_t i) const { return weights_[i]; }
    const vgc_ot_stats& stats() const { return stats_; }

private:
    double box_[6];
    double eps_ = 1e-3;
    bool convex_levels_ = false;
    std::vector<double> points_, nu_, weights_;
    vgc_ot_stats stats_{};
};

// Called once by the application (after GEO::initialize()): makes "VGC"
// selectable with algo:nn_search=VGC.
void vgc_register() {
    geo_register_NearestNeighborSearch_creator(VgcNearestNeighborSearch, "VGC");
}

}  // namespace GEO

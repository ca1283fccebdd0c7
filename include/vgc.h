/*
 * vgc — MI355X-native meshless Voronoi / Laguerre cells (C ABI).
 *
 * Drop-in boundary for geogram's meshless cell path.  Every entry point takes
 * plain pointers and sizes, never throws, returns VGC_OK (0), a positive
 * warning (VGC_PARTIAL) or a negative error, and leaves a message readable
 * with vgc_last_error().  Output buffers are owned and sized by the caller, as
 * in geogram (CentroidalVoronoiTesselation::Lloyd_iterations hands
 * RestrictedVoronoiDiagram::compute_centroids caller-allocated mg / m arrays,
 * voronoi/CVT.cpp:146-152).  With VGC_DEVICE_PTRS the pointers are device
 * memory of the context's GPU and nothing crosses PCIe.
 *
 * Streams: a context runs its kernels on a stream of its own.  Every call
 * first orders that stream after the work queued before the call on the
 * caller's stream (vgc_set_stream; default: the null stream), so device
 * inputs written by the caller's asynchronous kernels are complete, and
 * returns only when its own work is done, so device outputs are ready for
 * any stream.  A context is used by one host thread at a time.
 *
 * Numbering: seeds are indexed in the caller's order; neighbour ids too.
 * Coordinates are FP64 x,y,z triples (the layout of
 * CentroidalVoronoiTesselation::embedding, voronoi/CVT.h:333).
 */
#ifndef VGC_H
#define VGC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VGC_OK 0
#define VGC_PARTIAL 1        /* some cells are flagged, see vgc_stats */
#define VGC_ERR_ARG (-1)
#define VGC_ERR_DEVICE (-2)
#define VGC_ERR_CAPACITY (-3)
#define VGC_ERR_INTERNAL (-4)

#define VGC_DEVICE_PTRS 1u   /* all array arguments are device pointers */
#define VGC_NO_NEIGHBORS 2u  /* skip the neighbour CSR (volumes/centroids only) */
#define VGC_COUNT_WORK 4u    /* fill the algorithmic work counters of vgc_stats
                                (global atomics: slows the first pass down) */
#define VGC_OT_JUDGE_START 8u /* vgc_ot_solve: also measure the zero start and keep
                                 the start with the smaller mass residual */
#define VGC_OT_VERBOSE 16u    /* vgc_ot_solve: one progress line per Newton step
                                 on stderr */
#define VGC_AUTOTUNE 32u      /* time the first clipping pass with 1 and 4 wavefronts
                                 per work-group on the first call for a problem size
                                 and keep the faster (results do not depend on it) */
#define VGC_OT_INEXACT 64u    /* transport: inexact Newton — once the previous step
                                 was a full step (t = 1), the next step's CG stops
                                 at relative residual min(0.01, 0.01 x the current
                                 max relative mass error), never below the default
                                 1e-7; damped steps keep 1e-7 (Eisenstat-Walker
                                 forcing) */
#define VGC_OT_REPAIR 512u    /* transport: a start that leaves cells empty is
                                 repaired instead of dropped — each empty cell's
                                 weight is raised to the smallest value that puts
                                 its seed inside its cell with a margin of
                                 h^2 / 100 (h = (box volume / n)^(1/3)), for at most
                                 16 rounds (a raised weight can empty a
                                 neighbour); w = 0 only if cells stay empty.  No
                                 reference counterpart (the learned start). */
#define VGC_PERIODIC 256u     /* periodic domain: the box is one period of a
                                 3-torus (seeds inside it); cells, neighbour ids
                                 and facet areas are those of the periodic
                                 Voronoi / Laguerre diagram, a neighbour being
                                 named by its seed (an image's seed).  Replaces
                                 PeriodicDelaunay3d(periodic = true)
                                 (delaunay/periodic_delaunay_3d.cpp:4017-4320)
                                 for vgc_cells, vgc_lloyd (centroids wrapped
                                 into the box), vgc_cvt_func_grad,
                                 vgc_cvt_newton and vgc_ot_solve; one
                                 instance */
#define VGC_WAVES4 128u       /* run the first clipping pass with 4-wavefront
                                 work-groups instead of the default 1 (identical
                                 results; for tests and tuning) */
#define VGC_FP64_FIRST 1024u  /* run the first clipping pass with FP64 corner
                                 state.  Default: FP32 corner state, signs
                                 certified against the FP64 planes, and the
                                 cells it cannot certify redone in FP64
                                 (vgc_stats.nb_fp64_redo); identical results */
#define VGC_OT_CONVEX_LEVELS 2048u /* vgc_ot_solve_levels: each coarse level's
                                 seeds carry the masses of the seeds nearest
                                 to them (not their own, rescaled), and each
                                 level starts from the convex extension of the
                                 previous level's plan (vgc_ot_convex_start's
                                 soft maximum, no empty cell) instead of
                                 copied weights */

typedef struct vgc_ctx vgc_ctx;

typedef struct vgc_stats {
    uint64_t nb_cells;        /* cells computed                                   */
    uint64_t nb_neighbors;    /* total CSR entries                                */
    uint64_t nb_pass_b;       /* cells re-clipped with a complete radius list     */
    uint64_t nb_pass_c;       /* cells re-clipped with 128 / 256 corner slots     */
    uint64_t nb_pass_e;       /* cells re-clipped with exact signs (ties)         */
    uint64_t nb_uncertain;    /* cells left with a sign not decided exactly       */
    uint64_t nb_failed;       /* cells without valid output                       */
    uint64_t nb_empty;        /* empty Laguerre cells                             */
    uint64_t nb_disconnected; /* cells left unclipped by the coincident-seed rule */
    uint64_t nb_clips;        /* planes that cut a cell (VGC_COUNT_WORK)          */
    uint64_t nb_tests;        /* planes tested against a cell (VGC_COUNT_WORK)    */
    uint64_t radius_entries;  /* seeds gathered by the radius pass                */
    double t_build_ms, t_knn_ms, t_clip_ms, t_pass_bc_ms, t_compact_ms, t_total_ms;
    /* first clipping pass (one launch over every cell): GPU time from events
       on the launch stream, and (with VGC_COUNT_WORK) its algorithmic work:
       planes tested, live corners summed over the tests, corners created,
       integration tetrahedra and facets */
    double t_clip_a_gpu_ms, t_knn_gpu_ms;
    uint64_t a_cells, a_tests, a_live, a_new, a_tets, a_facets;
    uint64_t nb_fail_capacity; /* failed cells that exceed 256 corners / planes or
                                  the facet staging stride                          */
    /* first pass cells handed to the list tiers, by cause: ring limit of the
       weighted search, a tree walk beyond the wave's buffers, a ring that
       could not be narrowed */
    uint64_t why_rings, why_walk, why_narrow;
    uint64_t nb_regather; /* weighted cells gathered again without the ring limit */
    uint64_t tuned_waves; /* VGC_AUTOTUNE: wavefronts per work-group chosen (0: none) */
    uint64_t nb_periodic_retries; /* VGC_PERIODIC: passes redone with a wider image
                                     margin */
    uint64_t nb_pieces;   /* cells over 256 facets computed as eight octant pieces
                             (ConvexCell grows without bound,
                             voronoi/convex_cell.cpp:1086-1101) */
    uint64_t nb_fp64_redo; /* cells of the FP32 first pass redone by the FP64
                              first pass (a sign or corner not certified) */
    double t_redo_gpu_ms;  /* GPU time of that FP64 launch */
} vgc_stats;

/* Context on one GPU (device index as seen by HIP).  Replaces nothing in
   geogram: it owns the device workspace reused across calls. */
int vgc_create(vgc_ctx** ctx, int device);
void vgc_destroy(vgc_ctx* ctx);
const char* vgc_last_error(const vgc_ctx* ctx);
/* The stream (a hipStream_t) whose earlier work each later call waits for;
   NULL: the null stream.  Replaces nothing in geogram (its calls are
   synchronous CPU code). */
int vgc_set_stream(vgc_ctx* ctx, void* stream);

/* k nearest neighbours of every point (k <= n - 1), exact: FP64 squared
   distances, ascending, equal distances by ascending index; caller's
   indices.  The neighbours of point i are those of
   Delaunay_NearestNeighbors::get_neighbors_internal
   (delaunay/delaunay_nn.cpp:105-143): the k + 1 nearest points of i (i
   itself included) without i; if one of them coincides with i and has a
   smaller index, i gets no neighbours; coincident points of larger index
   are skipped.  Rows that end early are padded with nbr = -1, d2 = -1.
   nbr and d2 are n*k (d2 may be NULL).  Replaces
   NearestNeighborSearch::get_nearest_neighbors(nb, index_t, ...)
   (points/nn_search.h:180-185, points/kd_tree.cpp:156-241) called for every
   point, as Delaunay_NearestNeighbors::store_neighbors_CB does
   (delaunay/delaunay_nn.cpp:73-84).  Ties: geogram's kd-tree puts a point
   met later before an equally distant one (points/kd_tree.h:176-190), an
   order that depends on its tree; here the smaller index comes first. */
int vgc_knn(vgc_ctx* ctx, const double* xyz, uint32_t n, uint32_t k, uint32_t flags,
            int32_t* nbr, double* d2);

/* k nearest of the n points xyz (k <= n) to each of nq query points q
   (FP64 x,y,z triples; no point excluded, a point at the query itself comes
   first at distance 0): nbr / d2 are nq*k, ascending (FP64 squared
   distance, index).  Replaces NearestNeighborSearch::set_points +
   get_nearest_neighbors(nb_neighbors, query_point, neighbors,
   neighbors_sq_dist) (points/nn_search.h:96,129-134) called for each query
   point. */
int vgc_knn_search(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* q, uint32_t nq,
                   uint32_t k, uint32_t flags, int32_t* nbr, double* d2);

/* Voronoi (w == NULL) or Laguerre cells (power distance |x-p|^2 - w) of n
   seeds clipped to box = {xmin,ymin,zmin,xmax,ymax,zmax}.
   vol[n]: cell volumes; mg[3n]: integral of x over each cell (so mg/vol is
   the centroid, the quantities RestrictedVoronoiDiagram::compute_centroids
   accumulates, voronoi/RVD.cpp:429-528, and VBW::ConvexCell::volume /
   compute_mg return, voronoi/convex_cell.cpp:1215-1307);
   nbr_ptr[n+1], nbr[], farea[]: CSR of contributing neighbours (the
   v_global_index of contributing planes, convex_cell.h:728-741,830) and their
   facet areas (ConvexCell::facet_area, convex_cell.cpp:1160-1191).
   nbr / farea may be NULL; if cap is too small the call returns
   VGC_ERR_CAPACITY with the required size in stats->nb_neighbors.
   k0: 0 (default) lets each cell's wavefront gather its own neighbours,
   nearest first, ring after ring of distance, until the security radius is
   reached; 16, 24, 32, 48 or 64 first computes kNN lists of that size. */
int vgc_cells(vgc_ctx* ctx, const double* xyz, const double* w, uint32_t n,
              const double* box, uint32_t k0, uint32_t flags, double* vol, double* mg,
              uint64_t* nbr_ptr, int32_t* nbr, double* farea, uint64_t cap,
              vgc_stats* stats);

/* Same for n_inst independent instances in one launch: seeds of instance i
   are the inst_n[i] consecutive triples after those of instance i-1, boxes
   holds 6 doubles per instance.  Neighbour ids are global (caller's) indices. */
int vgc_cells_batch(vgc_ctx* ctx, const double* xyz, const double* w,
                    const uint32_t* inst_n, uint32_t n_inst, const double* boxes,
                    uint32_t k0, uint32_t flags, double* vol, double* mg,
                    uint64_t* nbr_ptr, int32_t* nbr, double* farea, uint64_t cap,
                    vgc_stats* stats);

/* CVT energy and its gradient over the Voronoi cells of xyz in the box,
   accumulated like RestrictedVoronoiDiagram::compute_CVT_func_grad_in_volume
   (voronoi/RVD.h:318, RVD.cpp:792-899; the caller zeroes f and g):
   f += sum_i integral over cell i of |x - p_i|^2, g[3i..3i+2] += 2 (m_i p_i -
   mg_i).  VGC_DEVICE_PTRS: xyz and g are device pointers (f stays on the
   host). */
int vgc_cvt_func_grad(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* box,
                      uint32_t flags, double* f, double* g, vgc_stats* stats);

/* iters Lloyd iterations in place with the security radius on: each seed
   moves to the centroid of its cell when the cell mass exceeds 1e-30.
   Replaces CentroidalVoronoiTesselation::Lloyd_iterations
   (voronoi/CVT.cpp:133-167), which would run with the security radius off. */
int vgc_lloyd(vgc_ctx* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
              uint32_t k0, uint32_t flags, vgc_stats* stats);

typedef struct vgc_mesh_sizes {
    uint64_t nb_vertices;        /* cell corners, all cells                     */
    uint64_t nb_facets;          /* polygons, all cells                         */
    uint64_t nb_facet_vertices;  /* entries of the polygons' vertex lists       */
    uint64_t nb_tets;            /* primal tetrahedra                           */
} vgc_mesh_sizes;

/* The cells as polyhedra, one per seed: VBW::ConvexCell::append_to_mesh /
   save (voronoi/convex_cell.cpp:282-358, 371-444), as compute_RVD's cube
   mode writes them (examples/geogram/compute_RVD/main.cpp:450-477).
   vertices[3 * nb_vertices]: the corners in absolute coordinates, pulled
   towards the cell's centroid by `shrink` (0: none); cell i owns vertices
   cell_vptr[i] .. cell_vptr[i+1]-1 and facets cell_fptr[i] ..
   cell_fptr[i+1]-1 (both n+1 entries).  Facet f is the polygon
   facet_vertices[facet_ptr[f] .. facet_ptr[f+1]-1] (global vertex numbers,
   counter-clockwise seen from outside), cut by seed facet_nbr[f] (caller's
   index) or by the box (-1).  Optionally tets[4 * nb_tets] (may be NULL):
   the primal tetrahedra, as vgc_rdt.  Two-call protocol: with NULL
   outputs, or capacities `cap` below the sizes, only the sizes are
   returned in `need` (VGC_ERR_CAPACITY when outputs were given). */
int vgc_cell_mesh(vgc_ctx* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
                  double shrink, uint32_t flags, uint64_t* cell_vptr, uint64_t* cell_fptr,
                  double* vertices, uint64_t* facet_ptr, int32_t* facet_vertices,
                  int32_t* facet_nbr, int32_t* tets, const vgc_mesh_sizes* cap,
                  vgc_mesh_sizes* need, vgc_stats* stats);

/* Restricted Delaunay triangulation of a volume (a box here): one
   tetrahedron (i, j, k, l) per cell corner whose three planes are bisectors
   (or power planes, w != NULL) with seeds j, k, l, emitted by the cell of
   the smallest index and oriented so that orient_3d(i, j, k, l) >= 0.
   Replaces RestrictedVoronoiDiagram::compute_RDT in volumetric mode
   (voronoi/RVD.h:504-510, RVD.cpp:2302-2327, for_each_primal_tetrahedron,
   generic_RVD.h:1011-1055).  tets: 4 * cap entries; the count is returned
   in *nb_tets (VGC_ERR_CAPACITY if cap is too small). */
int vgc_rdt(vgc_ctx* ctx, const double* xyz, const double* w, uint32_t n, const double* box,
            uint32_t flags, int32_t* tets, uint64_t cap, uint64_t* nb_tets, vgc_stats* stats);

typedef struct vgc_opt_stats {
    uint32_t iterations;      /* line searches completed                        */
    uint32_t evaluations;     /* energy / gradient evaluations (cell passes)    */
    int32_t line_search_info; /* last line search: 1 converged, 2 interval below
                                 xtol, 3 20 evaluations, 4/5 step at its bound,
                                 6 rounding, -1 not a descent direction         */
    uint32_t failed_cells;    /* flagged cells over all evaluations            */
    double f0, f;             /* CVT energy at the start and at the end        */
    double gnorm0, gnorm;     /* gradient norms                                */
    double t_total_ms;
} vgc_opt_stats;

/* Quasi-Newton CVT: minimises the CVT energy of the Voronoi cells of xyz in
   the box by L-BFGS with m correction pairs (m = 7 in geogram) and a
   More-Thuente line search, xyz updated in place; the iteration stops when
   a line search fails or after iters + 1 line searches (HLBFGS counts
   that way).  Replaces CentroidalVoronoiTesselation::Newton_iterations
   (voronoi/CVT.cpp:272-302, HLBFGS of third_party/HLBFGS with epsg = epsf =
   epsx = 0) with the energy and gradient of vgc_cvt_func_grad. */
int vgc_cvt_newton(vgc_ctx* ctx, double* xyz, uint32_t n, const double* box, uint32_t iters,
                   uint32_t m, uint32_t flags, vgc_opt_stats* stats);

/* Voronoi cells of the n seeds restricted to a triangulated surface
   (vertices vtx[3 nv], triangles tri[3 nt] of vertex indices, vw[nv] a
   density per vertex, linear on each triangle, or NULL for density 1): m[n]
   is the mass (area, or integral of the density) of each seed's part of the
   surface, mg[3 n] the integral of x over it (mg / m is its centroid); both
   are overwritten.  Replaces RestrictedVoronoiDiagram::
   compute_centroids_on_surface (voronoi/RVD.cpp:371-415; polygons of
   voronoi/generic_RVD.h:1319-1420 and 2133-2190, Geom::triangle_centroid of
   basic/geometry_nd.h:178-197 with a vertex "weight" attribute), which
   accumulates into caller-zeroed arrays.  FP64 throughout; each seed's
   polygons are summed in (triangle, split) order, so a call is
   reproducible bit for bit.  flags: VGC_DEVICE_PTRS. */
typedef struct vgc_surface_stats {
    uint64_t nb_items;   /* triangles processed, splits and retries included */
    uint64_t nb_split;   /* items split in four (polygon or seed list full)  */
    uint64_t nb_redo;    /* items redone with longer neighbour lists         */
    uint64_t nb_records; /* restricted polygons integrated (retried items'
                            first attempts included)                         */
    uint32_t k_final;    /* neighbours per seed in the last lists            */
    uint32_t pad;
    double t_ms;         /* wall time of the call                            */
} vgc_surface_stats;
int vgc_surface_centroids(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* vtx,
                          uint32_t nv, const int32_t* tri, uint32_t nt, const double* vw,
                          uint32_t flags, double* mg, double* m, vgc_surface_stats* stats);

/* Lloyd iterations of n_inst independent CVT problems at once (instance
   layout of vgc_cells_batch): each iteration clips every instance's cells in
   the same launches.  No counterpart in geogram, which runs one
   CentroidalVoronoiTesselation at a time. */
int vgc_lloyd_batch(vgc_ctx* ctx, double* xyz, const uint32_t* inst_n, uint32_t n_inst,
                    const double* boxes, uint32_t iters, uint32_t k0, uint32_t flags,
                    vgc_stats* stats);

typedef struct vgc_ot_stats {
    uint32_t newton_iters;        /* accepted Newton steps                       */
    uint32_t cg_iters;            /* conjugate-gradient iterations, all steps    */
    uint32_t cell_passes;         /* Laguerre cell evaluations (line search too) */
    uint32_t halvings;            /* step halvings of the damped Newton rule     */
    uint32_t used_fallback_start; /* the given start had an empty cell: w = 0    */
    uint32_t converged;
    uint32_t repair_rounds;       /* VGC_OT_REPAIR: rounds that raised the weights
                                     of empty cells of the start                 */
    uint32_t pad;
    double err0, final_err;       /* max_i |m_i - nu_i| / nu_i, first and last   */
    double t_total_ms, t_cells_ms, t_cg_ms;
} vgc_ot_stats;

/* Semi-discrete optimal transport from the uniform density on box to the
   masses nu[n] at the seeds (nu must sum to the box volume): finds Laguerre
   weights w with cell masses m_i(w) = nu_i up to
   max_i |m_i - nu_i| / nu_i <= tol, by damped Newton (Kitagawa-Merigot-
   Thibert step rule) with a Jacobi-preconditioned CG on the Laguerre
   Hessian.  w is the start on input (zeros: plain start; a start with an
   empty cell falls back to zeros) and the solution on output.
   Replaces OptimalTransportMap3d::set_points / set_epsilon / optimize
   (examples/exploragram/compute_OTM/main.cpp:283-297; the solver itself is
   Exploragram's, outside the geogram tree).  max_newton 0 = 100.
   Returns VGC_PARTIAL if not converged. */
int vgc_ot_solve(vgc_ctx* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                 const double* box, double tol, uint32_t max_newton, uint32_t flags,
                 vgc_ot_stats* stats);

/* n_inst independent transport problems at once (instance layout of
   vgc_cells_batch; nu of each instance sums to its box volume; w in: starts,
   out: solutions).  One cell pass per step covers every instance; each
   instance has its own damped-Newton state and PCG scalars.  stats: n_inst
   entries or NULL.  No counterpart in geogram (one OptimalTransportMap3d at a
   time). */
int vgc_ot_solve_batch(vgc_ctx* ctx, const double* xyz, const double* nu, double* w,
                       const uint32_t* inst_n, uint32_t n_inst, const double* boxes, double tol,
                       uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats);

/* Multilevel transport, OptimalTransportMap3d::optimize_levels
   (compute_OTM/main.cpp:294): the seeds come in levels (a random order, such
   as geogram's BRIO order, cut at levels[0] = 0 < levels[1] < ... <
   levels[n_levels] = n).  Level k solves the transport of the first
   levels[k+1] seeds (their masses rescaled to the box volume) from the
   previous level's weights, each new seed starting with the weight of its
   power-nearest seed of the previous level (VGC_OT_CONVEX_LEVELS: every seed
   of the level from the convex extension of the previous plan).  stats:
   n_levels entries (may be null).  Other arguments as vgc_ot_solve; w is
   output only. */
int vgc_ot_solve_levels(vgc_ctx* ctx, const double* xyz, const double* nu, double* w, uint32_t n,
                        const double* box, const uint32_t* levels, uint32_t n_levels, double tol,
                        uint32_t max_newton, uint32_t flags, vgc_ot_stats* stats);

/* ---- learned warm start of vgc_ot_solve (no counterpart in geogram) ----
   MLP 6 -> 64 -> 64 -> 1, leaky ReLU, FP32; VGC_MLP_PARAMS floats laid out
   W1[6][64], b1[64], W2[64][64], b2[64], W3[64], b3.  Features of a seed:
   position in the box (3), log(nu_i n / V), log of the mean and of the max
   distance to its 16 nearest neighbours over h = (V/n)^(1/3).  The network
   predicts w_i - mean(w) (box units). */
#define VGC_MLP_PARAMS 4673
#define VGC_MLP_FEATURES 6
int vgc_ot_features(vgc_ctx* ctx, const double* xyz, const double* nu, uint32_t n,
                    const double* box, uint32_t flags, float* feats);
int vgc_mlp_forward(vgc_ctx* ctx, const float* params, const float* feats, uint32_t n,
                    uint32_t flags, float* out);
/* Newton start from a coarse transport plan (geogram's counterpart is the
   multilevel start of OptimalTransportMap3d, compute_OTM/main.cpp:293-294):
   w_i = |p_i|^2 - s log sum_j exp(piece_j(p_i) / s) with
   piece_j(p) = a_j + g_j.p + mu_j H(|p - b_j|^2), H(r2) = r2 up to R_j^2 and
   R_j (2 sqrt(r2) - R_j) beyond; VGC_START_PIECE doubles per piece
   (a, g[3], mu, b[3], R), m pieces, s > 0.  The function is convex, so no
   cell of the start is empty.  flags: VGC_DEVICE_PTRS. */
#define VGC_START_PIECE 9
int vgc_ot_convex_start(vgc_ctx* ctx, const double* xyz, uint32_t n, const double* pieces,
                        uint32_t m, double s, uint32_t flags, double* w);
/* gradient of the mean squared error over the n rows; loss = that mean */
int vgc_mlp_grad(vgc_ctx* ctx, const float* params, const float* feats, const float* target,
                 uint32_t n, uint32_t flags, float* grad, double* loss);
/* vgc_mlp_grad without any allocation, for callers that cannot allocate
   device memory (code under HIP graph capture, training loops that forbid
   allocation).  No counterpart in geogram.  Device pointers only; loss is
   one device double (may be NULL), grad VGC_MLP_PARAMS floats; results equal
   vgc_mlp_grad's bit for bit.  n <= 32768 rows (VGC_ERR_CAPACITY beyond):
   the per-wave partials live in a static device array of 19.2 MB that each
   GPU holds for the life of the process and that every call on that GPU
   shares — calls that overlap on one GPU (several contexts or threads) race. */
int vgc_mlp_grad_noalloc(vgc_ctx* ctx, const float* params, const float* feats,
                         const float* target, uint32_t n, float* grad, double* loss);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8); device pointers only */
int vgc_mlp_adam(vgc_ctx* ctx, float* params, const float* grad, float* m, float* v,
                 uint32_t step, float lr, uint32_t flags);

/* count doubles of the stream std::mt19937_64(seed) +
   uniform_real_distribution<double>(0,1), the generator behind
   Numeric::random_reset(seed) / random_float64() (basic/numeric.cpp:76-103). */
int vgc_random_float64(uint64_t seed, double* out, uint64_t count);

/* Synthetic seeds: kind 0 = n uniform in [0,1]^3; kind 1 = n/2 uniform then
   n - n/2 from an 8-component isotropic Gaussian mixture (means U[0.2,0.8]^3,
   sigma U[0.01,0.1], rejection to the cube); kind 2 = 16 components,
   sigma U[0.02,0.15].  All draws from vgc_random_float64(seed). */
int vgc_sample_points(uint32_t kind, uint64_t seed, uint32_t n, double* xyz);

#ifdef __cplusplus
}
#endif
#endif

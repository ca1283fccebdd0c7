/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement, in plain C, of geogram's meshless Voronoi / Laguerre cell
 * path, used to check the HIP implementation.  Only tests, the smoke check and
 * the CPU-baseline leg of the benchmark may load it; the product never does.
 *
 * What it restates (reference files, relative to the geogram source tree):
 *  - kNN with self removal and the duplicate-point rule of
 *    Delaunay_NearestNeighbors::get_neighbors_internal
 *    (delaunay/delaunay_nn.cpp:105-143), over an exact FP64 kd-tree
 *    (points/kd_tree.cpp:156-387; ties ordered by index here);
 *  - VBW::ConvexCell in dual form (voronoi/convex_cell.h:419-1436):
 *    init_with_box (convex_cell.cpp:175-215), clip_by_plane with the exact
 *    conflict test det_4d >= 0 (convex_cell.cpp:461-516, 870-911),
 *    triangulate_conflict_zone (convex_cell.cpp:761-865), Cramer corners
 *    (convex_cell.cpp:971-1082), volume / compute_mg / facet_area fans
 *    (convex_cell.cpp:1160-1307), squared_radius (convex_cell.cpp:1312-1330);
 *  - the security-radius loop of RestrictedVoronoiDiagram::clip_by_cell_SR
 *    (voronoi/generic_RVD.h:2278-2347): stop when d(pi,pj)^2 > 4.1 R^2, grow
 *    k by k/8 (k > 8) or by 1;
 *  - the Laguerre plane of PeriodicDelaunay3d::copy_Laguerre_cell_facet_from_
 *    Delaunay (delaunay/periodic_delaunay_3d.cpp:3719-3726);
 *  - the Lloyd update of CentroidalVoronoiTesselation::Lloyd_iterations
 *    (voronoi/CVT.cpp:148-162), with the security radius on;
 *  - the Voronoi cells restricted to a triangulated surface and their
 *    masses / first moments (RestrictedVoronoiDiagram::
 *    compute_centroids_on_surface, voronoi/RVD.cpp:371-415; traversal of
 *    voronoi/generic_RVD.h:1319-1420, polygon clipping of
 *    voronoi/generic_RVD_polygon.h:242-366).
 */
#ifndef VGC_ORACLE_H
#define VGC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    long long nb_clip;          /* clip_by_plane calls                      */
    long long nb_knn_grow;      /* neighbourhood enlargements               */
    long long nb_disconnected;  /* cells left unclipped (duplicate rule)    */
    long long nb_empty;         /* empty (Laguerre) cells                   */
    long long nb_bad_border;    /* conflict zones that were not a disk      */
} orc_stats;

/* Exact FP64 k nearest neighbours of every point, self excluded, sorted by
   (squared distance, index).  idx / d2 are n*k arrays.  Returns 0 or <0. */
int orc_knn(const double* xyz, int n, int k, int* idx, double* d2);

/* Same, for arbitrary query points (self is not excluded). */
int orc_knn_query(const double* xyz, int n, const double* q, int nq, int k,
                  int* idx, double* d2);

/* Voronoi (w == NULL) or Laguerre (power distance |x-p|^2 - w) cells of every
   seed, clipped to box = {xmin,ymin,zmin,xmax,ymax,zmax}.
   mode 1: security radius with k0 initial neighbours, growing as geogram;
   mode 0: exactly the k0 nearest neighbours (no security radius);
   mode 2: all n-1 other seeds (brute force).
   Outputs: vol[n]; mg[3n] = sum over the cell of x dV (so mg/vol is the
   centroid); nbr_cnt[n]; nbr / farea [n*stride], contributing neighbours in
   clipping order and their facet areas.  Returns 0, or -1 if a cell has more
   than stride neighbours, -2 on allocation failure. */
int orc_cells(const double* xyz, const double* w, int n, const double* box,
              int k0, int mode, double* vol, double* mg, int* nbr_cnt,
              int* nbr, double* farea, int stride, orc_stats* st);

/* Cells of a subset of seeds (cell ids in `cells`, count nc): same outputs,
   indexed 0..nc-1.  Used for bounded CPU samples. */
int orc_cells_subset(const double* xyz, const double* w, int n,
                     const double* box, int k0, int mode, const int* cells,
                     int nc, double* vol, double* mg, int* nbr_cnt, int* nbr,
                     double* farea, int stride, orc_stats* st);

/* iters Lloyd iterations in place, security radius on (mode 1, k0). */
int orc_lloyd(double* xyz, int n, const double* box, int iters, int k0,
              orc_stats* st);

/* 1: form every cell's planes in the seed-centred frame (coordinates
   translated by -pi first), 0 (default): absolute coordinates, exactly as
   copy_Laguerre_cell_facet_from_Delaunay does.  Same algorithm; the centred
   frame avoids the |pi|^2 rounding of the plane offsets on tiny cells. */
void orc_set_local_planes(int on);
/* per-cell integral of |x - p_i|^2 of the next orc_cells* calls (NULL: off) */
void orc_set_e2_output(double* e2);

/* Cell polyhedra of the next orc_cells* calls (NULL: off): per cell c,
   ibuf[c * istride ...] = nv, nf, nt, then nf faces (neighbour or -1, size,
   size vertex numbers of the cell), then nt tetrahedra (4 seed indices);
   vbuf[c * vstride ...] = 3 nv vertex coordinates.  nv = -1 when a stride
   is too small. */
void orc_set_mesh_output(int* ibuf, int istride, double* vbuf, int vstride);

/* Threads used by the OpenMP loops (0 = runtime default). */
void orc_set_threads(int nt);
int orc_get_threads(void);

/* Voronoi cells of the n seeds restricted to the triangle mesh (V: nv
   vertices, F: nf triangles, vw: vertex weights = a linear density, or NULL
   for density 1): m[n] = mass (area) of each restricted cell, mg[3n] = its
   integral of x.  mode 0: the reference's traversal with the security
   radius; mode 1: every facet clipped by all bisectors of every seed (brute
   force).  npoly: restricted polygons with 3 or more vertices.  Returns 0
   or <0. */
int orc_surface_centroids(const double* xyz, int n, const double* V, int nv, const int* F, int nf,
                          const double* vw, int mode, double* mg, double* m, long long* npoly);

#ifdef __cplusplus
}
#endif
#endif

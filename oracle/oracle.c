/*
 * ORACLE — test infrastructure only (see oracle.h for what it restates and
 * where).  Plain C, FP64, exact conflict predicates; compiled with
 * -frounding-math -ffp-contract=off like the reference
 * (cmake/platforms/Linux-gcc.cmake).
 */
#include "oracle.h"
#include "exact.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;
static int g_local = 0;
void orc_set_local_planes(int on) { g_local = on; }
/* optional per-cell output of the next orc_cells* calls: integral of
   |x - p_i|^2 over cell i (NULL: off) */
static double* g_e2 = NULL;
void orc_set_e2_output(double* e2) { g_e2 = e2; }
/* cell meshes of the next orc_cells* calls (NULL: off), see oracle.h */
static int* g_mi = NULL;
static int g_mi_stride = 0;
static double* g_mv = NULL;
static int g_mv_stride = 0;
void orc_set_mesh_output(int* ibuf, int istride, double* vbuf, int vstride) {
    g_mi = ibuf;
    g_mi_stride = istride;
    g_mv = vbuf;
    g_mv_stride = vstride;
}
void orc_set_threads(int nt) { g_threads = nt; }
int orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================
 * kd-tree, exact FP64 (the role of BalancedKdTree, points/kd_tree.cpp:
 * median split on the widest axis, leaves of at most 16 points).
 * ====================================================================== */

typedef struct {
    int begin, end, left, right;
    double lo[3], hi[3];
} kd_node;

typedef struct {
    const double* xyz;
    int n;
    int* perm;
    kd_node* nodes;
    int nn;
} kd_tree;

static int kd_less(const double* xyz, int axis, int p, int q) {
    double a = xyz[3 * p + axis], b = xyz[3 * q + axis];
    return (a < b) || (a == b && p < q);
}

/* nth_element on perm[lo..hi] (inclusive) so that perm[k] is in place. */
static void kd_select(int* a, int lo, int hi, int k, const double* xyz,
                      int axis) {
    while (hi > lo) {
        int pv = a[lo + (hi - lo) / 2];
        int i = lo, j = hi;
        while (i <= j) {
            while (kd_less(xyz, axis, a[i], pv)) ++i;
            while (kd_less(xyz, axis, pv, a[j])) --j;
            if (i <= j) {
                int t = a[i];
                a[i] = a[j];
                a[j] = t;
                ++i;
                --j;
            }
        }
        if (k <= j)
            hi = j;
        else if (k >= i)
            lo = i;
        else
            return;
    }
}

static int kd_build_rec(kd_tree* T, int b, int e) {
    int id = T->nn++;
    kd_node* N = &T->nodes[id];
    N->begin = b;
    N->end = e;
    N->left = N->right = -1;
    for (int c = 0; c < 3; ++c) {
        N->lo[c] = DBL_MAX;
        N->hi[c] = -DBL_MAX;
    }
    for (int i = b; i < e; ++i) {
        const double* p = T->xyz + 3 * T->perm[i];
        for (int c = 0; c < 3; ++c) {
            if (p[c] < N->lo[c]) N->lo[c] = p[c];
            if (p[c] > N->hi[c]) N->hi[c] = p[c];
        }
    }
    if (e - b <= 16) return id;
    int axis = 0;
    double ext = N->hi[0] - N->lo[0];
    for (int c = 1; c < 3; ++c) {
        if (N->hi[c] - N->lo[c] > ext) {
            ext = N->hi[c] - N->lo[c];
            axis = c;
        }
    }
    int m = b + (e - b) / 2;
    kd_select(T->perm, b, e - 1, m, T->xyz, axis);
    int l = kd_build_rec(T, b, m);
    int r = kd_build_rec(T, m, e);
    T->nodes[id].left = l;
    T->nodes[id].right = r;
    return id;
}

static int kd_build(kd_tree* T, const double* xyz, int n) {
    T->xyz = xyz;
    T->n = n;
    T->perm = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    T->nodes = (kd_node*)malloc(sizeof(kd_node) * (size_t)(2 * n + 2));
    T->nn = 0;
    if (!T->perm || !T->nodes) return -2;
    for (int i = 0; i < n; ++i) T->perm[i] = i;
    if (n > 0) kd_build_rec(T, 0, n);
    return 0;
}

static void kd_free(kd_tree* T) {
    free(T->perm);
    free(T->nodes);
}

typedef struct {
    int k, cnt;
    int* idx;
    double* d2;
} knn_list;

static void knn_insert(knn_list* H, int p, double d) {
    if (H->cnt == H->k) {
        double wd = H->d2[H->k - 1];
        int wi = H->idx[H->k - 1];
        if (!(d < wd || (d == wd && p < wi))) return;
    }
    int i = (H->cnt < H->k) ? H->cnt : H->k - 1;
    while (i > 0 && (H->d2[i - 1] > d || (H->d2[i - 1] == d && H->idx[i - 1] > p))) {
        H->d2[i] = H->d2[i - 1];
        H->idx[i] = H->idx[i - 1];
        --i;
    }
    H->d2[i] = d;
    H->idx[i] = p;
    if (H->cnt < H->k) H->cnt++;
}

static double dist2(const double* a, const double* b) {
    double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return dx * dx + dy * dy + dz * dz;
}

static double box_dist2(const kd_node* N, const double* q) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
        double d = 0.0;
        if (q[c] < N->lo[c])
            d = N->lo[c] - q[c];
        else if (q[c] > N->hi[c])
            d = q[c] - N->hi[c];
        s += d * d;
    }
    return s;
}

static void kd_query_rec(const kd_tree* T, int id, const double* q,
                         knn_list* H) {
    const kd_node* N = &T->nodes[id];
    if (H->cnt == H->k && box_dist2(N, q) > H->d2[H->k - 1]) return;
    if (N->left < 0) {
        for (int i = N->begin; i < N->end; ++i) {
            int p = T->perm[i];
            knn_insert(H, p, dist2(q, T->xyz + 3 * p));
        }
        return;
    }
    double dl = box_dist2(&T->nodes[N->left], q);
    double dr = box_dist2(&T->nodes[N->right], q);
    if (dl <= dr) {
        kd_query_rec(T, N->left, q, H);
        kd_query_rec(T, N->right, q, H);
    } else {
        kd_query_rec(T, N->right, q, H);
        kd_query_rec(T, N->left, q, H);
    }
}

/* the kq nearest points of q (self included if q is a data point) */
static int kd_query(const kd_tree* T, const double* q, int kq, int* idx,
                    double* d2) {
    knn_list H = {kq, 0, idx, d2};
    if (kq <= 0 || T->n == 0) return 0;
    kd_query_rec(T, 0, q, &H);
    return H.cnt;
}

int orc_knn_query(const double* xyz, int n, const double* q, int nq, int k,
                  int* idx, double* d2) {
    if (k < 1 || k > n) return -1;
    kd_tree T;
    if (kd_build(&T, xyz, n) != 0) {
        kd_free(&T);
        return -2;
    }
#pragma omp parallel for schedule(dynamic, 256) num_threads(orc_get_threads())
    for (int i = 0; i < nq; ++i) {
        kd_query(&T, q + 3 * (size_t)i, k, idx + (size_t)i * k,
                 d2 + (size_t)i * k);
    }
    kd_free(&T);
    return 0;
}

int orc_knn(const double* xyz, int n, int k, int* idx, double* d2) {
    if (k < 1 || k > n - 1) return -1;
    kd_tree T;
    if (kd_build(&T, xyz, n) != 0) {
        kd_free(&T);
        return -2;
    }
    int err = 0;
#pragma omp parallel num_threads(orc_get_threads())
    {
        int* ti = (int*)malloc(sizeof(int) * (size_t)(k + 1));
        double* td = (double*)malloc(sizeof(double) * (size_t)(k + 1));
        if (!ti || !td) {
#pragma omp atomic write
            err = -2;
        } else {
#pragma omp for schedule(dynamic, 256)
            for (int i = 0; i < n; ++i) {
                int got = kd_query(&T, xyz + 3 * (size_t)i, k + 1, ti, td);
                int c = 0;
                for (int m = 0; m < got && c < k; ++m) {
                    if (ti[m] == i) continue;
                    idx[(size_t)i * k + c] = ti[m];
                    d2[(size_t)i * k + c] = td[m];
                    ++c;
                }
            }
        }
        free(ti);
        free(td);
    }
    kd_free(&T);
    return err;
}

/* ======================================================================
 * Convex cell in dual form (VBW::ConvexCell semantics).
 * Vertices of the dual are planes P = (a,b,c,d), kept side a x+b y+c z+d >= 0.
 * Triangles (i,j,k) are the corners; A[t][m] is the triangle across the edge
 * opposite to local vertex m.  Valid corners have w = det3(normals) < 0.
 * Plane 0 is unused (the reference's vertex at infinity); planes 1..6 are
 * the box, as in init_with_box (convex_cell.cpp:175-215).
 * ====================================================================== */

typedef struct {
    double x, y, z, w;
} v4;

typedef struct {
    int nv, cap_v;
    v4* P;
    int* vg;
    int* bt;      /* border triangle, per dual vertex */
    int* bm;      /* border edge, per dual vertex */
    int* bstamp;  /* clip counter stamp of bt/bm */
    int* v2t;
    int nt, cap_t;
    int (*T)[3];
    int (*A)[3];
    char* alive;
    char* conf;
    int* fstack;
    int nfree;
    double (*pt)[3];
    int empty;
    int bad;
    int stamp;
} cell;

static int cell_reserve_v(cell* C, int need) {
    if (need <= C->cap_v) return 0;
    int cap = C->cap_v ? C->cap_v : 32;
    while (cap < need) cap *= 2;
    C->P = (v4*)realloc(C->P, sizeof(v4) * (size_t)cap);
    C->vg = (int*)realloc(C->vg, sizeof(int) * (size_t)cap);
    C->bt = (int*)realloc(C->bt, sizeof(int) * (size_t)cap);
    C->bm = (int*)realloc(C->bm, sizeof(int) * (size_t)cap);
    C->bstamp = (int*)realloc(C->bstamp, sizeof(int) * (size_t)cap);
    C->v2t = (int*)realloc(C->v2t, sizeof(int) * (size_t)cap);
    if (!C->P || !C->vg || !C->bt || !C->bm || !C->bstamp || !C->v2t)
        return -2;
    for (int i = C->cap_v; i < cap; ++i) C->bstamp[i] = -1;
    C->cap_v = cap;
    return 0;
}

static int cell_reserve_t(cell* C, int need) {
    if (need <= C->cap_t) return 0;
    int cap = C->cap_t ? C->cap_t : 64;
    while (cap < need) cap *= 2;
    C->T = (int(*)[3])realloc(C->T, sizeof(int[3]) * (size_t)cap);
    C->A = (int(*)[3])realloc(C->A, sizeof(int[3]) * (size_t)cap);
    C->alive = (char*)realloc(C->alive, (size_t)cap);
    C->conf = (char*)realloc(C->conf, (size_t)cap);
    C->fstack = (int*)realloc(C->fstack, sizeof(int) * (size_t)cap);
    C->pt = (double(*)[3])realloc(C->pt, sizeof(double[3]) * (size_t)cap);
    if (!C->T || !C->A || !C->alive || !C->conf || !C->fstack || !C->pt)
        return -2;
    C->cap_t = cap;
    return 0;
}

static void cell_free(cell* C) {
    free(C->P);
    free(C->vg);
    free(C->bt);
    free(C->bm);
    free(C->bstamp);
    free(C->v2t);
    free(C->T);
    free(C->A);
    free(C->alive);
    free(C->conf);
    free(C->fstack);
    free(C->pt);
    memset(C, 0, sizeof(*C));
}

static int cell_new_triangle(cell* C, int i, int j, int k) {
    int t;
    if (C->nfree > 0) {
        t = C->fstack[--C->nfree];
    } else {
        if (cell_reserve_t(C, C->nt + 1) != 0) return -1;
        t = C->nt++;
    }
    C->T[t][0] = i;
    C->T[t][1] = j;
    C->T[t][2] = k;
    C->A[t][0] = C->A[t][1] = C->A[t][2] = -1;
    C->alive[t] = 1;
    return t;
}

static int cell_add_plane(cell* C, v4 P, int gid) {
    if (cell_reserve_v(C, C->nv + 1) != 0) return -1;
    int v = C->nv++;
    C->P[v] = P;
    C->vg[v] = gid;
    return v;
}

static int local_index(const int* T, int v) {
    return (T[0] == v) ? 0 : ((T[1] == v) ? 1 : 2);
}

/* Adjacency of all alive triangles by matching reversed edges (the role of
   ConvexCell::connect_triangles, convex_cell.cpp:1348-1411). */
static void cell_connect(cell* C) {
    for (int t = 0; t < C->nt; ++t) {
        if (!C->alive[t]) continue;
        for (int m = 0; m < 3; ++m) {
            int a = C->T[t][(m + 1) % 3], b = C->T[t][(m + 2) % 3];
            C->A[t][m] = -1;
            for (int u = 0; u < C->nt && C->A[t][m] < 0; ++u) {
                if (u == t || !C->alive[u]) continue;
                for (int mm = 0; mm < 3; ++mm) {
                    if (C->T[u][(mm + 1) % 3] == b && C->T[u][(mm + 2) % 3] == a) {
                        C->A[t][m] = u;
                        break;
                    }
                }
            }
        }
    }
}

static void cell_init_box(cell* C, const double* box) {
    C->nv = 0;
    C->nt = 0;
    C->nfree = 0;
    C->empty = 0;
    C->bad = 0;
    v4 inf = {0, 0, 0, 0};
    cell_add_plane(C, inf, -1);
    v4 p1 = {1.0, 0.0, 0.0, -box[0]};
    v4 p2 = {-1.0, 0.0, 0.0, box[3]};
    v4 p3 = {0.0, 1.0, 0.0, -box[1]};
    v4 p4 = {0.0, -1.0, 0.0, box[4]};
    v4 p5 = {0.0, 0.0, 1.0, -box[2]};
    v4 p6 = {0.0, 0.0, -1.0, box[5]};
    cell_add_plane(C, p1, -1);
    cell_add_plane(C, p2, -1);
    cell_add_plane(C, p3, -1);
    cell_add_plane(C, p4, -1);
    cell_add_plane(C, p5, -1);
    cell_add_plane(C, p6, -1);
    /* 8 corners: planes (X,Y,Z), X in {1 (min), 2 (max)}, etc.  The normal
       of a min plane points +, of a max plane -, so det3(nX,nY,nZ) is the
       product of the three signs; order (X,Y,Z) when it is negative and
       (Y,X,Z) otherwise, which gives w < 0 for every corner. */
    for (int s = 0; s < 8; ++s) {
        int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
        int X = 1 + sx, Y = 3 + sy, Z = 5 + sz;
        int sign = (sx ? -1 : 1) * (sy ? -1 : 1) * (sz ? -1 : 1);
        if (sign < 0)
            cell_new_triangle(C, X, Y, Z);
        else
            cell_new_triangle(C, Y, X, Z);
    }
    cell_connect(C);
}

static double det3(double a11, double a12, double a13, double a21, double a22,
                   double a23, double a31, double a32, double a33) {
    return a11 * (a22 * a33 - a23 * a32) - a12 * (a21 * a33 - a23 * a31) +
           a13 * (a21 * a32 - a22 * a31);
}

/* Homogeneous corner of triangle t by Cramer's rule (convex_cell.cpp:
   1040-1082): the intersection of its three planes. */
static v4 cell_corner(const cell* C, int t) {
    const v4 a = C->P[C->T[t][0]], b = C->P[C->T[t][1]], c = C->P[C->T[t][2]];
    v4 r;
    r.x = -det3(a.w, a.y, a.z, b.w, b.y, b.z, c.w, c.y, c.z);
    r.y = -det3(a.x, a.w, a.z, b.x, b.w, b.z, c.x, c.w, c.z);
    r.z = -det3(a.x, a.y, a.w, b.x, b.y, b.w, c.x, c.y, c.w);
    r.w = det3(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z);
    return r;
}

static double cell_sq_radius(const cell* C, const double* center) {
    double r = 0.0;
    for (int t = 0; t < C->nt; ++t) {
        if (!C->alive[t]) continue;
        v4 p = cell_corner(C, t);
        double q[3] = {p.x / p.w, p.y / p.w, p.z / p.w};
        double d = dist2(center, q);
        if (d > r) r = d;
    }
    return r;
}

/* clip_by_plane + triangulate_conflict_zone (convex_cell.cpp:461-516,
   761-865): exact conflict test det4(Pi,Pj,Pk,P) >= 0. */
static int cell_clip(cell* C, v4 P, int gid) {
    int lv = cell_add_plane(C, P, gid);
    if (lv < 0) return -2;
    if (C->empty) return 0;
    int nconf = 0, nalive = 0;
    double pe[4] = {P.x, P.y, P.z, P.w};
    for (int t = 0; t < C->nt; ++t) {
        C->conf[t] = 0;
        if (!C->alive[t]) continue;
        ++nalive;
        const v4 a = C->P[C->T[t][0]], b = C->P[C->T[t][1]], c = C->P[C->T[t][2]];
        double ra[4] = {a.x, a.y, a.z, a.w}, rb[4] = {b.x, b.y, b.z, b.w},
               rc[4] = {c.x, c.y, c.z, c.w};
        if (orc_det4_sign(ra, rb, rc, pe) >= 0) {
            C->conf[t] = 1;
            ++nconf;
        }
    }
    if (nconf == 0) return 0;
    if (nconf == nalive) {
        for (int t = 0; t < C->nt; ++t) C->alive[t] = 0;
        C->nfree = 0;
        C->empty = 1;
        return 0;
    }
    C->stamp++;
    int first = -1, nborder = 0;
    for (int t = 0; t < C->nt; ++t) {
        if (!C->alive[t] || C->conf[t]) continue;
        for (int m = 0; m < 3; ++m) {
            int u = C->A[t][m];
            if (u >= 0 && C->conf[u]) {
                int v1 = C->T[t][(m + 1) % 3];
                C->bt[v1] = t;
                C->bm[v1] = m;
                C->bstamp[v1] = C->stamp;
                first = v1;
                ++nborder;
            }
        }
    }
    /* walk the border, one new triangle (lv, v2, v1) per border edge */
    int v = first, prev = -1, first_new = -1, steps = 0;
    do {
        if (C->bstamp[v] != C->stamp || steps > nborder) {
            C->bad = 1;
            return -3;
        }
        int t = C->bt[v], m = C->bm[v];
        int v1 = C->T[t][(m + 1) % 3], v2 = C->T[t][(m + 2) % 3];
        int nt = cell_new_triangle(C, lv, v2, v1);
        if (nt < 0) return -2;
        C->conf[nt] = 0;
        C->A[nt][0] = t;
        C->A[t][m] = nt;
        if (prev < 0) {
            first_new = nt;
        } else {
            C->A[prev][2] = nt;
            C->A[nt][1] = prev;
        }
        prev = nt;
        v = v2;
        ++steps;
    } while (v != first);
    if (steps != nborder) {
        C->bad = 1;
        return -3;
    }
    C->A[prev][2] = first_new;
    C->A[first_new][1] = prev;
    for (int t = 0; t < C->nt; ++t) {
        if (C->alive[t] && C->conf[t]) {
            C->alive[t] = 0;
            C->conf[t] = 0;
            if (cell_reserve_t(C, C->nfree + 1) != 0) return -2;
            C->fstack[C->nfree++] = t;
        }
    }
    return 0;
}

/* compute_geometry (convex_cell.cpp:1125-1145) */
static void cell_geometry(cell* C) {
    for (int v = 0; v < C->nv; ++v) C->v2t[v] = -1;
    for (int t = 0; t < C->nt; ++t) {
        if (!C->alive[t]) continue;
        v4 p = cell_corner(C, t);
        C->pt[t][0] = p.x / p.w;
        C->pt[t][1] = p.y / p.w;
        C->pt[t][2] = p.z / p.w;
        for (int k = 0; k < 3; ++k) {
            if (C->v2t[C->T[t][k]] < 0) C->v2t[C->T[t][k]] = t;
        }
    }
}

static int walk_next(const cell* C, int t, int v) {
    int lv = local_index(C->T[t], v);
    return C->A[t][(lv + 1) % 3];
}

static double tet_vol(const double* p1, const double* p2, const double* p3,
                      const double* p4) {
    double ux = p2[0] - p1[0], uy = p2[1] - p1[1], uz = p2[2] - p1[2];
    double vx = p3[0] - p1[0], vy = p3[1] - p1[1], vz = p3[2] - p1[2];
    double wx = p4[0] - p1[0], wy = p4[1] - p1[1], wz = p4[2] - p1[2];
    double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
    return fabs(cx * wx + cy * wy + cz * wz) / 6.0;
}

static double tri_area(const double* p1, const double* p2, const double* p3) {
    double ux = p2[0] - p1[0], uy = p2[1] - p1[1], uz = p2[2] - p1[2];
    double vx = p3[0] - p1[0], vy = p3[1] - p1[1], vz = p3[2] - p1[2];
    double wx = uy * vz - uz * vy, wy = uz * vx - ux * vz, wz = ux * vy - uy * vx;
    return 0.5 * sqrt(wx * wx + wy * wy + wz * wz);
}

/* volume() and compute_mg() (convex_cell.cpp:1215-1307): fan of tets from
   the corner of the first contributing plane over every other face fan. */
static void cell_volume_mg(const cell* C, double* m, double* mg) {
    *m = 0.0;
    mg[0] = mg[1] = mg[2] = 0.0;
    int t_origin = -1;
    for (int v = 0; v < C->nv; ++v) {
        if (C->v2t[v] < 0) continue;
        if (t_origin < 0) {
            t_origin = C->v2t[v];
            continue;
        }
        int t0 = C->v2t[v];
        int t1 = walk_next(C, t0, v);
        int t = walk_next(C, t1, v);
        int guard = 0;
        while (t != t0 && guard++ < 100000) {
            const double *p = C->pt[t_origin], *q = C->pt[t0], *r = C->pt[t1],
                         *s = C->pt[t];
            double cm = tet_vol(p, q, r, s);
            *m += cm;
            mg[0] += cm * (p[0] + q[0] + r[0] + s[0]) / 4.0;
            mg[1] += cm * (p[1] + q[1] + r[1] + s[1]) / 4.0;
            mg[2] += cm * (p[2] + q[2] + r[2] + s[2]) / 4.0;
            t1 = t;
            t = walk_next(C, t, v);
        }
    }
}

/* Integral of |x - p0|^2 over the cell, as ComputeCVTFuncGradVolumetric
   (RVD.cpp:792-866) integrates it: tetrahedra (p0, p1, p2, p3) with apex at
   the seed over the triangles of each facet's fan, each contributing
   (m/10)(U.U + V.V + W.W + U.V + V.W + W.U), U, V, W = p1, p2, p3 - p0. */
static double cell_second_moment(const cell* C, const double* p0) {
    double f = 0.0;
    for (int v = 0; v < C->nv; ++v) {
        if (C->v2t[v] < 0) continue;
        int t0 = C->v2t[v];
        int t1 = walk_next(C, t0, v);
        int t = walk_next(C, t1, v);
        int guard = 0;
        while (t != t0 && guard++ < 100000) {
            const double *p1 = C->pt[t0], *p2 = C->pt[t1], *p3 = C->pt[t];
            double mi = tet_vol(p0, p1, p2, p3);
            double fi = 0.0;
            for (int c = 0; c < 3; ++c) {
                double Uc = p1[c] - p0[c], Vc = p2[c] - p0[c], Wc = p3[c] - p0[c];
                fi += Uc * Uc + Vc * Vc + Wc * Wc;
                fi += Uc * Vc + Vc * Wc + Wc * Uc;
            }
            f += fi * (mi / 10.0);
            t1 = t;
            t = walk_next(C, t, v);
        }
    }
    return f;
}

/* facet_area(v) (convex_cell.cpp:1160-1191) */
static double cell_facet_area(const cell* C, int v) {
    int t0 = C->v2t[v];
    if (t0 < 0) return 0.0;
    double a = 0.0;
    int t1 = walk_next(C, t0, v);
    int t = walk_next(C, t1, v);
    int guard = 0;
    while (t != t0 && guard++ < 100000) {
        a += tri_area(C->pt[t0], C->pt[t1], C->pt[t]);
        t1 = t;
        t = walk_next(C, t, v);
    }
    return a;
}

/* ======================================================================
 * Security-radius loop (generic_RVD.h:2278-2347) with the neighbour
 * retrieval of delaunay_nn.cpp:105-143.
 * ====================================================================== */

typedef struct {
    int cap;
    int* qi;
    double* qd;
    int* nb;
} nbr_scratch;

static int scratch_reserve(nbr_scratch* S, int k) {
    if (k + 1 <= S->cap) return 0;
    int cap = S->cap ? S->cap : 64;
    while (cap < k + 1) cap *= 2;
    S->qi = (int*)realloc(S->qi, sizeof(int) * (size_t)cap);
    S->qd = (double*)realloc(S->qd, sizeof(double) * (size_t)cap);
    S->nb = (int*)realloc(S->nb, sizeof(int) * (size_t)cap);
    if (!S->qi || !S->qd || !S->nb) return -2;
    S->cap = cap;
    return 0;
}

/* get_neighbors_internal(i, k): k+1 nearest including self, self removed,
   a coincident point with a smaller index disconnects i (returns 0), a
   coincident point with a larger index is skipped. */
static int get_neighbors(const kd_tree* T, int i, int k, nbr_scratch* S) {
    int n = T->n;
    int kq = k + 1;
    if (kq > n) kq = n;
    if (scratch_reserve(S, kq) != 0) return -2;
    int got = kd_query(T, T->xyz + 3 * (size_t)i, kq, S->qi, S->qd);
    int c = 0;
    for (int m = 0; m < got; ++m) {
        int j = S->qi[m];
        if (j == i) continue;
        if (S->qd[m] == 0.0) {
            if (j < i) return 0;
            continue;
        }
        S->nb[c++] = j;
        if (c == kq - 1) break;
    }
    return c;
}

static v4 laguerre_plane(const double* pi, double wi, const double* pj,
                         double wj) {
    /* periodic_delaunay_3d.cpp:3719-3726 */
    double pi_len2 = pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2];
    double pj_len2 = pj[0] * pj[0] + pj[1] * pj[1] + pj[2] * pj[2];
    v4 P;
    P.x = 2.0 * (pi[0] - pj[0]);
    P.y = 2.0 * (pi[1] - pj[1]);
    P.z = 2.0 * (pi[2] - pj[2]);
    P.w = ((wi - pi_len2) - (wj - pj_len2));
    return P;
}

/* true when no seed at squared distance >= dij can cut a cell of squared
   radius R2 around pi (c = wi - max_j wj <= 0; c = 0 for Voronoi). */
static int sr_reached(double dij, double R2, double c, int weighted) {
    if (!weighted) return dij > 4.1 * R2;
    double s = dij + c;
    return s > 0.0 && s * s > 4.1 * dij * R2;
}

typedef struct {
    const double* xyz;
    const double* w;
    double wmax;
    int n;
    const double* box;
    int k0;
    int mode;
    const kd_tree* T;
} cell_job;

/* Polyhedron of a finished cell as ConvexCell::save / append_to_mesh
   write it (voronoi/convex_cell.cpp:282-358, 371-444): vertices = alive
   triangles in index order (absolute coordinates), one facet per plane v
   with v2t[v] set and at least 3 corners, walked by triangle_adjacent(t,
   (lv + 1) % 3), tagged with the plane's global index (-1 for the box); and
   the primal tetrahedra (i, v1, v2, v3) of triangles whose three planes are
   bisectors with v1, v2, v3 > i (generic_RVD.h:1035-1055), reoriented so
   that orient3d >= 0 (RVD.cpp:2311-2322).  ib: [nv, nf, nt, faces (nbr,
   size, vertex...), tets (4 each)]; -1 in ib[0] when a stride is too
   small. */
/* orient_3d (numerics/predicates.cpp:1889-1899): sign of
   det(p1 - p0, p2 - p0, p3 - p0) = -det4 of the rows (p, 1), exact */
static int orient3d_sign_exact(const double* p0, const double* p1, const double* p2,
                               const double* p3) {
    double r0[4] = {p0[0], p0[1], p0[2], 1.0}, r1[4] = {p1[0], p1[1], p1[2], 1.0};
    double r2[4] = {p2[0], p2[1], p2[2], 1.0}, r3[4] = {p3[0], p3[1], p3[2], 1.0};
    return -orc_det4_sign(r0, r1, r2, r3);
}

static void cell_mesh_out(const cell* C, int i, const double* shift, const double* xyz,
                          int* ib, int istride, double* vb, int vstride) {
    int* rank = (int*)malloc(sizeof(int) * (size_t)(C->nt > 0 ? C->nt : 1));
    int nv = 0;
    for (int t = 0; t < C->nt; ++t) {
        rank[t] = -1;
        if (!C->alive[t]) continue;
        if (3 * (nv + 1) > vstride) {
            ib[0] = -1;
            free(rank);
            return;
        }
        for (int c = 0; c < 3; ++c) vb[3 * nv + c] = C->pt[t][c] + shift[c];
        rank[t] = nv++;
    }
    int pos = 3, nf = 0;
    for (int v = 1; v < C->nv; ++v) {
        int t0 = C->v2t[v];
        if (t0 < 0) continue;
        int sz = 0, t = t0, guard = 0;
        do {
            ++sz;
            t = walk_next(C, t, v);
        } while (t != t0 && guard++ < 100000);
        if (sz < 3) continue;
        if (pos + 2 + sz > istride) {
            ib[0] = -1;
            free(rank);
            return;
        }
        ib[pos++] = C->vg[v];
        ib[pos++] = sz;
        t = t0;
        do {
            ib[pos++] = rank[t];
            t = walk_next(C, t, v);
        } while (t != t0);
        ++nf;
    }
    int nt = 0;
    for (int t = 0; t < C->nt; ++t) {
        if (!C->alive[t]) continue;
        int a = C->vg[C->T[t][0]], b = C->vg[C->T[t][1]], c = C->vg[C->T[t][2]];
        if (a < 0 || b < 0 || c < 0 || !(a > i && b > i && c > i)) continue;
        if (pos + 4 > istride) {
            ib[0] = -1;
            free(rank);
            return;
        }
        int v0 = i, v1 = a;
        if (orient3d_sign_exact(xyz + 3 * (size_t)i, xyz + 3 * (size_t)a, xyz + 3 * (size_t)b,
                                xyz + 3 * (size_t)c) < 0) {
            v0 = a;
            v1 = i;
        }
        ib[pos++] = v0;
        ib[pos++] = v1;
        ib[pos++] = b;
        ib[pos++] = c;
        ++nt;
    }
    ib[0] = nv;
    ib[1] = nf;
    ib[2] = nt;
    free(rank);
}

static int compute_one(const cell_job* J, int i, cell* C, nbr_scratch* S,
                       double* vol, double* mg, int* cnt, int* nbr,
                       double* farea, int stride, orc_stats* st, double* e2,
                       int* mib, double* mvb) {
    const double* pi_abs = J->xyz + 3 * (size_t)i;
    /* seed-centred frame (optional): the same planes and box translated by
       -pi before they are formed, so the offsets carry no |pi|^2 rounding */
    const double zero3[3] = {0.0, 0.0, 0.0};
    const double* pi = g_local ? zero3 : pi_abs;
    double box_l[6];
    const double* box = J->box;
    if (g_local) {
        for (int c = 0; c < 3; ++c) {
            box_l[c] = J->box[c] - pi_abs[c];
            box_l[3 + c] = J->box[3 + c] - pi_abs[c];
        }
        box = box_l;
    }
    double wi = J->w ? J->w[i] : 0.0;
    double c = J->w ? (wi - J->wmax) : 0.0;
    int weighted = J->w != NULL;
    cell_init_box(C, box);
    int n = J->n;
    int k = J->k0;
    if (J->mode == 2) k = n - 1;
    if (k > n - 1) k = n - 1;
    int jj = 0, prev = 0, done = 0;
    while (!done) {
        int nb = get_neighbors(J->T, i, k, S);
        if (nb < 0) return -2;
        if (nb == 0) {
            st->nb_disconnected++;
            break;
        }
        if (prev == nb) break;
        for (; jj < nb; ++jj) {
            int j = S->nb[jj];
            const double* pj_abs = J->xyz + 3 * (size_t)j;
            double pj_l[3];
            const double* pj = pj_abs;
            if (g_local) {
                for (int c = 0; c < 3; ++c) pj_l[c] = pj_abs[c] - pi_abs[c];
                pj = pj_l;
            }
            double R2 = cell_sq_radius(C, pi);
            double dij = dist2(pi, pj);
            if (J->mode != 2 && sr_reached(dij, R2, c, weighted)) {
                done = 1;
                break;
            }
            v4 P = laguerre_plane(pi, wi, pj, J->w ? J->w[j] : 0.0);
            int e = cell_clip(C, P, j);
            st->nb_clip++;
            if (e == -2) return -2;
            if (C->empty) {
                done = 1;
                break;
            }
        }
        if (done || J->mode != 1 || nb >= n - 1) break;
        prev = nb;
        int kk = nb;
        if (kk > 8)
            kk += kk / 8;
        else
            kk++;
        if (kk > n - 1) kk = n - 1;
        k = kk;
        st->nb_knn_grow++;
    }
    *cnt = 0;
    if (mib) mib[0] = mib[1] = mib[2] = 0;
    if (C->empty) {
        st->nb_empty++;
        *vol = 0.0;
        mg[0] = mg[1] = mg[2] = 0.0;
        if (e2) *e2 = 0.0;
        return 0;
    }
    cell_geometry(C);
    if (mib) {
        const double zero3[3] = {0.0, 0.0, 0.0};
        cell_mesh_out(C, i, g_local ? pi_abs : zero3, J->xyz, mib, g_mi_stride, mvb, g_mv_stride);
    }
    cell_volume_mg(C, vol, mg);
    if (e2) *e2 = cell_second_moment(C, pi);
    if (g_local)
        for (int c = 0; c < 3; ++c) mg[c] += *vol * pi_abs[c];
    for (int v = 1; v < C->nv; ++v) {
        if (C->v2t[v] < 0 || C->vg[v] < 0) continue;
        if (*cnt >= stride) return -1;
        nbr[*cnt] = C->vg[v];
        farea[*cnt] = cell_facet_area(C, v);
        (*cnt)++;
    }
    if (C->bad) st->nb_bad_border++;
    return 0;
}

static void stats_add(orc_stats* a, const orc_stats* b) {
    a->nb_clip += b->nb_clip;
    a->nb_knn_grow += b->nb_knn_grow;
    a->nb_disconnected += b->nb_disconnected;
    a->nb_empty += b->nb_empty;
    a->nb_bad_border += b->nb_bad_border;
}

int orc_cells_subset(const double* xyz, const double* w, int n,
                     const double* box, int k0, int mode, const int* cells,
                     int nc, double* vol, double* mg, int* nbr_cnt, int* nbr,
                     double* farea, int stride, orc_stats* st) {
    if (n < 1) return 0;
    kd_tree T;
    if (kd_build(&T, xyz, n) != 0) {
        kd_free(&T);
        return -2;
    }
    cell_job J = {xyz, w, 0.0, n, box, k0 < 1 ? 1 : k0, mode, &T};
    if (w) {
        double m = -DBL_MAX;
        for (int i = 0; i < n; ++i)
            if (w[i] > m) m = w[i];
        J.wmax = m;
    }
    int err = 0;
    orc_stats total;
    memset(&total, 0, sizeof(total));
#pragma omp parallel num_threads(orc_get_threads())
    {
        cell C;
        memset(&C, 0, sizeof(C));
        nbr_scratch S;
        memset(&S, 0, sizeof(S));
        orc_stats loc;
        memset(&loc, 0, sizeof(loc));
        int lerr = 0;
#pragma omp for schedule(dynamic, 64)
        for (int c = 0; c < nc; ++c) {
            if (lerr) continue;
            int i = cells ? cells[c] : c;
            int e = compute_one(&J, i, &C, &S, vol + c, mg + 3 * (size_t)c,
                                nbr_cnt + c, nbr + (size_t)c * stride,
                                farea + (size_t)c * stride, stride, &loc,
                                g_e2 ? g_e2 + c : NULL,
                                g_mi ? g_mi + (size_t)c * g_mi_stride : NULL,
                                g_mv ? g_mv + (size_t)c * g_mv_stride : NULL);
            if (e) lerr = e;
        }
#pragma omp critical
        {
            stats_add(&total, &loc);
            if (lerr && !err) err = lerr;
        }
        cell_free(&C);
        free(S.qi);
        free(S.qd);
        free(S.nb);
    }
    kd_free(&T);
    if (st) *st = total;
    return err;
}

int orc_cells(const double* xyz, const double* w, int n, const double* box,
              int k0, int mode, double* vol, double* mg, int* nbr_cnt,
              int* nbr, double* farea, int stride, orc_stats* st) {
    return orc_cells_subset(xyz, w, n, box, k0, mode, NULL, n, vol, mg,
                            nbr_cnt, nbr, farea, stride, st);
}

int orc_lloyd(double* xyz, int n, const double* box, int iters, int k0,
              orc_stats* st) {
    const int stride = 128;
    double* vol = (double*)malloc(sizeof(double) * (size_t)n);
    double* mg = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    int* cnt = (int*)malloc(sizeof(int) * (size_t)n);
    int* nbr = (int*)malloc(sizeof(int) * (size_t)n * stride);
    double* fa = (double*)malloc(sizeof(double) * (size_t)n * stride);
    int err = 0;
    if (!vol || !mg || !cnt || !nbr || !fa) err = -2;
    orc_stats total;
    memset(&total, 0, sizeof(total));
    for (int it = 0; it < iters && !err; ++it) {
        orc_stats s;
        err = orc_cells(xyz, NULL, n, box, k0, 1, vol, mg, cnt, nbr, fa, stride,
                        &s);
        stats_add(&total, &s);
        if (err) break;
        /* CVT.cpp:154-161 */
        for (int j = 0; j < n; ++j) {
            if (vol[j] > 1e-30) {
                double s1 = 1.0 / vol[j];
                xyz[3 * (size_t)j + 0] = s1 * mg[3 * (size_t)j + 0];
                xyz[3 * (size_t)j + 1] = s1 * mg[3 * (size_t)j + 1];
                xyz[3 * (size_t)j + 2] = s1 * mg[3 * (size_t)j + 2];
            }
        }
    }
    if (st) *st = total;
    free(vol);
    free(mg);
    free(cnt);
    free(nbr);
    free(fa);
    return err;
}

/* ======================================================================
 * Voronoi cells restricted to a triangulated surface: masses and first
 * moments (RestrictedVoronoiDiagram::compute_centroids_on_surface,
 * voronoi/RVD.cpp:371-415), by the traversal of
 * compute_surfacic_with_seeds_priority (voronoi/generic_RVD.h:1319-1420):
 * facets in index order, each started from the seed nearest its first
 * vertex (find_seed_near_facet, generic_RVD.h:2010-2068) or from the seed
 * that reached it through a shared edge, and the Delaunay 1-skeleton walked
 * through the bisector edges of each restricted polygon.  A polygon is the
 * facet clipped by Sutherland-Hodgman against the bisectors of the seed's
 * neighbours in distance order until the security radius
 * (clip_by_cell_SR, generic_RVD.h:2133-2190; Polygon::clip_by_plane_fast,
 * voronoi/generic_RVD_polygon.h:242-366), and is integrated as a fan from
 * its first vertex (TriangleAction, generic_RVD.h:425-460) with Heron's
 * area (basic/geometry_nd.h:143-156) and, with vertex weights, the linear
 * density of Geom::triangle_centroid (basic/geometry_nd.h:178-197).
 * ====================================================================== */

typedef struct {
    double p[3];
    double w;
    int adj_f; /* facet across the edge leaving this vertex (-1: none) */
    int adj_s; /* seed across the edge arriving at this vertex (-1: none) */
} spoly_v;

typedef struct {
    int n, cap;
    spoly_v* v;
} spoly;

static int spoly_push(spoly* P, spoly_v x) {
    if (P->n == P->cap) {
        int cap = P->cap ? 2 * P->cap : 16;
        spoly_v* nv = (spoly_v*)realloc(P->v, sizeof(spoly_v) * (size_t)cap);
        if (!nv) return -2;
        P->v = nv;
        P->cap = cap;
    }
    P->v[P->n++] = x;
    return 0;
}

static int sgn(double x) { return (x > 0.0) - (x < 0.0); }

/* dst = src on pi's side of the bisector of [pi, pj] (seed j) */
static int spoly_clip(const spoly* src, spoly* dst, const double* pi, const double* pj, int j) {
    dst->n = 0;
    if (src->n == 0) return 0;
    double d = 0.0;
    for (int c = 0; c < 3; ++c) d += (pi[c] + pj[c]) * (pi[c] - pj[c]);
    const spoly_v* pv = &src->v[src->n - 1];
    double pl = 0.0;
    for (int c = 0; c < 3; ++c) pl += pv->p[c] * (pi[c] - pj[c]);
    int ps = sgn(2.0 * pl - d);
    for (int k = 0; k < src->n; ++k) {
        const spoly_v* vk = &src->v[k];
        double l = 0.0;
        for (int c = 0; c < 3; ++c) l += vk->p[c] * (pi[c] - pj[c]);
        const int s = sgn(2.0 * l - d);
        if (s != ps && ps != 0) {
            const double denom = 2.0 * (pl - l);
            double l1 = 0.5, l2 = 0.5;
            if (fabs(denom) >= 1e-20) {
                l1 = (d - 2.0 * l) / denom;
                l2 = 1.0 - l1;
            }
            spoly_v I;
            for (int c = 0; c < 3; ++c) I.p[c] = l1 * pv->p[c] + l2 * vk->p[c];
            I.w = l1 * pv->w + l2 * vk->w;
            if (s > 0) {
                I.adj_f = pv->adj_f;
                I.adj_s = j;
            } else {
                I.adj_f = -1;
                I.adj_s = vk->adj_s;
            }
            if (spoly_push(dst, I)) return -2;
        }
        if (s > 0 && spoly_push(dst, *vk)) return -2;
        pv = vk;
        pl = l;
        ps = s;
    }
    return 0;
}

static double heron(const double* a, const double* b, const double* c) {
    const double la = sqrt(dist2(a, b)), lb = sqrt(dist2(b, c)), lc = sqrt(dist2(c, a));
    const double s = 0.5 * (la + lb + lc);
    const double A2 = s * (s - la) * (s - lb) * (s - lc);
    return sqrt(A2 > 0.0 ? A2 : 0.0);
}

static void spoly_integrate(const spoly* P, int weighted, double* m, double* mg) {
    for (int i = 1; i + 1 < P->n; ++i) {
        const spoly_v *a = &P->v[0], *b = &P->v[i], *c = &P->v[i + 1];
        const double area = heron(a->p, b->p, c->p);
        if (!weighted) {
            const double s = area / 3.0;
            *m += area;
            for (int k = 0; k < 3; ++k) mg[k] += s * (a->p[k] + b->p[k] + c->p[k]);
        } else {
            const double abc = a->w + b->w + c->w;
            const double wa = a->w + abc, wb = b->w + abc, wc = c->w + abc, s = area / 12.0;
            *m += area / 3.0 * abc;
            for (int k = 0; k < 3; ++k) mg[k] += s * (wa * a->p[k] + wb * b->p[k] + wc * c->p[k]);
        }
    }
}

static int facet_polygon(const double* V, const int* F, const int* adj, const double* vw, int f,
                         spoly* P) {
    P->n = 0;
    for (int k = 0; k < 3; ++k) {
        const int v = F[3 * f + k];
        spoly_v x;
        for (int c = 0; c < 3; ++c) x.p[c] = V[3 * (size_t)v + c];
        x.w = vw ? vw[v] : 1.0;
        x.adj_f = adj[3 * f + k];
        x.adj_s = -1;
        if (spoly_push(P, x)) return -2;
    }
    return 0;
}

/* facet across edge (k, k+1) of each facet, -1 on the border: the
   facet_corners adjacency of a manifold triangle mesh */
static int facet_adjacency(const int* F, int nf, int* adj) {
    const int ne = 3 * nf;
    long long* key = (long long*)malloc(sizeof(long long) * (size_t)ne);
    int* ord = (int*)malloc(sizeof(int) * (size_t)ne);
    if (!key || !ord) {
        free(key);
        free(ord);
        return -2;
    }
    for (int e = 0; e < ne; ++e) {
        const int f = e / 3, k = e % 3;
        const long long a = F[3 * f + k], b = F[3 * f + (k + 1) % 3];
        key[e] = a < b ? (a << 32) | b : (b << 32) | a;
        ord[e] = e;
        adj[e] = -1;
    }
    /* insertion-free: shell sort of edge ids by key */
    for (int gap = ne / 2; gap > 0; gap /= 2)
        for (int i = gap; i < ne; ++i) {
            const int t = ord[i];
            int j = i;
            while (j >= gap && key[ord[j - gap]] > key[t]) {
                ord[j] = ord[j - gap];
                j -= gap;
            }
            ord[j] = t;
        }
    for (int i = 0; i + 1 < ne; ++i)
        if (key[ord[i]] == key[ord[i + 1]] && (i + 2 >= ne || key[ord[i + 2]] != key[ord[i]])) {
            adj[ord[i]] = ord[i + 1] / 3;
            adj[ord[i + 1]] = ord[i] / 3;
        }
    free(key);
    free(ord);
    return 0;
}

/* polygon of seed s on facet F: the SR loop of clip_by_cell_SR */
static int surface_clip_seed(const kd_tree* T, const double* xyz, int n, int s, const spoly* F,
                             spoly* ping, spoly* pong, nbr_scratch* S, int k0, long long* nclip) {
    ping->n = 0;
    for (int k = 0; k < F->n; ++k)
        if (spoly_push(ping, F->v[k])) return -2;
    const double* pi = xyz + 3 * (size_t)s;
    int k = k0 < n - 1 ? k0 : n - 1;
    int jj = 0, prev = -1;
    while (1) {
        const int got = get_neighbors(T, s, k, S);
        if (got < 0) return got;
        if (got == 0 || got == prev) return 0;
        for (; jj < got; ++jj) {
            const int j = S->nb[jj];
            double R2 = 0.0;
            for (int v = 0; v < ping->n; ++v) {
                const double dv = dist2(pi, ping->v[v].p);
                if (dv > R2) R2 = dv;
            }
            if (dist2(pi, xyz + 3 * (size_t)j) > 4.1 * R2) return 0;
            if (spoly_clip(ping, pong, pi, xyz + 3 * (size_t)j, j)) return -2;
            ++*nclip;
            spoly tmp = *ping;
            *ping = *pong;
            *pong = tmp;
        }
        prev = got;
        if (got >= n - 1) return 0;
        k = got > 8 ? got + got / 8 : got + 1;
        if (k > n - 1) k = n - 1;
    }
}

/* mode 0: the reference traversal above; mode 1: every (facet, seed) pair,
   the facet clipped by all n - 1 bisectors of the seed (brute force).
   mg[3n] / m[n] are overwritten.  Returns 0 or <0. */
int orc_surface_centroids(const double* xyz, int n, const double* V, int nv, const int* F, int nf,
                          const double* vw, int mode, double* mg, double* m, long long* npoly) {
    (void)nv;
    for (int i = 0; i < n; ++i) {
        m[i] = 0.0;
        mg[3 * i] = mg[3 * i + 1] = mg[3 * i + 2] = 0.0;
    }
    if (npoly) *npoly = 0;
    if (n == 0 || nf == 0) return 0;
    int err = 0;
    int* adj = (int*)malloc(sizeof(int) * 3 * (size_t)nf);
    if (!adj) return -2;
    err = facet_adjacency(F, nf, adj);
    spoly Fp = {0, 0, NULL}, ping = {0, 0, NULL}, pong = {0, 0, NULL};
    long long np = 0, nclip = 0;
    if (!err && mode == 1) {
        for (int f = 0; f < nf && !err; ++f) {
            err = facet_polygon(V, F, adj, vw, f, &Fp);
            for (int s = 0; s < n && !err; ++s) {
                ping.n = 0;
                for (int k = 0; k < Fp.n; ++k) spoly_push(&ping, Fp.v[k]);
                for (int j = 0; j < n && ping.n > 0 && !err; ++j) {
                    if (j == s) continue;
                    err = spoly_clip(&ping, &pong, xyz + 3 * (size_t)s, xyz + 3 * (size_t)j, j);
                    spoly tmp = ping;
                    ping = pong;
                    pong = tmp;
                }
                if (ping.n >= 3) {
                    spoly_integrate(&ping, vw != NULL, m + s, mg + 3 * (size_t)s);
                    ++np;
                }
            }
        }
    } else if (!err) {
        kd_tree T;
        nbr_scratch S = {0, NULL, NULL, NULL};
        char* fmark = (char*)calloc((size_t)nf, 1);
        int* stamp = (int*)malloc(sizeof(int) * (size_t)n);
        int* fstack_f = (int*)malloc(sizeof(int) * (size_t)nf);
        int* fstack_s = (int*)malloc(sizeof(int) * (size_t)nf);
        int* sstack = (int*)malloc(sizeof(int) * (size_t)n);
        if (kd_build(&T, xyz, n) != 0 || !fmark || !stamp || !fstack_f || !fstack_s || !sstack)
            err = -2;
        for (int i = 0; i < n && !err; ++i) stamp[i] = -1;
        for (int f0 = 0; f0 < nf && !err; ++f0) {
            if (fmark[f0]) continue;
            fmark[f0] = 1;
            int nfs = 0;
            {
                int near;
                double nd;
                kd_query(&T, V + 3 * (size_t)F[3 * f0], 1, &near, &nd);
                fstack_f[nfs] = f0;
                fstack_s[nfs++] = near;
            }
            while (nfs > 0 && !err) {
                --nfs;
                const int cf = fstack_f[nfs];
                int cs = fstack_s[nfs];
                err = facet_polygon(V, F, adj, vw, cf, &Fp);
                stamp[cs] = cf;
                int nss = 0;
                sstack[nss++] = cs;
                while (nss > 0 && !err) {
                    cs = sstack[--nss];
                    err = surface_clip_seed(&T, xyz, n, cs, &Fp, &ping, &pong, &S, 20, &nclip);
                    if (err) break;
                    if (ping.n >= 3) ++np;
                    spoly_integrate(&ping, vw != NULL, m + cs, mg + 3 * (size_t)cs);
                    for (int v = 0; v < ping.n; ++v) {
                        const int nf2 = ping.v[v].adj_f;
                        if (nf2 >= 0 && nf2 != cf && !fmark[nf2]) {
                            fmark[nf2] = 1;
                            fstack_f[nfs] = nf2;
                            fstack_s[nfs++] = cs;
                        }
                        const int ns = ping.v[v].adj_s;
                        if (ns >= 0 && stamp[ns] != cf) {
                            stamp[ns] = cf;
                            sstack[nss++] = ns;
                        }
                    }
                }
            }
        }
        kd_free(&T);
        free(S.qi);
        free(S.qd);
        free(S.nb);
        free(fmark);
        free(stamp);
        free(fstack_f);
        free(fstack_s);
        free(sstack);
    }
    free(Fp.v);
    free(ping.v);
    free(pong.v);
    free(adj);
    if (npoly) *npoly = np;
    return err;
}

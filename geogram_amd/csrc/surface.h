// Voronoi cells restricted to a triangulated surface: per seed, the area and
// first moment of the part of the surface in its cell (the masses and
// centroids of RestrictedVoronoiDiagram::compute_centroids_on_surface,
// voronoi/RVD.cpp:371-415).
//
// One lane per surface triangle ("item").  The lane walks the seeds whose
// cells meet its triangle: it starts from the seed nearest the triangle's
// first vertex (that vertex lies in its cell; find_seed_near_facet,
// voronoi/generic_RVD.h:2010-2068), clips the triangle by the bisectors of
// the seed's neighbours in distance order until the security radius
// (clip_by_cell_SR, generic_RVD.h:2133-2190; Sutherland-Hodgman as
// Polygon::clip_by_plane_fast, voronoi/generic_RVD_polygon.h:242-366), and
// queues the seeds across the bisector edges of the polygon (the walk on
// the Delaunay 1-skeleton, generic_RVD.h:1372-1420).  The restricted cells
// on a triangle form a connected partition of it, so the walk reaches every
// seed whose cell meets it.
//
// Each non-degenerate polygon becomes a record (seed, item) -> (m, mg); a
// final pass sorts the records by key and sums each seed's records in that
// order, so results do not depend on the order lanes write them.
//
// Polygons are kept in the frame of the seed being clipped (a bisector is
// e.y <= |e|^2 / 2 with e = pj - pi), as the cells are.  An item whose
// polygon or seed list outgrows the lane's arrays is split into four (the
// integrals add up, the linear density is interpolated exactly); an item
// whose neighbour lists end before the security radius is redone with
// longer lists.
#pragma once

#include "common.h"

namespace vgc {

constexpr int kSurfPV = 24;  // polygon vertices
constexpr int kSurfVC = 48;  // seeds per item
enum SurfFlag : int { SURF_NEED_K = 1, SURF_OVERFLOW = 2 };

struct SurfParams {
    const double* xyz;     // seeds, caller order
    int n;
    const int* knn;        // K nearest other seeds per seed, -1 after the last
    const double* knn_d2;  // their squared distances
    int K;
    int k_all;             // the lists hold every other seed
    const double* tv;      // items: 9 coordinates
    const double* tw;      // items: 3 vertex densities (nullptr: 1)
    const int* seed0;      // seed nearest each item's first vertex
    int n_items;
    uint32_t item_base;    // global id of item 0
    unsigned long long* rec_key;
    double* rec_val;       // 4 per record: m, mg
    unsigned long long* rec_count;
    unsigned long long rec_cap;
    int* flags;            // per item (local index)
};

struct SurfV {
    double x, y, z, w;
    int s;  // seed across the edge arriving at this vertex, -1: a triangle edge
};

VGC_SD unsigned long long surf_reserve(unsigned long long* c) {
#if defined(VGC_EMU)
    return __atomic_fetch_add(c, 1ull, __ATOMIC_RELAXED);
#else
    return atomicAdd(c, 1ull);
#endif
}

VGC_SD int surf_sgn(double x) { return (x > 0.0) - (x < 0.0); }

// dst = the part of src on the seed's side of the plane e.y = h (seed j on
// the other side).  false when dst would exceed kSurfPV vertices.
VGC_SD bool surf_clip(const SurfV* src, int ns, SurfV* dst, int& nd, double ex, double ey,
                      double ez, double h, int j) {
    nd = 0;
    if (ns == 0) return true;
    SurfV pv = src[ns - 1];
    double pu = h - (ex * pv.x + ey * pv.y + ez * pv.z);
    int ps = surf_sgn(pu);
    for (int k = 0; k < ns; ++k) {
        const SurfV v = src[k];
        const double u = h - (ex * v.x + ey * v.y + ez * v.z);
        const int s = surf_sgn(u);
        if (s != ps && ps != 0) {
            // the crossing point l1 prev + l2 v of the edge (prev, v)
            const double den = 2.0 * (pu - u);
            double l1 = 0.5, l2 = 0.5;
            if (fabs(den) >= 1e-20) {
                l1 = (-2.0 * u) / den;
                l2 = 1.0 - l1;
            }
            if (nd == kSurfPV) return false;
            SurfV I;
            I.x = l1 * pv.x + l2 * v.x;
            I.y = l1 * pv.y + l2 * v.y;
            I.z = l1 * pv.z + l2 * v.z;
            I.w = l1 * pv.w + l2 * v.w;
            I.s = s > 0 ? j : v.s;
            dst[nd++] = I;
        }
        if (s > 0) {
            if (nd == kSurfPV) return false;
            dst[nd++] = v;
        }
        pv = v;
        pu = u;
        ps = s;
    }
    return true;
}

// area and first moment (in the polygon's frame) of a fan from vertex 0;
// with densities, the integral of the linear density and of x times it
// (Geom::triangle_centroid, basic/geometry_nd.h:178-197)
VGC_SD void surf_integrate(const SurfV* A, int na, bool weighted, double& m, double& gx,
                           double& gy, double& gz) {
    m = gx = gy = gz = 0.0;
    const SurfV a = A[0];
    for (int i = 1; i + 1 < na; ++i) {
        const SurfV b = A[i], c = A[i + 1];
        const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
        const double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
        const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
        const double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        if (!weighted) {
            const double s = area / 3.0;
            m += area;
            gx += s * (a.x + b.x + c.x);
            gy += s * (a.y + b.y + c.y);
            gz += s * (a.z + b.z + c.z);
        } else {
            const double abc = a.w + b.w + c.w;
            const double wa = a.w + abc, wb = b.w + abc, wc = c.w + abc, s = area / 12.0;
            m += area / 3.0 * abc;
            gx += s * (wa * a.x + wb * b.x + wc * c.x);
            gy += s * (wa * a.y + wb * b.y + wc * c.y);
            gz += s * (wa * a.z + wb * b.z + wc * c.z);
        }
    }
}

VGC_SD void surface_item(const SurfParams& P, int t) {
    const double* tv = P.tv + 9 * (size_t)t;
    int vis[kSurfVC];
    SurfV buf[2][kSurfPV];
    int nv = 0, head = 0, flag = 0;
    vis[nv++] = P.seed0[t];
    while (head < nv && !flag) {
        const int s = vis[head++];
        const double px = P.xyz[3 * (size_t)s], py = P.xyz[3 * (size_t)s + 1],
                     pz = P.xyz[3 * (size_t)s + 2];
        int cur = 0, na = 3;
        for (int k = 0; k < 3; ++k) {
            SurfV v;
            v.x = tv[3 * k] - px;
            v.y = tv[3 * k + 1] - py;
            v.z = tv[3 * k + 2] - pz;
            v.w = P.tw ? P.tw[3 * (size_t)t + k] : 1.0;
            v.s = -1;
            buf[0][k] = v;
        }
        const int* L = P.knn + (size_t)s * P.K;
        const double* Ld = P.knn_d2 + (size_t)s * P.K;
        bool done = false;
        int m = 0;
        for (; m < P.K; ++m) {
            const int j = L[m];
            if (j < 0) break;
            double R2 = 0.0;
            for (int v = 0; v < na; ++v) {
                const SurfV q = buf[cur][v];
                R2 = fmax(R2, q.x * q.x + q.y * q.y + q.z * q.z);
            }
            if (Ld[m] > 4.1 * R2) {
                done = true;
                break;
            }
            const double ex = P.xyz[3 * (size_t)j] - px, ey = P.xyz[3 * (size_t)j + 1] - py,
                         ez = P.xyz[3 * (size_t)j + 2] - pz;
            const double h = 0.5 * (ex * ex + ey * ey + ez * ez);
            int nb = 0;
            if (!surf_clip(buf[cur], na, buf[1 - cur], nb, ex, ey, ez, h, j)) {
                flag = SURF_OVERFLOW;
                break;
            }
            cur = 1 - cur;
            na = nb;
        }
        if (flag) break;
        // an empty row (no other seed, or the coincident-seed rule) leaves
        // the triangle unclipped; a row that ends early is complete only
        // when it lists every other seed
        if (!done && m > 0 && !P.k_all) {
            flag = SURF_NEED_K;
            break;
        }
        if (na >= 3) {
            double mm, gx, gy, gz;
            surf_integrate(buf[cur], na, P.tw != nullptr, mm, gx, gy, gz);
            const unsigned long long r = surf_reserve(P.rec_count);
            if (r < P.rec_cap) {
                P.rec_key[r] = ((unsigned long long)(unsigned)s << 32) |
                               (unsigned long long)(P.item_base + (uint32_t)t);
                P.rec_val[4 * r] = mm;
                P.rec_val[4 * r + 1] = gx + mm * px;
                P.rec_val[4 * r + 2] = gy + mm * py;
                P.rec_val[4 * r + 3] = gz + mm * pz;
            }
        }
        for (int v = 0; v < na && !flag; ++v) {
            const int j = buf[cur][v].s;
            if (j < 0) continue;
            bool seen = false;
            for (int q = 0; q < nv; ++q) seen = seen || vis[q] == j;
            if (seen) continue;
            if (nv == kSurfVC) {
                flag = SURF_OVERFLOW;
                break;
            }
            vis[nv++] = j;
        }
    }
    P.flags[t] = flag;
}

// the four halves-of-edges children of item t (corner triangles first, the
// middle one last) into slots 4c .. 4c + 3 of the output
VGC_SD void surface_split_row(const double* tv, const double* tw, int t, int c, double* ov,
                              double* ow) {
    double p[6][3], w[6];
    for (int k = 0; k < 3; ++k) {
        for (int a = 0; a < 3; ++a) p[k][a] = tv[9 * (size_t)t + 3 * k + a];
        w[k] = tw ? tw[3 * (size_t)t + k] : 1.0;
    }
    for (int k = 0; k < 3; ++k) {
        const int k1 = (k + 1) % 3;
        for (int a = 0; a < 3; ++a) p[3 + k][a] = 0.5 * (p[k][a] + p[k1][a]);
        w[3 + k] = 0.5 * (w[k] + w[k1]);
    }
    // vertices: 0 1 2, midpoints 3 = (0,1), 4 = (1,2), 5 = (2,0)
    const int child[4][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}, {3, 4, 5}};
    for (int q = 0; q < 4; ++q)
        for (int k = 0; k < 3; ++k) {
            const int v = child[q][k];
            for (int a = 0; a < 3; ++a) ov[9 * (size_t)(4 * c + q) + 3 * k + a] = p[v][a];
            if (ow) ow[3 * (size_t)(4 * c + q) + k] = w[v];
        }
}

// the triangles of the mesh as items
VGC_SD void surface_items_row(const double* V, const int* F, const double* vw, int t, double* ov,
                              double* ow) {
    for (int k = 0; k < 3; ++k) {
        const int v = F[3 * (size_t)t + k];
        for (int a = 0; a < 3; ++a) ov[9 * (size_t)t + 3 * k + a] = V[3 * (size_t)v + a];
        if (ow) ow[3 * (size_t)t + k] = vw[v];
    }
}

// seed s's sum of its records (sorted keys; records of flagged items left
// out), in key order
VGC_SD void surface_reduce_row(int s, const unsigned long long* keys, const int* idx,
                               long long nrec, const double* val, const int* item_flag,
                               double* m, double* mg) {
    const unsigned long long k0 = (unsigned long long)(unsigned)s << 32;
    long long lo = 0, hi = nrec;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (keys[mid] < k0)
            lo = mid + 1;
        else
            hi = mid;
    }
    double a = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (long long r = lo; r < nrec && (keys[r] >> 32) == (unsigned long long)(unsigned)s; ++r) {
        if (item_flag[(uint32_t)keys[r]]) continue;
        const double* v = val + 4 * (size_t)idx[r];
        a += v[0];
        gx += v[1];
        gy += v[2];
        gz += v[3];
    }
    m[s] = a;
    mg[3 * (size_t)s] = gx;
    mg[3 * (size_t)s + 1] = gy;
    mg[3 * (size_t)s + 2] = gz;
}

}  // namespace vgc
